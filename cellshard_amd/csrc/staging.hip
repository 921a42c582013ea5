// Device staging: packed per-partition / per-shard uploads of host matrices,
// device view descriptors, and a per-process allocation cache.
// Reference: include/CellShard/runtime/device/sharded_device.cuh:22-50 (records),
// :199-243 (16-entry best-fit allocation cache), :343-755 (upload/upload_async),
// :958-990 (release), :1336-1432 (packed Blocked-ELL shard upload).
#include "common.hh"

#include <cstdlib>
#include <vector>

namespace csamd {
namespace {

constexpr unsigned long CACHE_LIMIT = 16ul;
constexpr size_t STAGE_ALIGN = 256u;

void *stage_alloc(csamd_allocation_cache *cache, int device, size_t bytes, size_t *out_bytes) {
    size_t got = 0u;
    void *p = cache != nullptr ? csamd_take_cached_allocation(cache, device, bytes, &got) : nullptr;
    if (p != nullptr) {
        *out_bytes = got;
        return p;
    }
    if (!hip_ok(hipMalloc(&p, bytes), "hipMalloc staging")) return nullptr;
    *out_bytes = bytes;
    return p;
}

int copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t stream) {
    if (bytes == 0u) return 1;
    if (stream == nullptr) return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D staging");
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D staging");
}

void zero_record(csamd_partition_record *r) {
    std::memset(r, 0, sizeof(*r));
    r->device_id = -1;
}

struct layout_builder {
    size_t off = 0;
    size_t take(size_t bytes) {
        off = align_up(off, STAGE_ALIGN);
        const size_t at = off;
        off += bytes;
        return at;
    }
};

int current_device() {
    int d = 0;
    (void) hipGetDevice(&d);
    return d;
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_allocation_cache_init(csamd_allocation_cache *c) {
    if (c != nullptr) std::memset(c, 0, sizeof(*c));
}

void csamd_allocation_cache_clear(csamd_allocation_cache *c) {
    if (c == nullptr) return;
    int cur = current_device();
    for (unsigned long i = 0; i < c->cache_count; ++i) {
        if (c->cache[i].ptr == nullptr) continue;
        (void) hipSetDevice(c->cache[i].device_id >= 0 ? c->cache[i].device_id : 0);
        (void) hipFree(c->cache[i].ptr);
    }
    (void) hipSetDevice(cur);
    std::free(c->cache);
    csamd_allocation_cache_init(c);
}

void *csamd_take_cached_allocation(csamd_allocation_cache *c, int device_id, size_t min_bytes, size_t *out_bytes) {
    if (out_bytes != nullptr) *out_bytes = 0u;
    if (c == nullptr || c->cache == nullptr) return nullptr;
    unsigned long best = (unsigned long) -1;
    for (unsigned long i = 0; i < c->cache_count; ++i) {
        if (c->cache[i].ptr == nullptr || c->cache[i].device_id != device_id || c->cache[i].bytes < min_bytes) continue;
        if (best == (unsigned long) -1 || c->cache[i].bytes < c->cache[best].bytes) best = i;
    }
    if (best == (unsigned long) -1) return nullptr;
    void *p = c->cache[best].ptr;
    if (out_bytes != nullptr) *out_bytes = c->cache[best].bytes;
    c->cache[best] = c->cache[c->cache_count - 1ul];
    --c->cache_count;
    return p;
}

int csamd_cache_allocation(csamd_allocation_cache *c, int device_id, void *ptr, size_t bytes) {
    if (ptr == nullptr) return 1;
    if (c == nullptr || bytes == 0u || c->cache_count >= CACHE_LIMIT) return hip_ok(hipFree(ptr), "hipFree cached allocation");
    if (c->cache_count == c->cache_capacity) {
        const unsigned long cap = c->cache_capacity != 0ul ? c->cache_capacity * 2ul : CACHE_LIMIT;
        void *g = std::realloc(c->cache, cap * sizeof(csamd_cached_allocation));
        if (g == nullptr) return hip_ok(hipFree(ptr), "hipFree cached allocation");
        c->cache = (csamd_cached_allocation *) g;
        c->cache_capacity = cap;
    }
    c->cache[c->cache_count].ptr = ptr;
    c->cache[c->cache_count].bytes = bytes;
    c->cache[c->cache_count].device_id = device_id;
    ++c->cache_count;
    return 1;
}

int csamd_release_record(csamd_partition_record *r, csamd_allocation_cache *cache) {
    CSAMD_REQUIRE(r != nullptr, "release_record: null record");
    int ok = 1;
    if (r->allocation != nullptr) ok = csamd_cache_allocation(cache, r->device_id, r->allocation, r->allocation_bytes);
    zero_record(r);
    return ok;
}

int csamd_upload_compressed(const csamd_host_compressed *src, csamd_partition_record *rec, csamd_compressed_view *out_view,
                            csamd_allocation_cache *cache, csamd_stream_t stream) {
    CSAMD_REQUIRE(src != nullptr && rec != nullptr, "upload_compressed: null argument");
    zero_record(rec);
    const unsigned int major = src->axis == CSAMD_COMPRESSED_BY_COL ? src->cols : src->rows;
    layout_builder L;
    const size_t o_view = L.take(sizeof(csamd_compressed_view));
    const size_t o_ptr = L.take((size_t) (major + 1u) * 4u);
    const size_t o_idx = L.take((size_t) src->nnz * 4u);
    const size_t o_val = L.take((size_t) src->nnz * 2u);
    const int dev = current_device();
    size_t got = 0u;
    char *base = (char *) stage_alloc(cache, dev, L.off, &got);
    if (base == nullptr) return 0;
    csamd_compressed_view v;
    std::memset(&v, 0, sizeof(v));
    v.rows = src->rows;
    v.cols = src->cols;
    v.nnz = src->nnz;
    v.axis = src->axis;
    v.majorPtr = (unsigned int *) (base + o_ptr);
    v.minorIdx = src->nnz != 0u ? (unsigned int *) (base + o_idx) : nullptr;
    v.val = src->nnz != 0u ? (csamd_half *) (base + o_val) : nullptr;
    hipStream_t s = (hipStream_t) stream;
    int ok = copy_h2d(v.majorPtr, src->majorPtr, (size_t) (major + 1u) * 4u, s) && copy_h2d(v.minorIdx, src->minorIdx, (size_t) src->nnz * 4u, s) &&
             copy_h2d(v.val, src->val, (size_t) src->nnz * 2u, s) && hip_ok(hipMemcpy(base + o_view, &v, sizeof(v), hipMemcpyHostToDevice), "view H2D");
    if (!ok) {
        (void) hipFree(base);
        return 0;
    }
    rec->allocation = base;
    rec->allocation_bytes = got;
    rec->storage = base;
    rec->view = base + o_view;
    rec->a0 = v.majorPtr;
    rec->a1 = v.minorIdx;
    rec->a2 = v.val;
    rec->device_id = dev;
    if (out_view != nullptr) *out_view = v;
    return 1;
}

int csamd_upload_blocked_ell(const csamd_host_blocked_ell *src, csamd_partition_record *rec, csamd_blocked_ell_view *out_view,
                             csamd_allocation_cache *cache, csamd_stream_t stream) {
    CSAMD_REQUIRE(src != nullptr && rec != nullptr, "upload_blocked_ell: null argument");
    zero_record(rec);
    const unsigned long long rb = src->block_size != 0u ? (src->rows + src->block_size - 1ull) / src->block_size : 0ull;
    const unsigned long long w = src->block_size != 0u ? src->ell_cols / src->block_size : 0ull;
    layout_builder L;
    const size_t o_view = L.take(sizeof(csamd_blocked_ell_view));
    const size_t o_bci = L.take(rb * w * 4u);
    const size_t o_val = L.take((size_t) src->rows * src->ell_cols * 2u);
    const int dev = current_device();
    size_t got = 0u;
    char *base = (char *) stage_alloc(cache, dev, L.off, &got);
    if (base == nullptr) return 0;
    csamd_blocked_ell_view v;
    std::memset(&v, 0, sizeof(v));
    v.rows = src->rows;
    v.cols = src->cols;
    v.nnz = src->nnz;
    v.block_size = src->block_size;
    v.ell_cols = src->ell_cols;
    v.blockColIdx = rb * w != 0ull ? (unsigned int *) (base + o_bci) : nullptr;
    v.val = (size_t) src->rows * src->ell_cols != 0u ? (csamd_half *) (base + o_val) : nullptr;
    hipStream_t s = (hipStream_t) stream;
    int ok = copy_h2d(v.blockColIdx, src->blockColIdx, rb * w * 4u, s) &&
             copy_h2d(v.val, src->val, (size_t) src->rows * src->ell_cols * 2u, s) &&
             hip_ok(hipMemcpy(base + o_view, &v, sizeof(v), hipMemcpyHostToDevice), "view H2D");
    if (!ok) {
        (void) hipFree(base);
        return 0;
    }
    rec->allocation = base;
    rec->allocation_bytes = got;
    rec->storage = base;
    rec->view = base + o_view;
    rec->a0 = v.blockColIdx;
    rec->a1 = v.val;
    rec->device_id = dev;
    if (out_view != nullptr) *out_view = v;
    return 1;
}

int csamd_upload_sliced_ell(const csamd_host_sliced_ell *src, csamd_partition_record *rec, csamd_sliced_ell_view *out_view,
                            csamd_allocation_cache *cache, csamd_stream_t stream) {
    CSAMD_REQUIRE(src != nullptr && rec != nullptr, "upload_sliced_ell: null argument");
    zero_record(rec);
    const unsigned int S = src->slice_count;
    std::vector<unsigned int> slots(S);
    unsigned long long running = 0ull;
    for (unsigned int s = 0; s < S; ++s) {
        CSAMD_REQUIRE(running <= 0xffffffffull, "upload_sliced_ell: slot offsets exceed u32");
        slots[s] = (unsigned int) running;
        running += (unsigned long long) (src->slice_row_offsets[s + 1u] - src->slice_row_offsets[s]) * src->slice_widths[s];
    }
    CSAMD_REQUIRE(running <= 0xffffffffull, "upload_sliced_ell: total slots exceed u32");
    const size_t total = (size_t) running;
    // uniform slice height (formats/sliced_ell.cuh:86-97)
    unsigned int uniform = 0u;
    if (S == 1u) uniform = src->rows;
    else if (S > 1u) {
        const unsigned int step = src->slice_row_offsets[1] - src->slice_row_offsets[0];
        uniform = step;
        for (unsigned int s = 1; s + 1u < S && uniform != 0u; ++s)
            if (src->slice_row_offsets[s + 1u] - src->slice_row_offsets[s] != step) uniform = 0u;
        if (uniform != 0u && src->slice_row_offsets[S] - src->slice_row_offsets[S - 1u] > step) uniform = 0u;
    }
    layout_builder L;
    const size_t o_view = L.take(sizeof(csamd_sliced_ell_view));
    const size_t o_off = L.take((size_t) (S + 1u) * 4u);
    const size_t o_w = L.take((size_t) S * 4u);
    const size_t o_slot = L.take((size_t) S * 4u);
    const size_t o_col = L.take(total * 4u);
    const size_t o_val = L.take(total * 2u);
    const int dev = current_device();
    size_t got = 0u;
    char *base = (char *) stage_alloc(cache, dev, L.off, &got);
    if (base == nullptr) return 0;
    csamd_sliced_ell_view v;
    std::memset(&v, 0, sizeof(v));
    v.rows = src->rows;
    v.cols = src->cols;
    v.nnz = src->nnz;
    v.slice_count = S;
    v.slice_rows = uniform;
    v.slice_row_offsets = (unsigned int *) (base + o_off);
    v.slice_widths = (unsigned int *) (base + o_w);
    v.slice_slot_offsets = (unsigned int *) (base + o_slot);
    v.col_idx = total != 0u ? (unsigned int *) (base + o_col) : nullptr;
    v.val = total != 0u ? (csamd_half *) (base + o_val) : nullptr;
    hipStream_t s = (hipStream_t) stream;
    int ok = copy_h2d(v.slice_row_offsets, src->slice_row_offsets, (size_t) (S + 1u) * 4u, s) &&
             copy_h2d(v.slice_widths, src->slice_widths, (size_t) S * 4u, s) &&
             hip_ok(hipMemcpy(v.slice_slot_offsets, slots.data(), (size_t) S * 4u, hipMemcpyHostToDevice), "slot offsets H2D") &&
             copy_h2d(v.col_idx, src->col_idx, total * 4u, s) && copy_h2d(v.val, src->val, total * 2u, s) &&
             hip_ok(hipMemcpy(base + o_view, &v, sizeof(v), hipMemcpyHostToDevice), "view H2D");
    if (!ok) {
        (void) hipFree(base);
        return 0;
    }
    rec->allocation = base;
    rec->allocation_bytes = got;
    rec->storage = base;
    rec->view = base + o_view;
    rec->a0 = v.slice_row_offsets;
    rec->a1 = v.slice_widths;
    rec->a2 = v.col_idx;
    rec->a3 = v.val;
    rec->device_id = dev;
    if (out_view != nullptr) *out_view = v;
    return 1;
}

int csamd_upload_blocked_ell_shard(const csamd_host_blocked_ell *const *parts, unsigned int count, csamd_partition_record *rec,
                                   csamd_blocked_ell_view *out_views, csamd_allocation_cache *cache, csamd_stream_t stream) {
    CSAMD_REQUIRE(parts != nullptr && rec != nullptr && out_views != nullptr, "upload_blocked_ell_shard: null argument");
    zero_record(rec);
    layout_builder L;
    const size_t o_views = L.take((size_t) count * sizeof(csamd_blocked_ell_view));
    std::vector<size_t> o_bci(count), o_val(count);
    for (unsigned int i = 0; i < count; ++i) {
        const csamd_host_blocked_ell *p = parts[i];
        CSAMD_REQUIRE(p != nullptr, "upload_blocked_ell_shard: null partition %u", i);
        const unsigned long long rb = p->block_size != 0u ? (p->rows + p->block_size - 1ull) / p->block_size : 0ull;
        const unsigned long long w = p->block_size != 0u ? p->ell_cols / p->block_size : 0ull;
        o_bci[i] = L.take(rb * w * 4u);
        o_val[i] = L.take((size_t) p->rows * p->ell_cols * 2u);
    }
    const int dev = current_device();
    size_t got = 0u;
    char *base = (char *) stage_alloc(cache, dev, L.off, &got);
    if (base == nullptr) return 0;
    hipStream_t s = (hipStream_t) stream;
    for (unsigned int i = 0; i < count; ++i) {
        const csamd_host_blocked_ell *p = parts[i];
        const unsigned long long rb = p->block_size != 0u ? (p->rows + p->block_size - 1ull) / p->block_size : 0ull;
        const unsigned long long w = p->block_size != 0u ? p->ell_cols / p->block_size : 0ull;
        csamd_blocked_ell_view &v = out_views[i];
        std::memset(&v, 0, sizeof(v));
        v.rows = p->rows;
        v.cols = p->cols;
        v.nnz = p->nnz;
        v.block_size = p->block_size;
        v.ell_cols = p->ell_cols;
        v.blockColIdx = (unsigned int *) (base + o_bci[i]);
        v.val = (csamd_half *) (base + o_val[i]);
        if (!copy_h2d(v.blockColIdx, p->blockColIdx, rb * w * 4u, s) || !copy_h2d(v.val, p->val, (size_t) p->rows * p->ell_cols * 2u, s)) {
            (void) hipFree(base);
            return 0;
        }
    }
    if (count != 0u && !hip_ok(hipMemcpy(base + o_views, out_views, (size_t) count * sizeof(csamd_blocked_ell_view), hipMemcpyHostToDevice),
                               "views H2D")) {
        (void) hipFree(base);
        return 0;
    }
    rec->allocation = base;
    rec->allocation_bytes = got;
    rec->storage = base;
    rec->view = base + o_views;
    rec->group_begin = 0ul;
    rec->group_end = count;
    rec->device_id = dev;
    return 1;
}

int csamd_upload_compressed_shard(const csamd_host_compressed *const *parts, unsigned int count, csamd_partition_record *rec,
                                  csamd_compressed_view *out_views, csamd_allocation_cache *cache, csamd_stream_t stream) {
    CSAMD_REQUIRE(parts != nullptr && rec != nullptr && out_views != nullptr, "upload_compressed_shard: null argument");
    zero_record(rec);
    layout_builder L;
    const size_t o_views = L.take((size_t) count * sizeof(csamd_compressed_view));
    std::vector<size_t> o_ptr(count), o_idx(count), o_val(count);
    for (unsigned int i = 0; i < count; ++i) {
        const csamd_host_compressed *p = parts[i];
        CSAMD_REQUIRE(p != nullptr, "upload_compressed_shard: null partition %u", i);
        const unsigned int major = p->axis == CSAMD_COMPRESSED_BY_COL ? p->cols : p->rows;
        o_ptr[i] = L.take((size_t) (major + 1u) * 4u);
        o_idx[i] = L.take((size_t) p->nnz * 4u);
        o_val[i] = L.take((size_t) p->nnz * 2u);
    }
    const int dev = current_device();
    size_t got = 0u;
    char *base = (char *) stage_alloc(cache, dev, L.off, &got);
    if (base == nullptr) return 0;
    hipStream_t s = (hipStream_t) stream;
    for (unsigned int i = 0; i < count; ++i) {
        const csamd_host_compressed *p = parts[i];
        const unsigned int major = p->axis == CSAMD_COMPRESSED_BY_COL ? p->cols : p->rows;
        csamd_compressed_view &v = out_views[i];
        std::memset(&v, 0, sizeof(v));
        v.rows = p->rows;
        v.cols = p->cols;
        v.nnz = p->nnz;
        v.axis = p->axis;
        v.majorPtr = (unsigned int *) (base + o_ptr[i]);
        v.minorIdx = (unsigned int *) (base + o_idx[i]);
        v.val = (csamd_half *) (base + o_val[i]);
        if (!copy_h2d(v.majorPtr, p->majorPtr, (size_t) (major + 1u) * 4u, s) || !copy_h2d(v.minorIdx, p->minorIdx, (size_t) p->nnz * 4u, s) ||
            !copy_h2d(v.val, p->val, (size_t) p->nnz * 2u, s)) {
            (void) hipFree(base);
            return 0;
        }
    }
    if (count != 0u && !hip_ok(hipMemcpy(base + o_views, out_views, (size_t) count * sizeof(csamd_compressed_view), hipMemcpyHostToDevice),
                               "views H2D")) {
        (void) hipFree(base);
        return 0;
    }
    rec->allocation = base;
    rec->allocation_bytes = got;
    rec->storage = base;
    rec->view = base + o_views;
    rec->group_begin = 0ul;
    rec->group_end = count;
    rec->device_id = dev;
    return 1;
}

} // extern "C"
