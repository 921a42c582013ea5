// sharded<MatrixT> metadata and sharded_device<MatrixT> residency.
//
// Reference: include/CellShard/runtime/layout/sharded.cuh (layout, lookups,
// byte estimates), runtime/host/sharded_host.cuh (metadata operations:
// reserve / define / append / concatenate / shard cuts / reshard / drop),
// runtime/device/sharded_device.cuh (residency table, partition and packed
// shard uploads, release through the allocation cache, stage and swap).
//
// Everything here is host code.  The metadata layout does not depend on the
// part type, so the operations take the part format where they read a part
// (its rows / nnz / aux word / host bytes) or destroy one.  Device bytes are
// this library's upload footprint (staging.hip: the device view at offset 0,
// then every array 256-byte aligned), which is what a shard cut by device
// bytes has to budget for.  Staging a shard whose host parts are not
// materialized reads the shard's CSPACK01 pack straight into HBM and
// rebuilds each canonical partition there (cspack.hip), where the reference
// fetches the parts into host memory first and uploads them.
#include "common.hh"

#include <cstdlib>
#include <vector>

namespace csamd {
namespace {

constexpr size_t STAGE_ALIGN = 256u;

struct part_meta {
    unsigned long rows, cols, nnz, aux;
};

bool format_ok(unsigned int f) {
    return f == CSAMD_FORMAT_COMPRESSED || f == CSAMD_FORMAT_BLOCKED_ELL || f == CSAMD_FORMAT_SLICED_ELL;
}

unsigned long sell_total_slots(const csamd_host_sliced_ell *m) {
    unsigned long long t = 0ull;
    if (m->slice_row_offsets == nullptr || m->slice_widths == nullptr) return 0ul;
    for (unsigned int s = 0; s < m->slice_count; ++s)
        t += (unsigned long long) (m->slice_row_offsets[s + 1u] - m->slice_row_offsets[s]) * m->slice_widths[s];
    return (unsigned long) (unsigned int) t;  // the reference packs a u32 total
}

unsigned long aux_of(unsigned int format, const void *p) {
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: return ((const csamd_host_compressed *) p)->axis;
        case CSAMD_FORMAT_BLOCKED_ELL: {
            const csamd_host_blocked_ell *m = (const csamd_host_blocked_ell *) p;
            const unsigned long w = m->block_size == 0u || m->ell_cols == 0u ? 0ul : m->ell_cols / m->block_size;
            return (w << 16) | (unsigned long) m->block_size;
        }
        case CSAMD_FORMAT_SLICED_ELL: {
            const csamd_host_sliced_ell *m = (const csamd_host_sliced_ell *) p;
            return (sell_total_slots(m) << 32) | (unsigned long) m->slice_count;
        }
        default: return 0ul;
    }
}

part_meta meta_of(unsigned int format, const void *p) {
    // every host struct starts with rows, cols, nnz (u32)
    const unsigned int *h = (const unsigned int *) p;
    return part_meta{h[0], h[1], h[2], aux_of(format, p)};
}

void destroy_part(unsigned int format, void *p) {
    if (p == nullptr || format == 0u) return;
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: csamd_host_compressed_clear((csamd_host_compressed *) p); break;
        case CSAMD_FORMAT_BLOCKED_ELL: csamd_host_blocked_ell_clear((csamd_host_blocked_ell *) p); break;
        case CSAMD_FORMAT_SLICED_ELL: csamd_host_sliced_ell_clear((csamd_host_sliced_ell *) p); break;
        default: return;
    }
    std::free(p);
}

// formats/*.cuh bytes() of a materialized part
size_t host_part_bytes(unsigned int format, const void *p) {
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: {
            const csamd_host_compressed *m = (const csamd_host_compressed *) p;
            const size_t major = m->axis == CSAMD_COMPRESSED_BY_COL ? m->cols : m->rows;
            return sizeof(*m) + (major + 1u) * 4u + (size_t) m->nnz * 6u;
        }
        case CSAMD_FORMAT_BLOCKED_ELL: {
            const csamd_host_blocked_ell *m = (const csamd_host_blocked_ell *) p;
            const size_t rb = m->block_size == 0u ? 0u : (m->rows + m->block_size - 1u) / m->block_size;
            const size_t w = m->block_size == 0u || m->ell_cols == 0u ? 0u : m->ell_cols / m->block_size;
            return sizeof(*m) + rb * w * 4u + (size_t) m->rows * m->ell_cols * 2u;
        }
        case CSAMD_FORMAT_SLICED_ELL: {
            const csamd_host_sliced_ell *m = (const csamd_host_sliced_ell *) p;
            const size_t slots = sell_total_slots(m);
            const size_t w_off = align_up(((size_t) m->slice_count + 1u) * 4u, 4u);
            const size_t col_off = align_up(w_off + (size_t) m->slice_count * 4u, 4u);
            const size_t val_off = align_up(col_off + slots * 4u, 2u);
            return sizeof(*m) + val_off + slots * 2u;
        }
        default: return 0u;
    }
}

// sharded.cuh:262-320: from the metadata when the part is not materialized
size_t partition_bytes(const csamd_sharded *m, unsigned int format, unsigned long p) {
    if (p >= m->num_partitions) return 0u;
    if (m->parts != nullptr && m->parts[p] != nullptr) return host_part_bytes(format, m->parts[p]);
    const unsigned long rows = m->partition_rows[p], nnz = m->partition_nnz[p], aux = m->partition_aux[p];
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: {
            const unsigned long major = aux == CSAMD_COMPRESSED_BY_COL ? m->cols : rows;
            return sizeof(csamd_host_compressed) + (size_t) (major + 1u) * 4u + (size_t) nnz * 6u;
        }
        case CSAMD_FORMAT_BLOCKED_ELL: {
            const unsigned long bs = aux & 0xfffful, w = aux >> 16;
            const size_t rb = bs == 0u ? 0u : (rows + bs - 1u) / bs;
            return sizeof(csamd_host_blocked_ell) + rb * w * 4u + (size_t) rows * (w * bs) * 2u;
        }
        case CSAMD_FORMAT_SLICED_ELL: {
            const size_t sc = aux & 0xfffffffful, slots = aux >> 32;
            return sizeof(csamd_host_sliced_ell) + (sc + 1u) * 4u + sc * 4u + slots * 6u;
        }
        default: return 0u;
    }
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

// The device footprint of one partition uploaded on its own (staging.hip)
size_t device_partition_bytes(const csamd_sharded *v, unsigned int format, unsigned long p) {
    if (p >= v->num_partitions) return 0u;
    const unsigned long rows = v->partition_rows[p], nnz = v->partition_nnz[p], aux = v->partition_aux[p];
    layout_builder L;
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: {
            const unsigned long major = aux == CSAMD_COMPRESSED_BY_COL ? v->cols : rows;
            L.take(sizeof(csamd_compressed_view));
            L.take((size_t) (major + 1u) * 4u);
            L.take((size_t) nnz * 4u);
            L.take((size_t) nnz * 2u);
            return L.off;
        }
        case CSAMD_FORMAT_BLOCKED_ELL: {
            const unsigned long bs = aux & 0xfffful, w = aux >> 16;
            const size_t rb = bs == 0u ? 0u : (rows + bs - 1u) / bs;
            L.take(sizeof(csamd_blocked_ell_view));
            L.take(rb * w * 4u);
            L.take((size_t) rows * (w * bs) * 2u);
            return L.off;
        }
        case CSAMD_FORMAT_SLICED_ELL: {
            const size_t sc = aux & 0xfffffffful, slots = aux >> 32;
            L.take(sizeof(csamd_sliced_ell_view));
            L.take((sc + 1u) * 4u);
            L.take(sc * 4u);
            L.take(sc * 4u);
            L.take(slots * 4u);
            L.take(slots * 2u);
            return L.off;
        }
        default: return 0u;
    }
}

unsigned long first_part(const csamd_sharded *m, unsigned long s) { return csamd_sharded_first_partition_in_shard(m, s); }
unsigned long last_part(const csamd_sharded *m, unsigned long s) { return csamd_sharded_last_partition_in_shard(m, s); }

bool shard_loaded(const csamd_sharded *m, unsigned long s) {
    if (s >= m->num_shards) return false;
    for (unsigned long i = first_part(m, s); i < last_part(m, s); ++i)
        if (m->parts == nullptr || m->parts[i] == nullptr) return false;
    return true;
}

// greedy cuts shared by set_shards_by_nnz / _by_part_bytes / _by_device_bytes
// (sharded_host.cuh:329-389, sharded_device.cuh:1111-1140): a partition of
// weight 0 joins the current shard; a new shard starts where the next
// partition would take the current one past max.
template<typename Weight>
int cut_shards(csamd_sharded *m, unsigned long long max, Weight weight) {
    if (max == 0ull) return csamd_sharded_set_shards_to_partitions(m);
    if (m->num_partitions == 0ul) {  // nothing to cut (and no table to write)
        m->num_shards = 0ul;
        return 1;
    }
    if (!csamd_sharded_reserve_shards(m, m->num_partitions)) return 0;
    m->shard_offsets[0] = 0ul;
    unsigned long count = 0ul;
    unsigned long long used = 0ull;
    for (unsigned long i = 0; i < m->num_partitions; ++i) {
        const unsigned long long w = weight(i);
        if (w == 0ull) continue;
        if (used != 0ull && used + w > max) {
            m->shard_offsets[++count] = m->partition_offsets[i];
            used = 0ull;
        }
        used += w;
    }
    if (m->num_partitions != 0ul) m->shard_offsets[++count] = m->rows;
    m->num_shards = count;
    csamd_sharded_rebuild_shard_parts(m);
    return 1;
}

// ---------------------------------------------------------------------------
// residency
csamd_allocation_cache cache_of(const csamd_sharded_device *s) {
    csamd_allocation_cache c;
    c.cache_count = s->cache_count;
    c.cache_capacity = s->cache_capacity;
    c.cache = s->cache;
    return c;
}
void cache_back(csamd_sharded_device *s, const csamd_allocation_cache &c) {
    s->cache_count = c.cache_count;
    s->cache_capacity = c.cache_capacity;
    s->cache = c.cache;
}

void zero_record(csamd_partition_record *r) {
    std::memset(r, 0, sizeof(*r));
    r->device_id = -1;
}

int set_device(int device) { return hip_ok(hipSetDevice(device), "hipSetDevice"); }

int release_group(csamd_sharded_device *s, unsigned long part) {
    csamd_partition_record &r = s->parts[part];
    if (!set_device(r.device_id >= 0 ? r.device_id : 0)) return 0;
    const unsigned long b = r.group_begin, e = r.group_end;
    csamd_allocation_cache c = cache_of(s);
    int ok = 1;
    if (e <= b || e > s->capacity) {
        // not part of a group: the record's own allocation is freed
        if (r.allocation != nullptr) ok = hip_ok(hipFree(r.allocation), "hipFree partition");
        zero_record(&r);
        return ok;
    }
    const csamd_partition_record &owner = s->parts[b];
    if (owner.allocation != nullptr) ok = csamd_cache_allocation(&c, owner.device_id, owner.allocation, owner.allocation_bytes);
    else if (owner.storage != nullptr) ok = hip_ok(hipFree(owner.storage), "hipFree partition storage");
    cache_back(s, c);
    for (unsigned long i = b; i < e; ++i) zero_record(&s->parts[i]);
    return ok;
}

int upload_one(csamd_sharded_device *s, const csamd_sharded *v, unsigned int format, unsigned long p, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(p < v->num_partitions && p < s->capacity, "upload_partition: partition %lu outside the table", p);
    CSAMD_REQUIRE(v->parts != nullptr && v->parts[p] != nullptr, "upload_partition: host partition %lu is not materialized", p);
    csamd_partition_record &r = s->parts[p];
    if (r.view != nullptr) {
        if (r.device_id == device) return 1;
        if (!release_group(s, p)) return 0;
    }
    if (!set_device(device)) return 0;
    csamd_allocation_cache c = cache_of(s);
    int ok = 0;
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: ok = csamd_upload_compressed((const csamd_host_compressed *) v->parts[p], &r, nullptr, &c, stream); break;
        case CSAMD_FORMAT_BLOCKED_ELL: ok = csamd_upload_blocked_ell((const csamd_host_blocked_ell *) v->parts[p], &r, nullptr, &c, stream); break;
        case CSAMD_FORMAT_SLICED_ELL: ok = csamd_upload_sliced_ell((const csamd_host_sliced_ell *) v->parts[p], &r, nullptr, &c, stream); break;
        default: set_error("upload_partition: unknown format %u", format);
    }
    cache_back(s, c);
    if (!ok) return 0;
    r.group_begin = p;
    r.group_end = p + 1u;
    r.device_id = device;
    return 1;
}

// The partition's device arrays in its record (a0.. as the single-partition
// uploads set them), for callers that read partitions of a packed shard.
inline void set_record_arrays(csamd_partition_record &r, const csamd_compressed_view &v) {
    r.a0 = v.majorPtr;
    r.a1 = v.minorIdx;
    r.a2 = v.val;
}
inline void set_record_arrays(csamd_partition_record &r, const csamd_blocked_ell_view &v) {
    r.a0 = v.blockColIdx;
    r.a1 = v.val;
}

template<typename HostT, typename ViewT, typename Upload>
int upload_packed(csamd_sharded_device *s, const csamd_sharded *v, unsigned long b, unsigned long e, int device, csamd_stream_t stream,
                  Upload upload) {
    const unsigned long n = e - b;
    CSAMD_REQUIRE(n <= 0xffffffffull, "upload_shard: too many partitions");
    std::vector<const HostT *> hp(n);
    std::vector<ViewT> views(n);
    for (unsigned long i = 0; i < n; ++i) hp[i] = (const HostT *) v->parts[b + i];
    csamd_partition_record rec;
    zero_record(&rec);
    csamd_allocation_cache c = cache_of(s);
    const int ok = upload(hp.data(), (unsigned int) n, &rec, views.data(), &c, stream);
    cache_back(s, c);
    if (!ok) return 0;
    for (unsigned long i = 0; i < n; ++i) {
        csamd_partition_record &r = s->parts[b + i];
        zero_record(&r);
        r.allocation = i == 0u ? rec.allocation : nullptr;
        r.allocation_bytes = i == 0u ? rec.allocation_bytes : 0u;
        r.storage = i == 0u ? rec.storage : nullptr;
        r.view = (char *) rec.view + i * sizeof(ViewT);
        set_record_arrays(r, views[i]);
        r.group_begin = b;
        r.group_end = e;
        r.device_id = device;
    }
    return 1;
}

// Rebuild partition p from the shard pack straight into a device record:
// stage the bucketed execution partition, its canonical CSR, then the
// canonical Blocked-ELL at the block size of the partition's aux word.
int stage_bell_from_pack(csamd_sharded_device *s, const csamd_sharded *v, const csamd_cspack_file *pack, unsigned long p,
                         unsigned long local, int device, csamd_stream_t stream) {
    const unsigned long aux = v->partition_aux[p];
    const unsigned int bs = (unsigned int) (aux & 0xfffful);
    const unsigned int w = (unsigned int) (aux >> 16);
    CSAMD_REQUIRE(bs != 0u, "stage_shard: partition %lu has no block size in its aux word", p);
    csamd_cspack_workspace pw;
    csamd_cspack_init(&pw);
    csamd_convert_workspace cw;
    csamd_convert_init(&cw);
    csamd_bucketed_blocked_ell part;
    std::memset(&part, 0, sizeof(part));
    unsigned int *d_ptr = nullptr, *d_idx = nullptr;
    void *d_val = nullptr;
    char *base = nullptr;
    size_t got = 0u;
    int ok = 0;
    do {
        if (!csamd_cspack_setup(&pw, device, stream) || !csamd_convert_setup(&cw, device, stream)) break;
        if (!csamd_cspack_stage_blocked_ell(pack, local, CSAMD_VALUE_F16, &pw, &part)) break;
        if (part.rows != v->partition_rows[p] || (v->cols != 0ul && part.cols != v->cols)) {
            set_error("stage_shard: pack partition %lu is %u x %u, the metadata says %lu x %lu", local, part.rows, part.cols,
                      v->partition_rows[p], v->cols);
            break;
        }
        if (!hip_ok(hipMalloc((void **) &d_ptr, ((size_t) part.rows + 1u) * 4u), "hipMalloc stage ptr")) break;
        unsigned long long nnz = 0ull, nnz2 = 0ull;
        if (!csamd_canonical_compressed_from_bucketed_blocked_ell(&part, CSAMD_VALUE_F16, d_ptr, 0, nullptr, nullptr, 0ull, &nnz, &cw)) break;
        if (!hip_ok(hipMalloc((void **) &d_idx, nnz * 4u + 4u), "hipMalloc stage idx") ||
            !hip_ok(hipMalloc(&d_val, nnz * 2u + 2u), "hipMalloc stage val"))
            break;
        if (!csamd_canonical_compressed_from_bucketed_blocked_ell(&part, CSAMD_VALUE_F16, d_ptr, 1, d_idx, d_val, nnz, &nnz2, &cw)) break;
        csamd_compressed_view csr;
        std::memset(&csr, 0, sizeof(csr));
        csr.rows = part.rows;
        csr.cols = part.cols;
        csr.nnz = (unsigned int) nnz;
        csr.axis = CSAMD_COMPRESSED_BY_ROW;
        csr.majorPtr = d_ptr;
        csr.minorIdx = d_idx;
        csr.val = (csamd_half *) d_val;
        unsigned int width = 0u;
        unsigned long long blocks = 0ull;
        double fill = 0.0;
        if (!csamd_blocked_ell_shape(&csr, bs, &cw, &width, &blocks, &fill)) break;
        if (width != w) {
            set_error("stage_shard: partition %lu rebuilds %u block columns wide, its aux word says %u", p, width, w);
            break;
        }
        const size_t rb = ((size_t) part.rows + bs - 1u) / bs;
        layout_builder L;
        const size_t o_view = L.take(sizeof(csamd_blocked_ell_view));
        const size_t o_bci = L.take(rb * w * 4u);
        const size_t o_val = L.take((size_t) part.rows * w * bs * 2u);
        csamd_allocation_cache c = cache_of(s);
        base = (char *) csamd_take_cached_allocation(&c, device, L.off, &got);
        cache_back(s, c);
        if (base == nullptr) {
            if (!hip_ok(hipMalloc((void **) &base, L.off), "hipMalloc staged partition")) break;
            got = L.off;
        }
        csamd_blocked_ell_view dst;
        std::memset(&dst, 0, sizeof(dst));
        dst.rows = part.rows;
        dst.cols = part.cols;
        dst.block_size = bs;
        dst.ell_cols = w * bs;
        dst.blockColIdx = rb * w != 0u ? (unsigned int *) (base + o_bci) : nullptr;
        dst.val = (size_t) part.rows * w != 0u ? (csamd_half *) (base + o_val) : nullptr;
        if (!csamd_blocked_ell_from_compressed(&csr, CSAMD_VALUE_F16, &dst, &cw)) break;
        if (!hip_ok(hipMemcpy(base + o_view, &dst, sizeof(dst), hipMemcpyHostToDevice), "staged view H2D")) break;
        csamd_partition_record &r = s->parts[p];
        zero_record(&r);
        r.allocation = base;
        r.allocation_bytes = got;
        r.storage = base;
        r.view = base + o_view;
        r.a0 = dst.blockColIdx;
        r.a1 = dst.val;
        r.group_begin = p;
        r.group_end = p + 1u;
        r.device_id = device;
        base = nullptr;  // owned by the record now
        ok = 1;
    } while (false);
    if (base != nullptr) (void) hipFree(base);
    if (d_val != nullptr) (void) hipFree(d_val);
    if (d_idx != nullptr) (void) hipFree(d_idx);
    if (d_ptr != nullptr) (void) hipFree(d_ptr);
    csamd_bucketed_blocked_ell_clear(&part);
    csamd_convert_clear(&cw);
    csamd_cspack_clear(&pw);
    return ok;
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_sharded_init(csamd_sharded *m) {
    if (m != nullptr) std::memset(m, 0, sizeof(*m));
}

void csamd_sharded_clear(csamd_sharded *m, unsigned int format) {
    if (m == nullptr) return;
    if (m->parts != nullptr)
        for (unsigned long i = 0; i < m->num_partitions; ++i) destroy_part(format, m->parts[i]);
    std::free(m->parts);
    std::free(m->partition_offsets);
    std::free(m->partition_rows);
    std::free(m->partition_nnz);
    std::free(m->partition_aux);
    std::free(m->shard_offsets);
    std::free(m->shard_parts);
    csamd_sharded_init(m);
}

int csamd_sharded_reserve_partitions(csamd_sharded *m, unsigned long capacity) {
    CSAMD_REQUIRE(m != nullptr, "reserve_partitions: null matrix");
    if (capacity <= m->partition_capacity) return 1;
    void **parts = (void **) std::calloc(capacity, sizeof(void *));
    unsigned long *off = (unsigned long *) std::calloc(capacity + 1u, sizeof(unsigned long));
    unsigned long *rows = (unsigned long *) std::calloc(capacity, sizeof(unsigned long));
    unsigned long *nnz = (unsigned long *) std::calloc(capacity, sizeof(unsigned long));
    unsigned long *aux = (unsigned long *) std::calloc(capacity, sizeof(unsigned long));
    if (parts == nullptr || off == nullptr || rows == nullptr || nnz == nullptr || aux == nullptr) {
        std::free(parts);
        std::free(off);
        std::free(rows);
        std::free(nnz);
        std::free(aux);
        set_error("reserve_partitions: out of host memory for %lu partitions", capacity);
        return 0;
    }
    const unsigned long n = m->num_partitions;
    if (n != 0ul) {
        std::memcpy(parts, m->parts, n * sizeof(void *));
        std::memcpy(off, m->partition_offsets, (n + 1u) * sizeof(unsigned long));
        std::memcpy(rows, m->partition_rows, n * sizeof(unsigned long));
        std::memcpy(nnz, m->partition_nnz, n * sizeof(unsigned long));
        std::memcpy(aux, m->partition_aux, n * sizeof(unsigned long));
    }
    std::free(m->parts);
    std::free(m->partition_offsets);
    std::free(m->partition_rows);
    std::free(m->partition_nnz);
    std::free(m->partition_aux);
    m->parts = parts;
    m->partition_offsets = off;
    m->partition_rows = rows;
    m->partition_nnz = nnz;
    m->partition_aux = aux;
    m->partition_capacity = capacity;
    return 1;
}

int csamd_sharded_reserve_shards(csamd_sharded *m, unsigned long capacity) {
    CSAMD_REQUIRE(m != nullptr, "reserve_shards: null matrix");
    if (capacity <= m->shard_capacity) return 1;
    unsigned long *off = (unsigned long *) std::calloc(capacity + 1u, sizeof(unsigned long));
    unsigned long *sp = (unsigned long *) std::calloc(capacity + 1u, sizeof(unsigned long));
    if (off == nullptr || sp == nullptr) {
        std::free(off);
        std::free(sp);
        set_error("reserve_shards: out of host memory for %lu shards", capacity);
        return 0;
    }
    if (m->shard_offsets != nullptr && m->num_shards != 0ul) {
        std::memcpy(off, m->shard_offsets, (m->num_shards + 1u) * sizeof(unsigned long));
        if (m->shard_parts != nullptr) std::memcpy(sp, m->shard_parts, (m->num_shards + 1u) * sizeof(unsigned long));
    }
    std::free(m->shard_offsets);
    std::free(m->shard_parts);
    m->shard_offsets = off;
    m->shard_parts = sp;
    m->shard_capacity = capacity;
    return 1;
}

void csamd_sharded_rebuild_shard_parts(csamd_sharded *m) {
    if (m == nullptr || m->shard_parts == nullptr || m->shard_offsets == nullptr || m->partition_offsets == nullptr) return;
    unsigned long part = 0ul;
    for (unsigned long s = 0; s < m->num_shards; ++s) {
        while (part < m->num_partitions && m->partition_offsets[part] < m->shard_offsets[s]) ++part;
        m->shard_parts[s] = part;
    }
    m->shard_parts[m->num_shards] = m->num_partitions;
}

void csamd_sharded_rebuild_partition_offsets(csamd_sharded *m) {
    if (m == nullptr) return;
    m->rows = 0ul;
    m->nnz = 0ul;
    if (m->partition_offsets == nullptr) return;
    m->partition_offsets[0] = 0ul;
    for (unsigned long i = 0; i < m->num_partitions; ++i) {
        m->partition_offsets[i + 1u] = m->partition_offsets[i] + m->partition_rows[i];
        m->nnz += m->partition_nnz[i];
    }
    m->rows = m->partition_offsets[m->num_partitions];
}

int csamd_sharded_set_shards_to_partitions(csamd_sharded *m) {
    CSAMD_REQUIRE(m != nullptr, "set_shards_to_partitions: null matrix");
    if (!csamd_sharded_reserve_shards(m, m->num_partitions)) return 0;
    m->num_shards = m->num_partitions;
    if (m->shard_offsets == nullptr) return 1;  // no partitions and nothing reserved yet
    for (unsigned long i = 0; i <= m->num_partitions; ++i) {
        m->shard_offsets[i] = m->partition_offsets != nullptr ? m->partition_offsets[i] : 0ul;
        m->shard_parts[i] = i;
    }
    return 1;
}

int csamd_sharded_define_partitions(csamd_sharded *m, unsigned int format, unsigned long cols, unsigned long n,
                                    const unsigned long *rows, const unsigned long *nnz, const unsigned long *aux) {
    CSAMD_REQUIRE(m != nullptr, "define_partitions: null matrix");
    csamd_sharded_clear(m, format);
    if (!csamd_sharded_reserve_partitions(m, n)) return 0;
    m->cols = cols;
    m->num_partitions = n;
    for (unsigned long i = 0; i < n; ++i) {
        m->parts[i] = nullptr;
        m->partition_rows[i] = rows != nullptr ? rows[i] : 0ul;
        m->partition_nnz[i] = nnz != nullptr ? nnz[i] : 0ul;
        m->partition_aux[i] = aux != nullptr ? aux[i] : 0ul;
    }
    csamd_sharded_rebuild_partition_offsets(m);
    return csamd_sharded_set_shards_to_partitions(m);
}

int csamd_sharded_append_partition(csamd_sharded *m, unsigned int format, void *part) {
    CSAMD_REQUIRE(m != nullptr, "append_partition: null matrix");
    CSAMD_REQUIRE(part == nullptr || format_ok(format), "append_partition: unknown format %u", format);
    if (m->num_partitions == m->partition_capacity &&
        !csamd_sharded_reserve_partitions(m, m->partition_capacity == 0ul ? 4ul : m->partition_capacity << 1))
        return 0;
    const unsigned long i = m->num_partitions;
    m->parts[i] = part;
    const part_meta pm = part != nullptr ? meta_of(format, part) : part_meta{0ul, 0ul, 0ul, 0ul};
    m->partition_rows[i] = pm.rows;
    m->partition_nnz[i] = pm.nnz;
    m->partition_aux[i] = pm.aux;
    ++m->num_partitions;
    if (part != nullptr && m->cols == 0ul) m->cols = pm.cols;
    csamd_sharded_rebuild_partition_offsets(m);
    return csamd_sharded_set_shards_to_partitions(m);
}

int csamd_sharded_concatenate(csamd_sharded *dst, csamd_sharded *src) {
    CSAMD_REQUIRE(dst != nullptr && src != nullptr, "concatenate: null matrix");
    if (src->num_partitions == 0ul) return 1;
    const unsigned long n0 = dst->num_partitions, n = src->num_partitions;
    if (!csamd_sharded_reserve_partitions(dst, n0 + n)) return 0;
    for (unsigned long i = 0; i < n; ++i) {
        dst->parts[n0 + i] = src->parts[i];
        dst->partition_rows[n0 + i] = src->partition_rows[i];
        dst->partition_nnz[n0 + i] = src->partition_nnz[i];
        dst->partition_aux[n0 + i] = src->partition_aux[i];
        src->parts[i] = nullptr;
        src->partition_rows[i] = src->partition_nnz[i] = src->partition_aux[i] = 0ul;
    }
    dst->num_partitions = n0 + n;
    src->num_partitions = 0ul;
    if (dst->cols == 0ul) dst->cols = src->cols;
    csamd_sharded_rebuild_partition_offsets(dst);
    if (!csamd_sharded_set_shards_to_partitions(dst)) return 0;
    csamd_sharded_rebuild_partition_offsets(src);
    src->num_shards = 0ul;
    if (src->shard_parts != nullptr) src->shard_parts[0] = 0ul;
    return 1;
}

int csamd_sharded_set_equal_shards(csamd_sharded *m, unsigned long count) {
    CSAMD_REQUIRE(m != nullptr, "set_equal_shards: null matrix");
    if (count == 0ul || m->num_partitions == 0ul) {
        m->num_shards = 0ul;
        return 1;
    }
    if (count >= m->num_partitions) return csamd_sharded_set_shards_to_partitions(m);
    if (!csamd_sharded_reserve_shards(m, count)) return 0;
    // a shard closes once it holds at least rows / count (rounded up) rows,
    // while enough partitions remain for the shards still to come
    const unsigned long target = (m->rows + count - 1u) / count;
    unsigned long made = 0ul, rows = 0ul;
    m->shard_offsets[0] = 0ul;
    for (unsigned long i = 0; i < m->num_partitions; ++i) {
        const unsigned long parts_left = m->num_partitions - (i + 1u);
        const unsigned long shards_left = count - (made + 1u);
        rows += m->partition_rows[i];
        if (shards_left == 0ul) continue;
        if (rows >= target && parts_left >= shards_left) {
            m->shard_offsets[++made] = m->partition_offsets[i + 1u];
            rows = 0ul;
        }
    }
    m->shard_offsets[++made] = m->rows;
    m->num_shards = made;
    csamd_sharded_rebuild_shard_parts(m);
    return 1;
}

int csamd_sharded_build_shard_offsets_by_rows(const csamd_sharded *m, unsigned long target, unsigned long **out_offsets,
                                              unsigned long *out_count) {
    CSAMD_REQUIRE(m != nullptr && out_offsets != nullptr && out_count != nullptr, "build_shard_offsets_by_rows: null argument");
    *out_offsets = nullptr;
    *out_count = 0ul;
    if (m->num_partitions == 0ul) return 1;
    unsigned long *off = (unsigned long *) std::calloc(m->num_partitions + 1u, sizeof(unsigned long));
    CSAMD_REQUIRE(off != nullptr, "build_shard_offsets_by_rows: out of host memory");
    if (target == 0ul) {
        for (unsigned long i = 0; i <= m->num_partitions; ++i) off[i] = m->partition_offsets[i];
        *out_offsets = off;
        *out_count = m->num_partitions;
        return 1;
    }
    unsigned long count = 0ul, used = 0ul;
    for (unsigned long i = 0; i < m->num_partitions; ++i) {
        const unsigned long r = m->partition_rows[i];
        if (used != 0ul && used + r > target) {
            off[++count] = m->partition_offsets[i];
            used = 0ul;
        }
        used += r;
    }
    off[++count] = m->rows;
    *out_offsets = off;
    *out_count = count;
    return 1;
}

int csamd_sharded_reshard(csamd_sharded *m, unsigned long count, const unsigned long *offsets) {
    CSAMD_REQUIRE(m != nullptr, "reshard: null matrix");
    if (count == 0ul || offsets == nullptr) {
        m->num_shards = 0ul;
        return 1;
    }
    CSAMD_REQUIRE(offsets[0] == 0ul && offsets[count] == m->rows, "reshard: cuts must run from 0 to %lu rows", m->rows);
    for (unsigned long i = 0; i < count; ++i) {
        CSAMD_REQUIRE(offsets[i] <= offsets[i + 1u], "reshard: cut %lu decreases", i);
        CSAMD_REQUIRE(csamd_sharded_row_is_partition_boundary(m, offsets[i]), "reshard: row %lu is not a partition boundary", offsets[i]);
    }
    if (!csamd_sharded_reserve_shards(m, count)) return 0;
    m->num_shards = count;
    for (unsigned long i = 0; i <= count; ++i) m->shard_offsets[i] = offsets[i];
    csamd_sharded_rebuild_shard_parts(m);
    return 1;
}

int csamd_sharded_set_shards_by_nnz(csamd_sharded *m, unsigned long max_nnz) {
    CSAMD_REQUIRE(m != nullptr, "set_shards_by_nnz: null matrix");
    return cut_shards(m, max_nnz, [&](unsigned long i) { return (unsigned long long) m->partition_nnz[i]; });
}

int csamd_sharded_set_shards_by_part_bytes(csamd_sharded *m, unsigned int format, size_t max_bytes) {
    CSAMD_REQUIRE(m != nullptr && format_ok(format), "set_shards_by_part_bytes: bad arguments");
    return cut_shards(m, max_bytes, [&](unsigned long i) { return (unsigned long long) partition_bytes(m, format, i); });
}

int csamd_sharded_drop_partition(csamd_sharded *m, unsigned int format, unsigned long part) {
    CSAMD_REQUIRE(m != nullptr && part < m->num_partitions, "drop_partition: partition %lu out of range", part);
    destroy_part(format, m->parts[part]);
    m->parts[part] = nullptr;
    return 1;
}

int csamd_sharded_drop_shard(csamd_sharded *m, unsigned int format, unsigned long shard) {
    CSAMD_REQUIRE(m != nullptr && shard < m->num_shards, "drop_shard: shard %lu out of range", shard);
    for (unsigned long i = first_part(m, shard); i < last_part(m, shard); ++i)
        if (!csamd_sharded_drop_partition(m, format, i)) return 0;
    return 1;
}

void *csamd_host_part_alloc(unsigned int format) {
    switch (format) {
        case CSAMD_FORMAT_COMPRESSED: return std::calloc(1, sizeof(csamd_host_compressed));
        case CSAMD_FORMAT_BLOCKED_ELL: return std::calloc(1, sizeof(csamd_host_blocked_ell));
        case CSAMD_FORMAT_SLICED_ELL: return std::calloc(1, sizeof(csamd_host_sliced_ell));
        default: set_error("host_part_alloc: unknown format %u", format); return nullptr;
    }
}

int csamd_host_compressed_allocate(csamd_host_compressed *m) {
    CSAMD_REQUIRE(m != nullptr, "compressed allocate: null matrix");
    const unsigned int rows = m->rows, cols = m->cols, nnz = m->nnz, axis = m->axis;
    csamd_host_compressed_clear(m);
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->axis = axis;
    const size_t major = axis == CSAMD_COMPRESSED_BY_COL ? cols : rows;
    const size_t idx_off = align_up((major + 1u) * 4u, 4u);
    const size_t val_off = align_up(idx_off + (size_t) nnz * 4u, 2u);
    const size_t total = val_off + (size_t) nnz * 2u;
    m->storage = std::malloc(total);
    CSAMD_REQUIRE(m->storage != nullptr, "compressed allocate: malloc(%zu) failed", total);
    m->majorPtr = (unsigned int *) m->storage;
    m->minorIdx = nnz != 0u ? (unsigned int *) ((char *) m->storage + idx_off) : nullptr;
    m->val = nnz != 0u ? (csamd_half *) ((char *) m->storage + val_off) : nullptr;
    return 1;
}

unsigned long csamd_partition_aux(unsigned int format, const void *part) { return part != nullptr ? aux_of(format, part) : 0ul; }

static unsigned long offset_span(unsigned long row, const unsigned long *offsets, unsigned long count) {
    unsigned long lo = 0ul, hi = count;
    while (lo < hi) {
        const unsigned long mid = lo + ((hi - lo) >> 1);
        if (row < offsets[mid]) hi = mid;
        else if (row >= offsets[mid + 1u]) lo = mid + 1u;
        else return mid;
    }
    return count;
}

unsigned long csamd_sharded_find_partition(const csamd_sharded *m, unsigned long row) {
    if (m->partition_offsets == nullptr || m->num_partitions == 0ul) return m->num_partitions;
    return offset_span(row, m->partition_offsets, m->num_partitions);
}

unsigned long csamd_sharded_find_shard(const csamd_sharded *m, unsigned long row) {
    if (m->shard_offsets == nullptr || m->num_shards == 0ul) return m->num_shards;
    return offset_span(row, m->shard_offsets, m->num_shards);
}

unsigned long csamd_sharded_first_partition_in_shard(const csamd_sharded *m, unsigned long s) {
    if (m->shard_parts != nullptr && s < m->num_shards) return m->shard_parts[s];
    if (s >= m->num_shards || m->num_partitions == 0ul) return m->num_partitions;
    return csamd_sharded_find_partition(m, m->shard_offsets[s]);
}

unsigned long csamd_sharded_last_partition_in_shard(const csamd_sharded *m, unsigned long s) {
    if (m->shard_parts != nullptr && s < m->num_shards) return m->shard_parts[s + 1u];
    if (s >= m->num_shards) return m->num_partitions;
    const unsigned long end = m->shard_offsets[s + 1u];
    return end == 0ul ? 0ul : csamd_sharded_find_partition(m, end - 1u) + 1u;
}

int csamd_sharded_row_is_partition_boundary(const csamd_sharded *m, unsigned long row) {
    if (row == 0ul || row == m->rows) return 1;
    if (m->partition_offsets == nullptr || m->num_partitions == 0ul) return 0;
    const unsigned long hit = csamd_sharded_find_partition(m, row);
    return hit < m->num_partitions && m->partition_offsets[hit] == row ? 1 : 0;
}

size_t csamd_sharded_partition_bytes(const csamd_sharded *m, unsigned int format, unsigned long part) {
    return m != nullptr && format_ok(format) ? partition_bytes(m, format, part) : 0u;
}

size_t csamd_sharded_shard_bytes(const csamd_sharded *m, unsigned int format, unsigned long shard) {
    if (m == nullptr || !format_ok(format) || shard >= m->num_shards) return 0u;
    size_t t = 0u;
    for (unsigned long i = first_part(m, shard); i < last_part(m, shard); ++i) t += partition_bytes(m, format, i);
    return t;
}

size_t csamd_sharded_bytes(const csamd_sharded *m, unsigned int format) {
    if (m == nullptr || !format_ok(format)) return 0u;
    size_t t = sizeof(*m) + (size_t) m->partition_capacity * sizeof(void *) + ((size_t) m->partition_capacity + 1u) * 8u +
               (size_t) m->partition_capacity * 8u * 3u + ((size_t) m->shard_capacity + 1u) * 8u * 2u;
    for (unsigned long i = 0; i < m->num_partitions; ++i) t += partition_bytes(m, format, i);
    return t;
}

// ---------------------------------------------------------------------------
void csamd_sharded_device_init(csamd_sharded_device *s) {
    if (s != nullptr) std::memset(s, 0, sizeof(*s));
}

int csamd_sharded_device_clear(csamd_sharded_device *s) {
    if (s == nullptr) return 1;
    int ok = 1;
    if (s->parts != nullptr)
        for (unsigned long i = 0; i < s->capacity; ++i)
            if (s->parts[i].view != nullptr) ok &= release_group(s, i);
    csamd_allocation_cache c = cache_of(s);
    csamd_allocation_cache_clear(&c);
    std::free(s->parts);
    csamd_sharded_device_init(s);
    return ok;
}

int csamd_sharded_device_reserve(csamd_sharded_device *s, unsigned long capacity) {
    CSAMD_REQUIRE(s != nullptr, "sharded_device reserve: null state");
    if (capacity <= s->capacity) return 1;
    csamd_partition_record *r = (csamd_partition_record *) std::malloc(capacity * sizeof(csamd_partition_record));
    CSAMD_REQUIRE(r != nullptr, "sharded_device reserve: out of host memory");
    for (unsigned long i = 0; i < capacity; ++i) {
        if (i < s->capacity) r[i] = s->parts[i];
        else zero_record(&r[i]);
    }
    std::free(s->parts);
    s->parts = r;
    s->capacity = capacity;
    return 1;
}

size_t csamd_sharded_device_partition_bytes(const csamd_sharded *view, unsigned int format, unsigned long part) {
    return view != nullptr && format_ok(format) ? device_partition_bytes(view, format, part) : 0u;
}

size_t csamd_sharded_device_shard_bytes(const csamd_sharded *view, unsigned int format, unsigned long shard) {
    if (view == nullptr || !format_ok(format) || shard >= view->num_shards) return 0u;
    size_t t = 0u;
    for (unsigned long i = first_part(view, shard); i < last_part(view, shard); ++i) t += device_partition_bytes(view, format, i);
    return t;
}

int csamd_sharded_set_shards_by_device_bytes(csamd_sharded *view, unsigned int format, size_t max_bytes) {
    CSAMD_REQUIRE(view != nullptr && format_ok(format), "set_shards_by_device_bytes: bad arguments");
    return cut_shards(view, max_bytes, [&](unsigned long i) { return (unsigned long long) device_partition_bytes(view, format, i); });
}

int csamd_sharded_device_upload_partition(csamd_sharded_device *s, const csamd_sharded *view, unsigned int format,
                                          unsigned long part, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(s != nullptr && view != nullptr && format_ok(format), "upload_partition: bad arguments");
    return upload_one(s, view, format, part, device, stream);
}

int csamd_sharded_device_release_partition(csamd_sharded_device *s, unsigned long part) {
    CSAMD_REQUIRE(s != nullptr, "release_partition: null state");
    if (part >= s->capacity || s->parts[part].view == nullptr) return 1;
    return release_group(s, part);
}

int csamd_sharded_device_upload_shard(csamd_sharded_device *s, const csamd_sharded *view, unsigned int format, unsigned long shard,
                                      int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(s != nullptr && view != nullptr && format_ok(format), "upload_shard: bad arguments");
    CSAMD_REQUIRE(shard < view->num_shards, "upload_shard: shard %lu out of range", shard);
    const unsigned long b = first_part(view, shard), e = last_part(view, shard);
    if (b >= e) return 1;
    bool resident = true;
    for (unsigned long p = b; p < e; ++p) {
        CSAMD_REQUIRE(p < s->capacity && view->parts != nullptr && view->parts[p] != nullptr,
                      "upload_shard: partition %lu is outside the table or not materialized", p);
        resident = resident && s->parts[p].view != nullptr && s->parts[p].device_id == device;
    }
    if (resident) return 1;
    if (format == CSAMD_FORMAT_SLICED_ELL) {  // one partition at a time, as the reference does
        if (!set_device(device)) return 0;
        for (unsigned long p = b; p < e; ++p)
            if (!upload_one(s, view, format, p, device, stream)) return 0;
        return 1;
    }
    for (unsigned long p = b; p < e; ++p)
        if (s->parts[p].view != nullptr && !release_group(s, p)) return 0;
    if (!set_device(device)) return 0;
    if (format == CSAMD_FORMAT_COMPRESSED)
        return upload_packed<csamd_host_compressed, csamd_compressed_view>(s, view, b, e, device, stream, csamd_upload_compressed_shard);
    return upload_packed<csamd_host_blocked_ell, csamd_blocked_ell_view>(s, view, b, e, device, stream, csamd_upload_blocked_ell_shard);
}

int csamd_sharded_device_release_shard(csamd_sharded_device *s, const csamd_sharded *view, unsigned long shard) {
    CSAMD_REQUIRE(s != nullptr && view != nullptr, "release_shard: null argument");
    CSAMD_REQUIRE(shard < view->num_shards, "release_shard: shard %lu out of range", shard);
    for (unsigned long p = first_part(view, shard); p < last_part(view, shard); ++p) {
        if (p >= s->capacity || s->parts[p].view == nullptr) continue;
        if (!release_group(s, p)) return 0;
    }
    return 1;
}

int csamd_sharded_device_stage_shard(csamd_sharded_device *s, csamd_sharded *view, unsigned int format, const csamd_cspack_file *pack,
                                     unsigned long shard, int device, csamd_stream_t stream, int drop_host_after_upload) {
    CSAMD_REQUIRE(s != nullptr && view != nullptr && format_ok(format), "stage_shard: bad arguments");
    CSAMD_REQUIRE(shard < view->num_shards, "stage_shard: shard %lu out of range", shard);
    if (shard_loaded(view, shard)) {
        if (!csamd_sharded_device_upload_shard(s, view, format, shard, device, stream)) return 0;
    } else {
        CSAMD_REQUIRE(pack != nullptr, "stage_shard: shard %lu has parts that are not materialized and no pack to read", shard);
        CSAMD_REQUIRE(format == CSAMD_FORMAT_BLOCKED_ELL, "stage_shard: packs rebuild Blocked-ELL partitions only");
        const unsigned long b = first_part(view, shard), e = last_part(view, shard);
        CSAMD_REQUIRE(pack->partition_count == e - b, "stage_shard: the pack holds %llu partitions, shard %lu has %lu",
                      pack->partition_count, shard, e - b);
        for (unsigned long p = b; p < e; ++p) {
            CSAMD_REQUIRE(p < s->capacity, "stage_shard: partition %lu is outside the table", p);
            csamd_partition_record &r = s->parts[p];
            if (r.view != nullptr && r.device_id == device) continue;
            if (r.view != nullptr && !release_group(s, p)) return 0;
            if (view->parts != nullptr && view->parts[p] != nullptr) {
                if (!upload_one(s, view, format, p, device, stream)) return 0;
            } else {
                if (!set_device(device) || !stage_bell_from_pack(s, view, pack, p, p - b, device, stream)) return 0;
            }
        }
        if (stream != nullptr) CSAMD_HIP(hipStreamSynchronize((hipStream_t) stream), "stage_shard sync");
    }
    if (drop_host_after_upload) return csamd_sharded_drop_shard(view, format, shard);
    return 1;
}

int csamd_sharded_device_stage_partition(csamd_sharded_device *s, csamd_sharded *view, unsigned int format, const csamd_cspack_file *pack,
                                         unsigned long part, int device, csamd_stream_t stream, int drop_host_after_upload) {
    CSAMD_REQUIRE(s != nullptr && view != nullptr && format_ok(format), "stage_partition: bad arguments");
    CSAMD_REQUIRE(part < view->num_partitions && part < s->capacity, "stage_partition: partition %lu out of range", part);
    csamd_partition_record &r = s->parts[part];
    if (r.view == nullptr || r.device_id != device) {
        if (view->parts != nullptr && view->parts[part] != nullptr) {
            if (!upload_one(s, view, format, part, device, stream)) return 0;
        } else {
            CSAMD_REQUIRE(pack != nullptr, "stage_partition: partition %lu is not materialized and no pack to read", part);
            CSAMD_REQUIRE(format == CSAMD_FORMAT_BLOCKED_ELL, "stage_partition: packs rebuild Blocked-ELL partitions only");
            const unsigned long shard = csamd_sharded_find_shard(view, view->partition_offsets[part]);
            CSAMD_REQUIRE(shard < view->num_shards, "stage_partition: partition %lu is in no shard", part);
            const unsigned long b = first_part(view, shard), e = last_part(view, shard);
            CSAMD_REQUIRE(part >= b && part < e && pack->partition_count == e - b,
                          "stage_partition: the pack does not hold shard %lu (%llu partitions, shard has %lu)", shard,
                          pack->partition_count, e - b);
            if (r.view != nullptr && !release_group(s, part)) return 0;
            if (!set_device(device) || !stage_bell_from_pack(s, view, pack, part, part - b, device, stream)) return 0;
        }
    }
    if (drop_host_after_upload && view->parts != nullptr && view->parts[part] != nullptr)
        return csamd_sharded_drop_partition(view, format, part);
    return 1;
}

int csamd_sharded_device_swap_shard(csamd_sharded_device *s, csamd_sharded *view, unsigned int format, const csamd_cspack_file *pack,
                                    unsigned long out_shard, unsigned long in_shard, int device, csamd_stream_t stream,
                                    int drop_host_after_upload, int drop_host_after_release) {
    if (!csamd_sharded_device_stage_shard(s, view, format, pack, in_shard, device, stream, drop_host_after_upload)) return 0;
    if (!csamd_sharded_device_release_shard(s, view, out_shard)) return 0;
    if (drop_host_after_release) return csamd_sharded_drop_shard(view, format, out_shard);
    return 1;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Repack a sharded CSR into a sharded Blocked-ELL of parts of
// target_block_rows row blocks (convert/blocked_ell_from_compressed.cuh:480-600).
// The rows of a part are stitched from the source partitions into one host
// CSR (a copy per partition span, not per row), which the GPU converter turns
// into the part.
// ---------------------------------------------------------------------------

namespace csamd {
namespace {

struct stitched_csr {
    std::vector<unsigned int> ptr, idx;
    std::vector<uint16_t> val;
    csamd_host_compressed h;
};

// Rows [r0, r1) of src (compressed by row) as one CSR.
int stitch_rows(const csamd_sharded *src, unsigned long r0, unsigned long r1, stitched_csr &out) {
    const unsigned long rows = r1 - r0;
    CSAMD_REQUIRE(rows <= 0xffffffffull, "repack: %lu rows in one part", rows);
    out.ptr.assign(rows + 1u, 0u);
    unsigned long long nnz = 0ull;
    // first pass: the row pointers, partition span by partition span
    for (unsigned long p = r0 < r1 ? csamd_sharded_find_partition(src, r0) : src->num_partitions; p < src->num_partitions; ++p) {
        const unsigned long pb = src->partition_offsets[p], pe = pb + src->partition_rows[p];
        if (pb >= r1) break;
        const csamd_host_compressed *part = (const csamd_host_compressed *) src->parts[p];
        CSAMD_REQUIRE(part != nullptr, "repack: partition %lu is not loaded", p);
        CSAMD_REQUIRE(part->axis == CSAMD_COMPRESSED_BY_ROW, "repack: partition %lu is not compressed by row", p);
        const unsigned long a = pb > r0 ? pb : r0, b = pe < r1 ? pe : r1;
        for (unsigned long r = a; r < b; ++r) {
            const unsigned long local = r - pb;
            nnz += part->majorPtr[local + 1u] - part->majorPtr[local];
            CSAMD_REQUIRE(nnz <= 0xffffffffull, "repack: part of more than 2^32 entries");
            out.ptr[r - r0 + 1u] = (unsigned int) nnz;
        }
    }
    out.idx.resize(nnz);
    out.val.resize(nnz);
    for (unsigned long p = r0 < r1 ? csamd_sharded_find_partition(src, r0) : src->num_partitions; p < src->num_partitions; ++p) {
        const unsigned long pb = src->partition_offsets[p], pe = pb + src->partition_rows[p];
        if (pb >= r1) break;
        const csamd_host_compressed *part = (const csamd_host_compressed *) src->parts[p];
        const unsigned long a = pb > r0 ? pb : r0, b = pe < r1 ? pe : r1;
        if (a >= b) continue;
        const unsigned int sb = part->majorPtr[a - pb], se = part->majorPtr[b - pb];
        const unsigned int ob = out.ptr[a - r0];
        if (se > sb) {
            std::memcpy(out.idx.data() + ob, part->minorIdx + sb, (size_t) (se - sb) * 4u);
            std::memcpy(out.val.data() + ob, part->val + sb, (size_t) (se - sb) * 2u);
        }
    }
    std::memset(&out.h, 0, sizeof(out.h));
    out.h.rows = (unsigned int) rows;
    out.h.cols = (unsigned int) src->cols;
    out.h.nnz = (unsigned int) nnz;
    out.h.axis = CSAMD_COMPRESSED_BY_ROW;
    out.h.majorPtr = out.ptr.data();
    out.h.minorIdx = nnz != 0u ? out.idx.data() : nullptr;
    out.h.val = nnz != 0u ? (csamd_half *) out.val.data() : nullptr;
    return 1;
}

} // namespace
} // namespace csamd

extern "C" {

int csamd_repack_sharded_compressed_to_blocked_ell(const csamd_sharded *src, unsigned int block_size, unsigned long target_block_rows,
                                                   csamd_sharded *dst) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && block_size != 0u, "repack_sharded_compressed_to_blocked_ell: bad arguments");
    CSAMD_REQUIRE(src->cols <= 0xffffffffull, "repack_sharded_compressed_to_blocked_ell: %lu columns", src->cols);
    const unsigned long rows_per_part = target_block_rows == 0ul ? src->rows : target_block_rows * (unsigned long) block_size;
    csamd_sharded_clear(dst, CSAMD_FORMAT_BLOCKED_ELL);
    dst->cols = src->cols;
    for (unsigned long r0 = 0ul; r0 < src->rows;) {
        const unsigned long r1 = rows_per_part == 0ul || r0 + rows_per_part >= src->rows ? src->rows : r0 + rows_per_part;
        stitched_csr part_src;
        if (!stitch_rows(src, r0, r1, part_src)) return 0;
        csamd_host_blocked_ell *part = (csamd_host_blocked_ell *) csamd_host_part_alloc(CSAMD_FORMAT_BLOCKED_ELL);
        CSAMD_REQUIRE(part != nullptr, "repack_sharded_compressed_to_blocked_ell: out of host memory");
        if (!csamd_host_blocked_ell_from_compressed(&part_src.h, block_size, part) ||
            !csamd_sharded_append_partition(dst, CSAMD_FORMAT_BLOCKED_ELL, part)) {
            csamd_host_blocked_ell_clear(part);
            std::free(part);
            return 0;
        }
        r0 = r1;
    }
    return 1;
}

int csamd_repack_sharded_compressed_to_blocked_ell_auto(const csamd_sharded *src, const unsigned int *candidates,
                                                        unsigned int candidate_count, unsigned long target_block_rows,
                                                        csamd_sharded *dst, csamd_blocked_ell_tune_result *picked) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && candidates != nullptr && candidate_count != 0u,
                  "repack_sharded_compressed_to_blocked_ell_auto: bad arguments");
    // the block size is chosen over the whole matrix, as the reference does
    stitched_csr whole;
    if (!stitch_rows(src, 0ul, src->rows, whole)) return 0;
    csamd_blocked_ell_tune_result tune;
    std::memset(&tune, 0, sizeof(tune));
    if (!csamd_host_choose_blocked_ell_block_size(&whole.h, candidates, candidate_count, &tune)) return 0;
    if (!csamd_repack_sharded_compressed_to_blocked_ell(src, tune.block_size, target_block_rows, dst)) return 0;
    if (picked != nullptr) *picked = tune;
    return 1;
}

} // extern "C"
