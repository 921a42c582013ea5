// Workspace-owned major-nnz bucketing: the operator layer over the raw plan
// and rebuild (bucket.hip), and the same for one shard of a sharded
// compressed matrix whose partitions are resident on the device.
//
// Reference: src/bucket/operators/major_nnz.cuh:11-213 (workspace, reserve,
// build_plan, rebuild, build_bucketed_major_view) and
// src/bucket/operators/sharded_major_nnz.cuh:35-273 with the shard kernels of
// src/bucket/major_nnz_raw.cuh:122-186,252-320 (count per shard row across
// partitions, gather counts in plan order, reorder segments from the
// partitions).  The shard's rows are numbered across its partitions in
// partition order; a kernel finds a row's partition by binary search over
// the partitions' row offsets, so the shard is bucketed and rebuilt as one
// matrix without first being stitched.
#include "common.hh"

namespace csamd {
namespace {

constexpr unsigned int BO_THREADS = 256u;

// Partition of shard row `row`: the last p with offs[p] <= row (empty
// partitions share their offset with the next one and are skipped).
__device__ __forceinline__ unsigned int bo_part_of(const unsigned int *__restrict__ offs, unsigned int parts, unsigned int row) {
    unsigned int lo = 0u, hi = parts - 1u;
    while (lo < hi) {
        const unsigned int mid = (lo + hi + 1u) >> 1;
        if (offs[mid] <= row) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

struct bo_parts {
    const unsigned int *const *ptr;
    const unsigned int *const *idx;
    const void *const *val;
    const unsigned int *offs;  // [parts + 1] shard row offsets
    unsigned int parts;
};

// Entry range of shard row `row` in its partition's arrays.
__device__ __forceinline__ void bo_row(const bo_parts &P, unsigned int row, unsigned int &p, unsigned int &b, unsigned int &e) {
    p = bo_part_of(P.offs, P.parts, row);
    const unsigned int local = row - P.offs[p];
    const CSAMD_GLOBAL unsigned int *ptr = gmem(P.ptr[p]);
    b = ptr[local];
    e = ptr[local + 1u];
}

// nnz per shard row and the identity order (count_shard_major_nnz + init_major_identity).
__global__ void bo_shard_count_kernel(bo_parts P, unsigned int rows, unsigned int *__restrict__ nnz, unsigned int *__restrict__ ids) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) {
        unsigned int p, b, e;
        bo_row(P, i, p, b, e);
        nnz[i] = e - b;
        ids[i] = i;
    }
}

// counts[i] = nnz of shard row order[i].
__global__ void bo_shard_gather_counts_kernel(bo_parts P, const unsigned int *__restrict__ order, unsigned int rows,
                                              unsigned int *__restrict__ counts) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) {
        unsigned int p, b, e;
        bo_row(P, order[i], p, b, e);
        counts[i] = e - b;
    }
}

// One wave per rebuilt row: the segment of shard row order[d] from its partition.
template<typename S>
__global__ void bo_shard_reorder_kernel(bo_parts P, const unsigned int *__restrict__ order, unsigned int rows,
                                        const unsigned int *__restrict__ dst_ptr, unsigned int *__restrict__ dst_idx,
                                        S *__restrict__ dst_val) {
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (unsigned int d = blockIdx.x * nw + wave; d < rows; d += gridDim.x * nw) {
        unsigned int p, b, e;
        bo_row(P, order[d], p, b, e);
        const CSAMD_GLOBAL unsigned int *idx = gmem(P.idx[p]);
        const CSAMD_GLOBAL S *val = gmem((const S *) P.val[p]);
        const unsigned int o = dst_ptr[d];
        for (unsigned int k = lane_id(); k < e - b; k += 64u) {
            dst_idx[o + k] = idx[b + k];
            dst_val[o + k] = val[b + k];
        }
    }
}

__global__ void bo_equal_offsets_kernel(unsigned int n, unsigned int buckets, unsigned int *__restrict__ offsets) {
    const unsigned int size = n / buckets, rem = n % buckets;
    for (unsigned int b = blockIdx.x * blockDim.x + threadIdx.x; b <= buckets; b += gridDim.x * blockDim.x)
        offsets[b] = b == buckets ? n : b * size + (b < rem ? b : rem);
}

__global__ void bo_inverse_kernel(const unsigned int *__restrict__ order, unsigned int n, unsigned int *__restrict__ inv) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) inv[order[i]] = i;
}

unsigned int bo_grid(unsigned int n) {
    return grid_for(n, BO_THREADS, (unsigned int) device_cu_count() * 32u);
}

// Frees a device block after the stream's earlier work (which may still read
// it) has finished.
void bo_free(csamd_compressed_major_bucket_workspace *ws, void *&p, bool &synced) {
    if (p == nullptr) return;
    if (!synced && ws->stream != nullptr) (void) hipStreamSynchronize((hipStream_t) ws->stream);
    synced = true;
    (void) hipFree(p);
    p = nullptr;
}

int bo_reserve(csamd_compressed_major_bucket_workspace *ws, unsigned int major_dim, unsigned int nnz, unsigned int bucket_count,
               unsigned int value_bytes) {
    CSAMD_REQUIRE(ws != nullptr, "major bucket reserve: null workspace");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    size_t sort_bytes = 0u, scan_bytes = 0u;
    if (!csamd_major_nnz_bucket_sort_scratch_bytes(major_dim, &sort_bytes)) return 0;
    if (!csamd_major_nnz_bucket_scan_scratch_bytes(major_dim, &scan_bytes)) return 0;
    bool synced = false;
    if (major_dim > ws->major_capacity || ws->d_major_block == nullptr) {
        bo_free(ws, ws->d_major_block, synced);
        const size_t a = align_up((size_t) major_dim * 4u, 256u), total = 5u * a + align_up(((size_t) major_dim + 1u) * 4u, 256u);
        CSAMD_HIP(hipMalloc(&ws->d_major_block, total), "hipMalloc major bucket block");
        char *b = (char *) ws->d_major_block;
        ws->d_major_nnz = (unsigned int *) b;
        ws->d_major_nnz_sorted = (unsigned int *) (b + a);
        ws->d_major_order_in = (unsigned int *) (b + 2u * a);
        ws->d_major_order_out = (unsigned int *) (b + 3u * a);
        ws->d_inverse_major_order = (unsigned int *) (b + 4u * a);
        ws->d_rebuilt_major_ptr = (unsigned int *) (b + 5u * a);
        ws->major_capacity = major_dim;
    }
    if (nnz > ws->nnz_capacity || value_bytes > ws->value_bytes_capacity) {
        const unsigned int n = nnz > ws->nnz_capacity ? nnz : ws->nnz_capacity;
        const unsigned int vb = value_bytes > ws->value_bytes_capacity ? value_bytes : ws->value_bytes_capacity;
        bo_free(ws, ws->d_rebuilt_block, synced);
        ws->d_rebuilt_minor_idx = nullptr;
        ws->d_rebuilt_val = nullptr;
        ws->nnz_capacity = 0u;
        if (n != 0u) {
            const size_t off = align_up((size_t) n * 4u, 256u);
            CSAMD_HIP(hipMalloc(&ws->d_rebuilt_block, off + (size_t) n * vb), "hipMalloc major bucket rebuilt block");
            ws->d_rebuilt_minor_idx = (unsigned int *) ws->d_rebuilt_block;
            ws->d_rebuilt_val = (char *) ws->d_rebuilt_block + off;
        }
        ws->nnz_capacity = n;
        ws->value_bytes_capacity = vb;
    }
    if (bucket_count > ws->bucket_capacity || ws->d_bucket_offsets == nullptr) {
        void *p = ws->d_bucket_offsets;
        bo_free(ws, p, synced);
        ws->d_bucket_offsets = nullptr;
        const unsigned int bc = bucket_count > ws->bucket_capacity ? bucket_count : ws->bucket_capacity;
        CSAMD_HIP(hipMalloc((void **) &ws->d_bucket_offsets, ((size_t) bc + 1u) * 4u), "hipMalloc major bucket offsets");
        ws->bucket_capacity = bc;
    }
    if (sort_bytes > ws->sort_capacity) {
        bo_free(ws, ws->d_sort_tmp, synced);
        CSAMD_HIP(hipMalloc(&ws->d_sort_tmp, sort_bytes), "hipMalloc major bucket sort scratch");
        ws->sort_capacity = sort_bytes;
    }
    if (scan_bytes > ws->scan_capacity) {
        bo_free(ws, ws->d_scan_tmp, synced);
        CSAMD_HIP(hipMalloc(&ws->d_scan_tmp, scan_bytes), "hipMalloc major bucket scan scratch");
        ws->scan_capacity = scan_bytes;
    }
    return 1;
}

unsigned int value_bytes_of(unsigned int value_type) {
    const int vb = value_type_bytes(value_type);
    return vb == 1 || vb == 2 || vb == 4 ? (unsigned int) vb : 0u;
}

int bo_stream_setup(int device, csamd_stream_t stream, csamd_stream_t *out, int *owns) {
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice major bucket setup");
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags major bucket");
        *out = (csamd_stream_t) s;
        *owns = 1;
    } else {
        *out = stream;
        *owns = 0;
    }
    return 1;
}

template<typename S>
int bo_launch_reorder(const bo_parts &P, const unsigned int *order, unsigned int rows, const unsigned int *dst_ptr,
                      unsigned int *dst_idx, void *dst_val, hipStream_t s) {
    bo_shard_reorder_kernel<S><<<grid_for((rows + 3u) / 4u, 1u, (unsigned int) device_cu_count() * 32u), BO_THREADS, 0, s>>>(
        P, order, rows, dst_ptr, dst_idx, (S *) dst_val);
    return hip_ok(hipGetLastError(), "bo_shard_reorder_kernel");
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_major_bucket_init(csamd_compressed_major_bucket_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_major_bucket_clear(csamd_compressed_major_bucket_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    if (ws->stream != nullptr) (void) hipStreamSynchronize((hipStream_t) ws->stream);
    void *blocks[] = {ws->d_scan_tmp, ws->d_sort_tmp, ws->d_rebuilt_block, ws->d_bucket_offsets, ws->d_major_block};
    for (void *b : blocks)
        if (b != nullptr) (void) hipFree(b);
    if (ws->owns_stream != 0 && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    csamd_major_bucket_init(ws);
}

int csamd_major_bucket_setup(csamd_compressed_major_bucket_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "major bucket setup: null workspace");
    csamd_major_bucket_clear(ws);
    ws->device = device;
    return bo_stream_setup(device, stream, &ws->stream, &ws->owns_stream);
}

int csamd_major_bucket_reserve(csamd_compressed_major_bucket_workspace *ws, unsigned int major_dim, unsigned int nnz,
                               unsigned int bucket_count) {
    return bo_reserve(ws, major_dim, nnz, bucket_count, 2u);
}

int csamd_major_bucket_reserve_typed(csamd_compressed_major_bucket_workspace *ws, unsigned int major_dim, unsigned int nnz,
                                     unsigned int bucket_count, unsigned int value_type) {
    const unsigned int vb = value_bytes_of(value_type);
    CSAMD_REQUIRE(vb != 0u, "major bucket reserve: unsupported value type %u", value_type);
    return bo_reserve(ws, major_dim, nnz, bucket_count, vb);
}

int csamd_major_bucket_build_plan(const csamd_compressed_view *src, unsigned int requested_bucket_count,
                                  csamd_compressed_major_bucket_workspace *ws, csamd_major_nnz_bucket_plan_view *out) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr && out != nullptr, "build_plan: null argument");
    const unsigned int major_dim = src->axis == CSAMD_COMPRESSED_BY_COL ? src->cols : src->rows;
    const unsigned int bucket_count = csamd_clamp_bucket_count(major_dim, requested_bucket_count);
    if (!bo_reserve(ws, major_dim, src->nnz, bucket_count, ws->value_bytes_capacity != 0u ? ws->value_bytes_capacity : 2u)) return 0;
    if (!csamd_build_major_nnz_bucket_plan_raw(src->majorPtr, major_dim, ws->d_major_nnz, ws->d_major_nnz_sorted, ws->d_major_order_in,
                                               ws->d_major_order_out, ws->d_bucket_offsets, bucket_count, ws->d_sort_tmp,
                                               ws->sort_capacity, ws->stream))
        return 0;
    out->major_dim = major_dim;
    out->bucket_count = bucket_count;
    out->major_order = ws->d_major_order_out;
    out->major_nnz_sorted = ws->d_major_nnz_sorted;
    out->bucket_offsets = ws->d_bucket_offsets;
    out->inverse_major_order = nullptr;
    return 1;
}

int csamd_major_bucket_rebuild_typed(const csamd_compressed_view *src, unsigned int value_type,
                                     const csamd_major_nnz_bucket_plan_view *plan, csamd_compressed_major_bucket_workspace *ws,
                                     csamd_compressed_major_bucket_result *out) {
    CSAMD_REQUIRE(src != nullptr && plan != nullptr && ws != nullptr && out != nullptr, "rebuild: null argument");
    const unsigned int vb = value_bytes_of(value_type);
    CSAMD_REQUIRE(vb != 0u, "rebuild: unsupported value type %u", value_type);
    const unsigned int major_dim = src->axis == CSAMD_COMPRESSED_BY_COL ? src->cols : src->rows;
    CSAMD_REQUIRE(plan->major_dim == major_dim, "rebuild: plan of %u majors for a source of %u", plan->major_dim, major_dim);
    if (!bo_reserve(ws, major_dim, src->nnz, ws->bucket_capacity, vb)) return 0;
    // an empty source has no payload arrays: any non-null pointer does (no entry is read or written)
    void *any = ws->d_major_nnz;
    if (!csamd_rebuild_compressed_major_order_raw(src->majorPtr, src->minorIdx != nullptr ? src->minorIdx : (const unsigned int *) any,
                                                  src->val != nullptr ? (const void *) src->val : any, value_type, major_dim,
                                                  plan->major_order, ws->d_rebuilt_major_ptr,
                                                  ws->d_rebuilt_minor_idx != nullptr ? ws->d_rebuilt_minor_idx : (unsigned int *) any,
                                                  ws->d_rebuilt_val != nullptr ? ws->d_rebuilt_val : any, ws->d_inverse_major_order,
                                                  ws->d_scan_tmp, ws->scan_capacity, ws->stream))
        return 0;
    out->rebuilt.rows = src->rows;
    out->rebuilt.cols = src->cols;
    out->rebuilt.nnz = src->nnz;
    out->rebuilt.axis = src->axis;
    out->rebuilt.majorPtr = ws->d_rebuilt_major_ptr;
    out->rebuilt.minorIdx = ws->d_rebuilt_minor_idx;
    out->rebuilt.val = (csamd_half *) ws->d_rebuilt_val;
    out->plan = *plan;
    out->plan.inverse_major_order = ws->d_inverse_major_order;
    return 1;
}

int csamd_major_bucket_rebuild(const csamd_compressed_view *src, const csamd_major_nnz_bucket_plan_view *plan,
                               csamd_compressed_major_bucket_workspace *ws, csamd_compressed_major_bucket_result *out) {
    return csamd_major_bucket_rebuild_typed(src, CSAMD_VALUE_F16, plan, ws, out);
}

int csamd_build_bucketed_major_view_typed(const csamd_compressed_view *src, unsigned int value_type,
                                          unsigned int requested_bucket_count, csamd_compressed_major_bucket_workspace *ws,
                                          csamd_compressed_major_bucket_result *out) {
    csamd_major_nnz_bucket_plan_view plan;
    if (!csamd_major_bucket_build_plan(src, requested_bucket_count, ws, &plan)) return 0;
    return csamd_major_bucket_rebuild_typed(src, value_type, &plan, ws, out);
}

int csamd_build_bucketed_major_view(const csamd_compressed_view *src, unsigned int requested_bucket_count,
                                    csamd_compressed_major_bucket_workspace *ws, csamd_compressed_major_bucket_result *out) {
    return csamd_build_bucketed_major_view_typed(src, CSAMD_VALUE_F16, requested_bucket_count, ws, out);
}

// ---------------------------------------------------------------------------
// Shard operators.
// ---------------------------------------------------------------------------
void csamd_sharded_major_bucket_init(csamd_sharded_major_bucket_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    csamd_major_bucket_init(&ws->base);
}

void csamd_sharded_major_bucket_clear(csamd_sharded_major_bucket_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->base.device >= 0) (void) hipSetDevice(ws->base.device);
    if (ws->base.stream != nullptr) (void) hipStreamSynchronize((hipStream_t) ws->base.stream);
    if (ws->tables_event != nullptr) (void) hipEventDestroy((hipEvent_t) ws->tables_event);
    if (ws->h_part_block != nullptr) (void) hipHostFree(ws->h_part_block);
    if (ws->d_part_block != nullptr) (void) hipFree(ws->d_part_block);
    csamd_major_bucket_clear(&ws->base);
    csamd_sharded_major_bucket_init(ws);
}

int csamd_sharded_major_bucket_setup(csamd_sharded_major_bucket_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "shard bucket setup: null workspace");
    csamd_sharded_major_bucket_clear(ws);
    if (!csamd_major_bucket_setup(&ws->base, device, stream)) return 0;
    hipEvent_t e = nullptr;
    CSAMD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate shard bucket tables");
    ws->tables_event = (void *) e;
    return 1;
}

int csamd_sharded_major_bucket_reserve_parts(csamd_sharded_major_bucket_workspace *ws, unsigned long part_count) {
    CSAMD_REQUIRE(ws != nullptr, "shard bucket reserve_parts: null workspace");
    if (!ensure_device(ws->base.device >= 0 ? ws->base.device : 0)) return 0;
    if (part_count <= ws->part_capacity && (part_count == 0ul || ws->h_part_block != nullptr)) return 1;
    // earlier copies and kernels may still read the old tables
    if (ws->base.stream != nullptr) CSAMD_HIP(hipStreamSynchronize((hipStream_t) ws->base.stream), "shard bucket tables sync");
    if (ws->h_part_block != nullptr) (void) hipHostFree(ws->h_part_block);
    if (ws->d_part_block != nullptr) (void) hipFree(ws->d_part_block);
    ws->h_part_block = ws->d_part_block = nullptr;
    ws->part_capacity = 0ul;
    if (part_count != 0ul) {
        const size_t pb = align_up((size_t) part_count * sizeof(void *), 64u);
        const size_t total = 3u * pb + ((size_t) part_count + 1u) * 4u;
        CSAMD_HIP(hipHostMalloc(&ws->h_part_block, total, hipHostMallocDefault), "hipHostMalloc shard bucket tables");
        std::memset(ws->h_part_block, 0, total);
        CSAMD_HIP(hipMalloc(&ws->d_part_block, total), "hipMalloc shard bucket tables");
        char *h = (char *) ws->h_part_block, *d = (char *) ws->d_part_block;
        ws->h_part_major_ptr = (const unsigned int **) h;
        ws->h_part_minor_idx = (const unsigned int **) (h + pb);
        ws->h_part_val = (const void **) (h + 2u * pb);
        ws->h_part_row_offsets = (unsigned int *) (h + 3u * pb);
        ws->d_part_major_ptr = (const unsigned int **) d;
        ws->d_part_minor_idx = (const unsigned int **) (d + pb);
        ws->d_part_val = (const void **) (d + 2u * pb);
        ws->d_part_row_offsets = (unsigned int *) (d + 3u * pb);
    }
    ws->part_capacity = part_count;
    return 1;
}

int csamd_build_bucketed_shard_major_view_typed(const csamd_sharded_device *state, const csamd_sharded *view,
                                                unsigned long shard_id, unsigned int value_type,
                                                unsigned int requested_bucket_count, csamd_sharded_major_bucket_workspace *ws,
                                                csamd_sharded_major_bucket_result *out) {
    CSAMD_REQUIRE(state != nullptr && view != nullptr && ws != nullptr && out != nullptr, "build_bucketed_shard_major_view: null argument");
    CSAMD_REQUIRE(shard_id < view->num_shards, "build_bucketed_shard_major_view: shard %lu of %lu", shard_id, view->num_shards);
    const unsigned int vb = value_bytes_of(value_type);
    CSAMD_REQUIRE(vb != 0u, "build_bucketed_shard_major_view: unsupported value type %u", value_type);
    const unsigned long pb = csamd_sharded_first_partition_in_shard(view, shard_id);
    const unsigned long pe = csamd_sharded_last_partition_in_shard(view, shard_id);
    CSAMD_REQUIRE(pb <= pe && pe <= view->num_partitions, "build_bucketed_shard_major_view: bad partition range");
    const unsigned long rows_l = view->shard_offsets[shard_id + 1u] - view->shard_offsets[shard_id];
    unsigned long long nnz_l = 0ull, rows_sum = 0ull;
    for (unsigned long p = pb; p < pe; ++p) {
        nnz_l += view->partition_nnz[p];
        rows_sum += view->partition_rows[p];
    }
    CSAMD_REQUIRE(rows_l <= 0xffffffffull && nnz_l <= 0xffffffffull && rows_sum == rows_l,
                  "build_bucketed_shard_major_view: shard of %lu rows and %llu entries (partition rows %llu)", rows_l, nnz_l, rows_sum);
    const unsigned int rows = (unsigned int) rows_l, nnz = (unsigned int) nnz_l, parts = (unsigned int) (pe - pb);
    const unsigned int bucket_count = csamd_clamp_bucket_count(rows, requested_bucket_count);
    if (!csamd_sharded_major_bucket_reserve_parts(ws, parts)) return 0;
    if (!bo_reserve(&ws->base, rows, nnz, bucket_count, vb)) return 0;
    hipStream_t s = (hipStream_t) ws->base.stream;
    // the pinned tables are rewritten only after the previous call's copies read them
    if (ws->tables_event != nullptr) CSAMD_HIP(hipEventSynchronize((hipEvent_t) ws->tables_event), "shard bucket tables wait");
    if (parts != 0u) {
        ws->h_part_row_offsets[0] = 0u;
        for (unsigned int l = 0; l < parts; ++l) {
            const unsigned long p = pb + l;
            CSAMD_REQUIRE(p < state->capacity, "build_bucketed_shard_major_view: partition %lu has no record", p);
            const csamd_partition_record &r = state->parts[p];
            CSAMD_REQUIRE(r.view != nullptr && r.a0 != nullptr, "build_bucketed_shard_major_view: partition %lu is not resident", p);
            CSAMD_REQUIRE(view->partition_aux == nullptr || view->partition_aux[p] == CSAMD_COMPRESSED_BY_ROW,
                          "build_bucketed_shard_major_view: partition %lu is not compressed by row", p);
            CSAMD_REQUIRE(r.device_id == ws->base.device, "build_bucketed_shard_major_view: partition %lu on device %d, workspace on %d",
                          p, r.device_id, ws->base.device);
            CSAMD_REQUIRE(view->partition_nnz[p] == 0ul || (r.a1 != nullptr && r.a2 != nullptr),
                          "build_bucketed_shard_major_view: partition %lu has no entry arrays", p);
            ws->h_part_major_ptr[l] = (const unsigned int *) r.a0;
            ws->h_part_minor_idx[l] = (const unsigned int *) r.a1;
            ws->h_part_val[l] = r.a2;
            ws->h_part_row_offsets[l + 1u] = ws->h_part_row_offsets[l] + (unsigned int) view->partition_rows[p];
        }
        const size_t tb = (char *) ws->h_part_row_offsets - (char *) ws->h_part_block + ((size_t) parts + 1u) * 4u;
        CSAMD_HIP(hipMemcpyAsync(ws->d_part_block, ws->h_part_block, tb, hipMemcpyHostToDevice, s), "H2D shard bucket tables");
        if (ws->tables_event != nullptr) CSAMD_HIP(hipEventRecord((hipEvent_t) ws->tables_event, s), "shard bucket tables event");
    }
    csamd_compressed_major_bucket_workspace &B = ws->base;
    if (rows != 0u) {
        const bo_parts P{ws->d_part_major_ptr, ws->d_part_minor_idx, ws->d_part_val, ws->d_part_row_offsets, parts};
        bo_shard_count_kernel<<<bo_grid(rows), BO_THREADS, 0, s>>>(P, rows, B.d_major_nnz, B.d_major_order_in);
        CSAMD_HIP(hipGetLastError(), "bo_shard_count_kernel");
        if (!radix_sort_pairs_u32(B.d_major_nnz, B.d_major_nnz_sorted, B.d_major_order_in, B.d_major_order_out, rows, 32u, B.d_sort_tmp,
                                  B.sort_capacity, s))
            return 0;
        bo_equal_offsets_kernel<<<grid_for(bucket_count + 1u, BO_THREADS, 256u), BO_THREADS, 0, s>>>(rows, bucket_count,
                                                                                                      B.d_bucket_offsets);
        CSAMD_HIP(hipGetLastError(), "bo_equal_offsets_kernel");
        unsigned int *counts = (unsigned int *) B.d_scan_tmp;
        void *scan = (char *) B.d_scan_tmp + align_up(((size_t) rows + 1u) * 4u, 256u);
        bo_shard_gather_counts_kernel<<<bo_grid(rows), BO_THREADS, 0, s>>>(P, B.d_major_order_out, rows, counts);
        CSAMD_HIP(hipGetLastError(), "bo_shard_gather_counts_kernel");
        if (!exclusive_scan_u32(counts, rows, B.d_rebuilt_major_ptr, nullptr, scan, exclusive_scan_scratch_bytes(rows + 1u), s)) return 0;
        if (nnz != 0u) {
            int ok = 0;
            switch (vb) {
                case 1u: ok = bo_launch_reorder<uint8_t>(P, B.d_major_order_out, rows, B.d_rebuilt_major_ptr, B.d_rebuilt_minor_idx,
                                                         B.d_rebuilt_val, s); break;
                case 2u: ok = bo_launch_reorder<uint16_t>(P, B.d_major_order_out, rows, B.d_rebuilt_major_ptr, B.d_rebuilt_minor_idx,
                                                          B.d_rebuilt_val, s); break;
                default: ok = bo_launch_reorder<float>(P, B.d_major_order_out, rows, B.d_rebuilt_major_ptr, B.d_rebuilt_minor_idx,
                                                       B.d_rebuilt_val, s); break;
            }
            if (!ok) return 0;
        }
        bo_inverse_kernel<<<bo_grid(rows), BO_THREADS, 0, s>>>(B.d_major_order_out, rows, B.d_inverse_major_order);
        CSAMD_HIP(hipGetLastError(), "bo_inverse_kernel");
    } else {
        if (!zero_bytes(B.d_bucket_offsets, 8u, s)) return 0;  // [0, 0]: one empty bucket
        if (!zero_bytes(B.d_rebuilt_major_ptr, 4u, s)) return 0;
    }
    std::memset(out, 0, sizeof(*out));
    out->rebuilt.rows = rows;
    out->rebuilt.cols = (unsigned int) view->cols;
    out->rebuilt.nnz = nnz;
    out->rebuilt.axis = CSAMD_COMPRESSED_BY_ROW;
    out->rebuilt.majorPtr = B.d_rebuilt_major_ptr;
    out->rebuilt.minorIdx = B.d_rebuilt_minor_idx;
    out->rebuilt.val = (csamd_half *) B.d_rebuilt_val;
    out->plan.major_dim = rows;
    out->plan.bucket_count = bucket_count;
    out->plan.major_order = B.d_major_order_out;
    out->plan.major_nnz_sorted = B.d_major_nnz_sorted;
    out->plan.bucket_offsets = B.d_bucket_offsets;
    out->plan.inverse_major_order = B.d_inverse_major_order;
    out->shard_id = shard_id;
    return 1;
}

int csamd_build_bucketed_shard_major_view(const csamd_sharded_device *state, const csamd_sharded *view, unsigned long shard_id,
                                          unsigned int requested_bucket_count, csamd_sharded_major_bucket_workspace *ws,
                                          csamd_sharded_major_bucket_result *out) {
    return csamd_build_bucketed_shard_major_view_typed(state, view, shard_id, CSAMD_VALUE_F16, requested_bucket_count, ws, out);
}

} // extern "C"
