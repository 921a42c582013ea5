// Major-axis nnz bucketing: per-major nnz, stable ascending order (ties by
// index), equal-count bucket offsets (reference rule) and log2 bucket offsets
// (project rule), segment rebuild in bucket order, and the Blocked-ELL
// row-block bucket plan.
//
// Reference: src/bucket/major_nnz_raw.cuh:27-320 and
// src/bucket/kernels/csBucket/*.cuh (count, identity, CUB stable radix sort
// of u32 nnz keys, fill_equal_count_bucket_offsets, gather / scan / reorder /
// inverse), src/bucket/operators/blocked_ell_bucket.cuh:77-210 (row-block keys).
// The reference's rebuild scans shifted counts with an exclusive sum
// (major_nnz_raw.cuh:214-226), which yields majorPtr one row late; the rebuilt
// CSR here has the documented meaning: dst row i = src row order[i].
#include "common.hh"

namespace csamd {
namespace {

__global__ void count_major_nnz_kernel(const unsigned int *__restrict__ ptr, unsigned int n, unsigned int *__restrict__ nnz,
                                       unsigned int *__restrict__ ids) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        nnz[i] = ptr[i + 1u] - ptr[i];
        ids[i] = i;
    }
}

__global__ void equal_count_offsets_kernel(unsigned int n, unsigned int buckets, unsigned int *__restrict__ offsets) {
    const unsigned int size = n / buckets, rem = n % buckets;
    for (unsigned int b = blockIdx.x * blockDim.x + threadIdx.x; b <= buckets; b += gridDim.x * blockDim.x) {
        offsets[b] = b == buckets ? n : b * size + (b < rem ? b : rem);
    }
}

// offsets[b] = first sorted position whose nnz >= 2^(b + min_log2 - 1), b >= 1.
__global__ void log2_offsets_kernel(const unsigned int *__restrict__ sorted_nnz, unsigned int n, unsigned int buckets,
                                    unsigned int min_log2, unsigned int *__restrict__ offsets) {
    for (unsigned int b = blockIdx.x * blockDim.x + threadIdx.x; b <= buckets; b += gridDim.x * blockDim.x) {
        if (b == 0u) {
            offsets[0] = 0u;
            continue;
        }
        if (b == buckets) {
            offsets[b] = n;
            continue;
        }
        const unsigned int e = b + min_log2 - 1u;
        const unsigned long long thr = e >= 32u ? (1ull << 32) : (1ull << e);
        unsigned int lo = 0u, hi = n;
        while (lo < hi) {
            const unsigned int mid = lo + ((hi - lo) >> 1);
            if ((unsigned long long) sorted_nnz[mid] < thr) lo = mid + 1u;
            else hi = mid;
        }
        offsets[b] = lo;
    }
}

__global__ void gather_counts_kernel(const unsigned int *__restrict__ order, const unsigned int *__restrict__ src_ptr, unsigned int n,
                                     unsigned int *__restrict__ counts) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned int s = order[i];
        counts[i] = src_ptr[s + 1u] - src_ptr[s];
    }
}

// One wave per destination segment; coalesced copy.
template<typename S>
__global__ void reorder_segments_kernel(const unsigned int *__restrict__ order, const unsigned int *__restrict__ src_ptr,
                                        const unsigned int *__restrict__ src_idx, const S *__restrict__ src_val, unsigned int n,
                                        const unsigned int *__restrict__ dst_ptr, unsigned int *__restrict__ dst_idx,
                                        S *__restrict__ dst_val) {
    constexpr unsigned int U = 8u;  // entries per lane in flight: all loads of a step before its stores
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = lane_id();
    const CSAMD_GLOBAL unsigned int *si = gmem(src_idx);
    const CSAMD_GLOBAL S *sv = gmem(src_val);
    CSAMD_GLOBAL unsigned int *di = gmem(dst_idx);
    CSAMD_GLOBAL S *dv = gmem(dst_val);
    for (unsigned int d = blockIdx.x * nw + wave; d < n; d += gridDim.x * nw) {
        const unsigned int s = order[d];
        const unsigned int b = src_ptr[s], len = src_ptr[s + 1u] - b;
        const unsigned int o = dst_ptr[d];
        for (unsigned int k0 = 0u; k0 < len; k0 += 64u * U) {
            unsigned int ix[U];
            S vv[U];
#pragma unroll
            for (unsigned int u = 0; u < U; ++u) {
                const unsigned int k = min(k0 + 64u * u + lane, len - 1u);  // clamped: every lane loads
                ix[u] = si[b + k];
                vv[u] = sv[b + k];
            }
#pragma unroll
            for (unsigned int u = 0; u < U; ++u) {
                const unsigned int k = k0 + 64u * u + lane;
                if (k < len) {
                    di[o + k] = ix[u];
                    dv[o + k] = vv[u];
                }
            }
        }
    }
}

__global__ void inverse_order_kernel(const unsigned int *__restrict__ order, unsigned int n, unsigned int *__restrict__ inv) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) inv[order[i]] = i;
}

// Row-block keys: (valid slots << 32) | hash of valid block columns in slot order.
__global__ void bell_bucket_keys_kernel(unsigned int row_blocks, unsigned int width, const unsigned int *__restrict__ bci,
                                        unsigned long long *__restrict__ keys, unsigned int *__restrict__ ids) {
    for (unsigned int rb = blockIdx.x * blockDim.x + threadIdx.x; rb < row_blocks; rb += gridDim.x * blockDim.x) {
        const unsigned long long base = (unsigned long long) rb * width;
        unsigned int valid = 0u, h = 2166136261u;
        for (unsigned int s = 0; s < width; ++s) {
            const unsigned int c = bci[base + s];
            if (c == CSAMD_INVALID_COL) continue;
            ++valid;
            h ^= c + 0x9e3779b9u + (h << 6) + (h >> 2);
            h *= 16777619u;
        }
        keys[rb] = ((unsigned long long) valid << 32) | (unsigned long long) h;
        ids[rb] = rb;
    }
}

__global__ void key_high_kernel(const unsigned long long *__restrict__ keys, unsigned int n, unsigned int *__restrict__ out) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = (unsigned int) (keys[i] >> 32);
}

template<typename S>
int launch_reorder(const unsigned int *order, const unsigned int *src_ptr, const unsigned int *src_idx, const void *src_val,
                   unsigned int n, const unsigned int *dst_ptr, unsigned int *dst_idx, void *dst_val, hipStream_t s) {
    reorder_segments_kernel<S><<<grid_for((n + 3u) / 4u, 1u, (unsigned int) device_cu_count() * 32u), 256, 0, s>>>(
        order, src_ptr, src_idx, (const S *) src_val, n, dst_ptr, dst_idx, (S *) dst_val);
    return hip_ok(hipGetLastError(), "reorder_segments_kernel");
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

unsigned int csamd_clamp_bucket_count(unsigned int major_dim, unsigned int requested) {
    if (major_dim == 0u) return 1u;
    if (requested < 1u) return 1u;
    if (requested > major_dim) return major_dim;
    return requested;
}

int csamd_major_nnz_bucket_sort_scratch_bytes(unsigned int major_dim, size_t *out_bytes) {
    CSAMD_REQUIRE(out_bytes != nullptr, "sort scratch bytes: null out");
    *out_bytes = radix_sort_pairs_u32_scratch_bytes(major_dim);
    return 1;
}

int csamd_major_nnz_bucket_scan_scratch_bytes(unsigned int major_dim, size_t *out_bytes) {
    CSAMD_REQUIRE(out_bytes != nullptr, "scan scratch bytes: null out");
    *out_bytes = exclusive_scan_scratch_bytes(major_dim + 1u) + align_up((size_t) (major_dim + 1u) * 4u, 256u);
    return 1;
}

int csamd_build_major_nnz_bucket_plan_raw(const unsigned int *d_major_ptr, unsigned int major_dim, unsigned int *d_major_nnz,
                                          unsigned int *d_major_nnz_sorted, unsigned int *d_major_order_in,
                                          unsigned int *d_major_order_out, unsigned int *d_bucket_offsets,
                                          unsigned int requested_bucket_count, void *d_sort_tmp, size_t sort_tmp_bytes,
                                          csamd_stream_t stream) {
    hipStream_t s = (hipStream_t) stream;
    CSAMD_REQUIRE(d_major_ptr != nullptr && d_major_nnz != nullptr && d_major_nnz_sorted != nullptr && d_major_order_in != nullptr &&
                      d_major_order_out != nullptr,
                  "build_major_nnz_bucket_plan_raw: null argument");
    if (major_dim == 0u) {
        if (d_bucket_offsets != nullptr) return zero_bytes(d_bucket_offsets, 4u, s);
        return 1;
    }
    count_major_nnz_kernel<<<grid_for(major_dim, 256u, 8192u), 256, 0, s>>>(d_major_ptr, major_dim, d_major_nnz, d_major_order_in);
    CSAMD_HIP(hipGetLastError(), "count_major_nnz_kernel");
    if (!radix_sort_pairs_u32(d_major_nnz, d_major_nnz_sorted, d_major_order_in, d_major_order_out, major_dim, 32u, d_sort_tmp,
                              sort_tmp_bytes, s))
        return 0;
    if (d_bucket_offsets != nullptr) {
        const unsigned int b = csamd_clamp_bucket_count(major_dim, requested_bucket_count);
        equal_count_offsets_kernel<<<grid_for(b + 1u, 256u, 256u), 256, 0, s>>>(major_dim, b, d_bucket_offsets);
        CSAMD_HIP(hipGetLastError(), "equal_count_offsets_kernel");
    }
    return 1;
}

int csamd_fill_log2_bucket_offsets_raw(const unsigned int *d_major_nnz_sorted, unsigned int major_dim, unsigned int bucket_count,
                                       unsigned int min_log2, unsigned int *d_bucket_offsets, csamd_stream_t stream) {
    CSAMD_REQUIRE(d_bucket_offsets != nullptr && (major_dim == 0u || d_major_nnz_sorted != nullptr),
                  "fill_log2_bucket_offsets_raw: null argument");
    CSAMD_REQUIRE(bucket_count >= 1u && bucket_count <= 64u, "fill_log2_bucket_offsets_raw: bucket_count %u out of range", bucket_count);
    CSAMD_REQUIRE(min_log2 >= 1u && min_log2 <= 32u, "fill_log2_bucket_offsets_raw: min_log2 %u out of range", min_log2);
    hipStream_t s = (hipStream_t) stream;
    log2_offsets_kernel<<<1, 256, 0, s>>>(d_major_nnz_sorted, major_dim, bucket_count, min_log2, d_bucket_offsets);
    return hip_ok(hipGetLastError(), "log2_offsets_kernel");
}

int csamd_rebuild_compressed_major_order_raw(const unsigned int *d_src_major_ptr, const unsigned int *d_src_minor_idx,
                                             const void *d_src_val, unsigned int value_type, unsigned int major_dim,
                                             const unsigned int *d_major_order, unsigned int *d_dst_major_ptr,
                                             unsigned int *d_dst_minor_idx, void *d_dst_val, unsigned int *d_inverse_major_order,
                                             void *d_scan_tmp, size_t scan_tmp_bytes, csamd_stream_t stream) {
    hipStream_t s = (hipStream_t) stream;
    CSAMD_REQUIRE(d_src_major_ptr != nullptr && d_major_order != nullptr && d_dst_major_ptr != nullptr,
                  "rebuild_compressed_major_order_raw: null argument");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "rebuild: unsupported value type %u", value_type);
    if (major_dim == 0u) return 1;
    size_t need = 0u;
    (void) csamd_major_nnz_bucket_scan_scratch_bytes(major_dim, &need);
    CSAMD_REQUIRE(d_scan_tmp != nullptr && scan_tmp_bytes >= need, "rebuild: scan scratch %zu < %zu", scan_tmp_bytes, need);
    unsigned int *counts = (unsigned int *) d_scan_tmp;
    void *scan = (char *) d_scan_tmp + align_up((size_t) (major_dim + 1u) * 4u, 256u);
    gather_counts_kernel<<<grid_for(major_dim, 256u, 8192u), 256, 0, s>>>(d_major_order, d_src_major_ptr, major_dim, counts);
    CSAMD_HIP(hipGetLastError(), "gather_counts_kernel");
    if (!exclusive_scan_u32(counts, major_dim, d_dst_major_ptr, nullptr, scan, exclusive_scan_scratch_bytes(major_dim + 1u), s))
        return 0;
    CSAMD_REQUIRE(d_src_minor_idx != nullptr && d_src_val != nullptr && d_dst_minor_idx != nullptr && d_dst_val != nullptr,
                  "rebuild: null payload arrays");
    int ok = 0;
    switch (vb) {
        case 1: ok = launch_reorder<uint8_t>(d_major_order, d_src_major_ptr, d_src_minor_idx, d_src_val, major_dim, d_dst_major_ptr,
                                             d_dst_minor_idx, d_dst_val, s); break;
        case 2: ok = launch_reorder<uint16_t>(d_major_order, d_src_major_ptr, d_src_minor_idx, d_src_val, major_dim, d_dst_major_ptr,
                                              d_dst_minor_idx, d_dst_val, s); break;
        default: ok = launch_reorder<float>(d_major_order, d_src_major_ptr, d_src_minor_idx, d_src_val, major_dim, d_dst_major_ptr,
                                            d_dst_minor_idx, d_dst_val, s); break;
    }
    if (!ok) return 0;
    if (d_inverse_major_order != nullptr) {
        inverse_order_kernel<<<grid_for(major_dim, 256u, 8192u), 256, 0, s>>>(d_major_order, major_dim, d_inverse_major_order);
        CSAMD_HIP(hipGetLastError(), "inverse_order_kernel");
    }
    return 1;
}

int csamd_blocked_ell_bucket_plan_raw(const unsigned int *d_block_col_idx, unsigned int row_block_count, unsigned int width_blocks,
                                      unsigned int *d_order_out, unsigned int *d_width_sorted, unsigned long long *d_keys_sorted,
                                      csamd_convert_workspace *ws) {
    CSAMD_REQUIRE(ws != nullptr && d_order_out != nullptr, "blocked_ell_bucket_plan_raw: null argument");
    if (row_block_count == 0u) return 1;
    CSAMD_REQUIRE(d_block_col_idx != nullptr || width_blocks == 0u, "blocked_ell_bucket_plan_raw: null blockColIdx");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const size_t n = row_block_count;
    const size_t keys_bytes = align_up(n * 8u, 256u), ids_bytes = align_up(n * 4u, 256u);
    const size_t sort_bytes = radix_sort_pairs_u64_scratch_bytes(row_block_count);
    char *p = (char *) convert_scratch(ws, 2u * keys_bytes + ids_bytes + sort_bytes);
    if (p == nullptr) return 0;
    unsigned long long *keys_in = (unsigned long long *) p;
    unsigned long long *keys_out = d_keys_sorted != nullptr ? d_keys_sorted : (unsigned long long *) (p + keys_bytes);
    unsigned int *ids = (unsigned int *) (p + 2u * keys_bytes);
    void *sort_tmp = p + 2u * keys_bytes + ids_bytes;
    bell_bucket_keys_kernel<<<grid_for(row_block_count, 256u, 8192u), 256, 0, s>>>(row_block_count, width_blocks, d_block_col_idx,
                                                                                     keys_in, ids);
    CSAMD_HIP(hipGetLastError(), "bell_bucket_keys_kernel");
    if (!radix_sort_pairs_u64(keys_in, keys_out, ids, d_order_out, row_block_count, 64u, sort_tmp, sort_bytes, s)) return 0;
    if (d_width_sorted != nullptr) {
        key_high_kernel<<<grid_for(row_block_count, 256u, 8192u), 256, 0, s>>>(keys_out, row_block_count, d_width_sorted);
        CSAMD_HIP(hipGetLastError(), "key_high_kernel");
    }
    CSAMD_HIP(hipStreamSynchronize(s), "blocked_ell_bucket_plan sync");
    return 1;
}

} // extern "C"
