// Column (and row) orders that bring columns with similar row support next
// to each other: the lever against Blocked-ELL padding.
//
// Reference:
//   include/CellShard/ingest/internal/dataset/internal/convert_layout_support_part.hh:3-86
//     (shard_column_signature, mix_signature, build_shard_column_maps): per
//     column, the row blocks of the shard's Blocked-ELL partitions in which
//     it has a nonzero stored value, folded in ascending global row-block
//     order into two 64-bit FNV-1a style hashes, the support count and the
//     first row block; columns stably sorted by (supported first, first row
//     block, hash a, hash b, support descending, column).
//   src/bucket/operators/blocked_ell_bipartite_optimize.cuh:223-500
//     (accumulate_{column,row}_score, build_entity_sort_keys, sort_entities,
//     build_column_order_from_sampled_coo, build_row_order_from_coo): an
//     entity's barycenter is the mean rank of the entities it touches
//     (FLT_MAX when it touches none); entities sorted by (barycenter,
//     index).
//
// gfx950 plan.  Signatures: a presence bitmap [global row block][column]
// built by one workgroup per row block (threads across the ell positions,
// an OR over the block's rows), then one thread per column folds its bitmap
// column in row-block order; the stable sort of the column signatures (one
// per column) runs on the host, as in the reference.  Barycenters: exact
// u64 rank sums and u32 counts (atomics), the quotient rounded once as the
// reference's float division of its float sum would be when that sum is
// exact (below 2^24), 64-bit keys (ordered barycenter bits, index) sorted by
// the library's stable radix sort.  The reference's float atomics make its
// order depend on the order of additions once a sum passes 2^24; this one
// does not.
#include "common.hh"

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <vector>

namespace csamd {
namespace {

constexpr unsigned int CO_INVALID = 0xffffffffu;

template<typename S> __device__ __forceinline__ bool co_nonzero(S raw);
template<> __device__ __forceinline__ bool co_nonzero<uint16_t>(uint16_t raw) { return (raw & 0x7fffu) != 0u; }
template<> __device__ __forceinline__ bool co_nonzero<float>(float raw) { return raw != 0.0f; }
template<> __device__ __forceinline__ bool co_nonzero<uint8_t>(uint8_t raw) { return (raw & 0x7fu) != 0u; }

// bits[(g0 + rb) * words + col / 32] bit col % 32: column col has a nonzero
// stored value in row block rb of this partition.
template<typename S>
__global__ void __launch_bounds__(256) co_presence_kernel(csamd_blocked_ell_view part, const S *__restrict__ val, unsigned int g0,
                                                          unsigned int cols, unsigned int words, unsigned int *__restrict__ bits) {
    const unsigned int bs = part.block_size;
    const unsigned int width = part.ell_cols / bs;
    const unsigned int row_blocks = (part.rows + bs - 1u) / bs;
    for (unsigned int rb = blockIdx.x; rb < row_blocks; rb += gridDim.x) {
        const unsigned int r0 = rb * bs;
        const unsigned int nr = part.rows - r0 < bs ? part.rows - r0 : bs;
        unsigned int *brow = bits + (size_t) (g0 + rb) * words;
        for (unsigned int pos = threadIdx.x; pos < part.ell_cols; pos += blockDim.x) {
            const unsigned int slot = pos / bs;
            const unsigned int bc = part.blockColIdx[(size_t) rb * width + slot];
            if (bc == CO_INVALID) continue;
            const unsigned int col = bc * bs + (pos - slot * bs);
            if (col >= cols) continue;
            bool any = false;
            for (unsigned int i = 0; i < nr && !any; ++i) any = co_nonzero<S>(val[(size_t) (r0 + i) * part.ell_cols + pos]);
            if (any) atomicOr(&brow[col >> 5], 1u << (col & 31u));
        }
    }
}

// mix_signature: the value's 8 bytes, low first, each xored in and
// multiplied by the 64-bit FNV prime.
__device__ __forceinline__ unsigned long long co_mix(unsigned long long seed, unsigned long long value) {
#pragma unroll
    for (unsigned int k = 0; k < 8u; ++k) {
        seed ^= (value >> (8u * k)) & 0xffull;
        seed *= 1099511628211ull;
    }
    return seed;
}

// One thread per column: its signature over the row blocks in order.
__global__ void co_signature_kernel(unsigned int cols, unsigned int words, unsigned int row_blocks,
                                    const unsigned int *__restrict__ bits, unsigned int *__restrict__ support,
                                    unsigned long long *__restrict__ hash_a, unsigned long long *__restrict__ hash_b,
                                    unsigned int *__restrict__ min_rb) {
    for (unsigned int col = blockIdx.x * blockDim.x + threadIdx.x; col < cols; col += gridDim.x * blockDim.x) {
        unsigned int n = 0u, first = CO_INVALID;
        unsigned long long a = 1469598103934665603ull, b = 1099511628211ull;
        const unsigned int w = col >> 5, m = 1u << (col & 31u);
        for (unsigned int g = 0; g < row_blocks; ++g) {
            if ((bits[(size_t) g * words + w] & m) == 0u) continue;
            ++n;
            a = co_mix(a, (unsigned long long) g + 1ull);
            b = co_mix(b, ((unsigned long long) g + 1ull) * 1315423911ull);
            first = first < g ? first : g;
        }
        support[col] = n;
        hash_a[col] = a;
        hash_b[col] = b;
        min_rb[col] = first;
    }
}

// Exact rank sums of the entities (columns when by_col, else rows).
__global__ void co_accumulate_kernel(unsigned int nnz, const unsigned int *__restrict__ row_idx, const unsigned int *__restrict__ col_idx,
                                     const unsigned int *__restrict__ other_rank, int by_col, unsigned int entities,
                                     unsigned int others, unsigned long long *__restrict__ sum, unsigned int *__restrict__ count,
                                     unsigned int *__restrict__ bad) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
        const unsigned int e = by_col ? col_idx[i] : row_idx[i];
        const unsigned int o = by_col ? row_idx[i] : col_idx[i];
        if (e >= entities || o >= others) {
            *bad = 1u;
            continue;
        }
        const unsigned int rank = other_rank[o];
        if (rank == CO_INVALID) continue;
        atomicAdd(&sum[e], (unsigned long long) rank);
        atomicAdd(&count[e], 1u);
    }
}

__device__ __forceinline__ unsigned int co_ordered(float v) {
    const unsigned int bits = __float_as_uint(v);
    return bits ^ ((bits & 0x80000000u) != 0u ? 0xffffffffu : 0x80000000u);
}

__global__ void co_keys_kernel(unsigned int n, const unsigned long long *__restrict__ sum, const unsigned int *__restrict__ count,
                               unsigned long long *__restrict__ keys, unsigned int *__restrict__ order) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned int w = count[i];
        const float bary = w != 0u ? __fdiv_rn((float) sum[i], (float) w) : FLT_MAX;
        keys[i] = ((unsigned long long) co_ordered(bary) << 32) | i;
        order[i] = i;
    }
}

struct co_signature {
    unsigned int canonical_col, support, min_row_block;
    unsigned long long hash_a, hash_b;
};

// Device buffer freed at scope exit.
struct dev_buf {
    void *p = nullptr;
    ~dev_buf() {
        if (p != nullptr) (void) hipFree(p);
    }
    int alloc(size_t n, const char *what) {
        if (hipMalloc(&p, n == 0u ? 256u : n) != hipSuccess) {
            set_error("hipMalloc %s (%zu bytes) failed", what, n);
            p = nullptr;
            return 0;
        }
        return 1;
    }
};

template<typename S> struct co_presence_launch {
    static int run(const csamd_blocked_ell_view *part, unsigned int g0, unsigned int cols, unsigned int words, unsigned int *bits,
                   hipStream_t s) {
        const unsigned int row_blocks = (part->rows + part->block_size - 1u) / part->block_size;
        co_presence_kernel<S><<<grid_for(row_blocks, 1u, (unsigned int) device_cu_count() * 8u), 256, 0, s>>>(
            *part, (const S *) part->val, g0, cols, words, bits);
        return hip_ok(hipGetLastError(), "co_presence_kernel");
    }
};

int invert(const unsigned int *exec_to_canonical, unsigned int n, unsigned int *canonical_to_exec, const char *who) {
    for (unsigned int i = 0; i < n; ++i) canonical_to_exec[i] = CO_INVALID;
    for (unsigned int e = 0; e < n; ++e) {
        const unsigned int c = exec_to_canonical[e];
        CSAMD_REQUIRE(c < n, "%s: order entry %u out of range", who, c);
        canonical_to_exec[c] = e;
    }
    return 1;
}

// Barycenter order of `entities` (columns when by_col) from a host COO and
// the host ranks of the other side.
int barycenter_order(const csamd_host_coo *coo, const unsigned int *h_other_rank, int by_col, int device, unsigned int *h_exec_to_canonical,
                     unsigned int *h_canonical_to_exec, const char *who) {
    CSAMD_REQUIRE(coo != nullptr && h_other_rank != nullptr && h_exec_to_canonical != nullptr && h_canonical_to_exec != nullptr,
                  "%s: null argument", who);
    CSAMD_REQUIRE(coo->nnz == 0u || (coo->rowIdx != nullptr && coo->colIdx != nullptr), "%s: COO arrays missing", who);
    const unsigned int entities = by_col ? coo->cols : coo->rows;
    const unsigned int others = by_col ? coo->rows : coo->cols;
    if (entities == 0u) return 1;
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice column order");
    hipStream_t s = nullptr;
    CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags column order");
    struct stream_guard {
        hipStream_t s;
        ~stream_guard() { (void) hipStreamDestroy(s); }
    } guard{s};
    const size_t nnz = coo->nnz;
    const size_t sort_bytes = radix_sort_pairs_u64_scratch_bytes(entities);
    dev_buf idx, rank, acc, keys, ord, tmp;
    if (!idx.alloc(2u * nnz * 4u, "COO indices") || !rank.alloc((size_t) others * 4u, "ranks") ||
        !acc.alloc((size_t) entities * 12u + 256u, "rank sums") || !keys.alloc((size_t) entities * 16u, "sort keys") ||
        !ord.alloc((size_t) entities * 8u, "orders") || !tmp.alloc(sort_bytes, "sort scratch"))
        return 0;
    unsigned int *d_row = (unsigned int *) idx.p, *d_col = d_row + nnz;
    unsigned long long *d_sum = (unsigned long long *) acc.p;
    unsigned int *d_count = (unsigned int *) (d_sum + entities);
    unsigned int *d_bad = d_count + entities;
    unsigned long long *d_keys_in = (unsigned long long *) keys.p, *d_keys_out = d_keys_in + entities;
    unsigned int *d_ord_in = (unsigned int *) ord.p, *d_ord_out = d_ord_in + entities;
    if (nnz != 0u) {
        CSAMD_HIP(hipMemcpyAsync(d_row, coo->rowIdx, nnz * 4u, hipMemcpyHostToDevice, s), "H2D COO rows");
        CSAMD_HIP(hipMemcpyAsync(d_col, coo->colIdx, nnz * 4u, hipMemcpyHostToDevice, s), "H2D COO cols");
    }
    if (others != 0u) CSAMD_HIP(hipMemcpyAsync(rank.p, h_other_rank, (size_t) others * 4u, hipMemcpyHostToDevice, s), "H2D ranks");
    if (!zero_bytes(acc.p, (size_t) entities * 12u + 4u, s)) return 0;
    if (nnz != 0u) {
        co_accumulate_kernel<<<grid_for(nnz, 256u, 4096u), 256, 0, s>>>((unsigned int) nnz, d_row, d_col, (const unsigned int *) rank.p,
                                                                       by_col, entities, others, d_sum, d_count, d_bad);
        CSAMD_HIP(hipGetLastError(), "co_accumulate_kernel");
    }
    co_keys_kernel<<<grid_for(entities, 256u, 4096u), 256, 0, s>>>(entities, d_sum, d_count, d_keys_in, d_ord_in);
    CSAMD_HIP(hipGetLastError(), "co_keys_kernel");
    if (!radix_sort_pairs_u64(d_keys_in, d_keys_out, d_ord_in, d_ord_out, entities, 64u, tmp.p, sort_bytes, s)) return 0;
    unsigned int bad = 0u;
    CSAMD_HIP(hipMemcpyAsync(&bad, d_bad, 4u, hipMemcpyDeviceToHost, s), "D2H column order check");
    CSAMD_HIP(hipMemcpyAsync(h_exec_to_canonical, d_ord_out, (size_t) entities * 4u, hipMemcpyDeviceToHost, s), "D2H order");
    CSAMD_HIP(hipStreamSynchronize(s), "column order sync");
    CSAMD_REQUIRE(bad == 0u, "%s: a COO index exceeds the matrix", who);
    return invert(h_exec_to_canonical, entities, h_canonical_to_exec, who);
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

int csamd_build_shard_column_maps(const csamd_blocked_ell_view *parts, unsigned int part_count, unsigned int value_type,
                                  unsigned int cols, unsigned int *h_exec_to_canonical, unsigned int *h_canonical_to_exec,
                                  csamd_convert_workspace *ws) {
    const char *who = "build_shard_column_maps";
    CSAMD_REQUIRE(h_exec_to_canonical != nullptr && h_canonical_to_exec != nullptr && ws != nullptr, "%s: null argument", who);
    CSAMD_REQUIRE(part_count == 0u || parts != nullptr, "%s: null partitions", who);
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "%s: unsupported value type %u", who, value_type);
    if (cols == 0u) return 1;
    unsigned long long row_blocks = 0ull;
    for (unsigned int p = 0; p < part_count; ++p) {
        CSAMD_REQUIRE(parts[p].rows == 0u || parts[p].block_size != 0u, "%s: block_size is 0", who);
        CSAMD_REQUIRE(parts[p].block_size == 0u || parts[p].ell_cols % parts[p].block_size == 0u,
                      "%s: ell_cols not a multiple of block_size", who);
        if (parts[p].rows != 0u) row_blocks += (parts[p].rows + parts[p].block_size - 1u) / parts[p].block_size;
    }
    CSAMD_REQUIRE(row_blocks <= 0xffffffffull, "%s: too many row blocks", who);
    CSAMD_HIP(hipSetDevice(ws->device >= 0 ? ws->device : 0), "hipSetDevice shard column maps");
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int words = (cols + 31u) / 32u;
    const size_t bits_bytes = align_up((size_t) row_blocks * words * 4u, 256u);
    const size_t sig_bytes = (size_t) cols * (4u + 8u + 8u + 4u);
    dev_buf buf;
    if (!buf.alloc(bits_bytes + sig_bytes, "column presence")) return 0;
    unsigned int *bits = (unsigned int *) buf.p;
    unsigned long long *ha = (unsigned long long *) ((char *) buf.p + bits_bytes);
    unsigned long long *hb = ha + cols;
    unsigned int *support = (unsigned int *) (hb + cols);
    unsigned int *minrb = support + cols;
    if (!zero_bytes(bits, bits_bytes, s)) return 0;
    unsigned int g0 = 0u;
    for (unsigned int p = 0; p < part_count; ++p) {
        const csamd_blocked_ell_view &v = parts[p];
        if (v.rows == 0u) continue;
        CSAMD_REQUIRE(v.ell_cols == 0u || (v.blockColIdx != nullptr && v.val != nullptr), "%s: partition arrays missing", who);
        if (v.ell_cols != 0u) {
            int ok = 0;
            switch (vb) {
                case 1: ok = co_presence_launch<uint8_t>::run(&v, g0, cols, words, bits, s); break;
                case 2: ok = co_presence_launch<uint16_t>::run(&v, g0, cols, words, bits, s); break;
                default: ok = co_presence_launch<float>::run(&v, g0, cols, words, bits, s); break;
            }
            if (!ok) return 0;
        }
        g0 += (v.rows + v.block_size - 1u) / v.block_size;
    }
    co_signature_kernel<<<grid_for(cols, 64u, 4096u), 64, 0, s>>>(cols, words, (unsigned int) row_blocks, bits, support, ha, hb, minrb);
    CSAMD_HIP(hipGetLastError(), "co_signature_kernel");
    std::vector<unsigned int> h_support(cols), h_min(cols);
    std::vector<unsigned long long> h_a(cols), h_b(cols);
    CSAMD_HIP(hipMemcpyAsync(h_support.data(), support, (size_t) cols * 4u, hipMemcpyDeviceToHost, s), "D2H support");
    CSAMD_HIP(hipMemcpyAsync(h_min.data(), minrb, (size_t) cols * 4u, hipMemcpyDeviceToHost, s), "D2H first row block");
    CSAMD_HIP(hipMemcpyAsync(h_a.data(), ha, (size_t) cols * 8u, hipMemcpyDeviceToHost, s), "D2H hash a");
    CSAMD_HIP(hipMemcpyAsync(h_b.data(), hb, (size_t) cols * 8u, hipMemcpyDeviceToHost, s), "D2H hash b");
    CSAMD_HIP(hipStreamSynchronize(s), "shard column maps sync");
    std::vector<co_signature> sig(cols);
    for (unsigned int c = 0; c < cols; ++c) sig[c] = co_signature{c, h_support[c], h_min[c], h_a[c], h_b[c]};
    std::stable_sort(sig.begin(), sig.end(), [](const co_signature &l, const co_signature &r) {
        if ((l.support == 0u) != (r.support == 0u)) return l.support != 0u;  // supported columns first
        if (l.min_row_block != r.min_row_block) return l.min_row_block < r.min_row_block;
        if (l.hash_a != r.hash_a) return l.hash_a < r.hash_a;
        if (l.hash_b != r.hash_b) return l.hash_b < r.hash_b;
        if (l.support != r.support) return l.support > r.support;
        return l.canonical_col < r.canonical_col;
    });
    for (unsigned int e = 0; e < cols; ++e) {
        h_exec_to_canonical[e] = sig[e].canonical_col;
        h_canonical_to_exec[sig[e].canonical_col] = e;
    }
    return 1;
}

int csamd_build_column_order_from_sampled_coo(const csamd_host_coo *sampled, const unsigned int *h_sample_row_rank, int device,
                                              unsigned int *h_exec_to_canonical_cols, unsigned int *h_canonical_to_exec_cols) {
    return barycenter_order(sampled, h_sample_row_rank, 1, device, h_exec_to_canonical_cols, h_canonical_to_exec_cols,
                            "build_column_order_from_sampled_coo");
}

int csamd_build_row_order_from_coo(const csamd_host_coo *src, const unsigned int *h_canonical_to_exec_cols, int device,
                                   unsigned int *h_exec_to_canonical_rows, unsigned int *h_canonical_to_exec_rows) {
    return barycenter_order(src, h_canonical_to_exec_cols, 0, device, h_exec_to_canonical_rows, h_canonical_to_exec_rows,
                            "build_row_order_from_coo");
}

} // extern "C"
