// Layout autotuner over a batch of CSR partitions, one launch for all of
// them: per partition, the Blocked-ELL shape at block sizes 8, 16 and 32,
// the Sliced-ELL padded slots at slice heights 32 and 64, and the log2
// histogram of row lengths; then, on the host, the reference's own choice
// rules per partition and the layout with the fewest format bytes.
//
// Reference rules:
//   Blocked-ELL width W = max over row blocks of the distinct column blocks
//   touched by stored entries (explicit zeros included), fill = sum / (row
//   blocks x W); choice: fewest rows x W x B x value bytes, then higher
//   fill (1e-9 slack), then larger B
//   (src/convert/blocked_ell_from_compressed.cuh:32-91, :157-190);
//   Sliced-ELL: slice height S = min(S, rows), width = max row length in
//   the slice, score = padded / nnz + 0.25 x slices / rows, ties to fewer
//   padded bytes, then larger S (src/convert/sliced_ell_from_blocked_ell.cuh:64-142,
//   :196-221).
//
// Design: one 256-thread workgroup per 64-row group of a partition (64 is
// a multiple of every candidate, so row blocks and slices never straddle
// groups).  The group's distinct 8-column blocks are marked in 8 LDS
// bitmaps, one per 8-row block; the 16- and 32-row blocks' 16- and
// 32-column blocks are ORs of 2 / 4 bitmaps folded over bit pairs / quads,
// so one pass over the index array serves all three block sizes.  A wave
// reads its row in 16-byte units and ORs each unit's bits per bitmap word
// in registers before the LDS atomic: the 4 indices of a sorted unit
// mostly share a word, and per-index atomics on one word from many lanes
// serialise.
#include "common.hh"

#include <vector>

namespace csamd {
namespace {

constexpr unsigned int PT_THREADS = 256u;
constexpr unsigned int PT_GROUP_ROWS = 64u;
constexpr unsigned int PT_MAX_WORDS = 2048u;  // 8 bitmaps x 2048 words = 64 KB of LDS: cols <= 524,288
constexpr unsigned int PT_HIST = 32u;
constexpr unsigned int PT_UNROLL = 4u;  // index loads in flight per lane

struct pt_part {
    const unsigned int *ptr;
    const unsigned int *idx;
    unsigned int rows;
    unsigned int group0;  // first 64-row group of the partition in the launch
};

struct pt_acc {
    unsigned long long sell_padded[2];  // S = 32, 64
    unsigned long long bell_total[3];   // B = 8, 16, 32: sum over row blocks of distinct column blocks
    unsigned int bell_width[3];         // max over row blocks
    unsigned int pad;
    unsigned int hist[PT_HIST];         // rows by log2 bucket of length
};

__global__ void pt_zero_kernel(pt_acc *__restrict__ acc, unsigned int n) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        acc[i] = pt_acc{};
    }
}

typedef unsigned int pt_u32x4 __attribute__((ext_vector_type(4)));

// Marks the 8-column blocks of one 16-byte unit of indices (a lane's 4
// consecutive entries of a row): bits that fall in the same bitmap word are
// ORed in registers first, so a sorted unit costs one LDS atomic.
__device__ __forceinline__ void pt_mark4(unsigned int *bm, pt_u32x4 q, bool valid, unsigned int cols) {
    unsigned int w[4], m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned int c = q[k];
        const bool ok = valid && c < cols;
        w[k] = ok ? (c >> 8) : 0xffffffffu;
        m[k] = ok ? 1u << ((c >> 3) & 31u) : 0u;
    }
    unsigned int cw = w[0], cm = m[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        if (w[k] == cw) {
            cm |= m[k];
        } else {
            if (cw != 0xffffffffu) atomicOr(&bm[cw], cm);
            cw = w[k];
            cm = m[k];
        }
    }
    if (cw != 0xffffffffu) atomicOr(&bm[cw], cm);
}

__device__ __forceinline__ unsigned int fold2(unsigned int x) { return (x | (x >> 1)) & 0x55555555u; }
__device__ __forceinline__ unsigned int fold4(unsigned int x) {
    const unsigned int y = x | (x >> 1);
    return (y | (y >> 2)) & 0x11111111u;
}

__device__ __forceinline__ unsigned int block_sum(unsigned int v, unsigned int *red) {
    v = wave_sum_u32(v);
    __syncthreads();
    if (lane_id() == 0u) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned int t = 0u;
    for (unsigned int k = 0; k < PT_THREADS / 64u; ++k) t += red[k];
    return t;
}

__global__ void __launch_bounds__(PT_THREADS) pt_tune_kernel(const pt_part *__restrict__ parts, unsigned int nparts,
                                                             unsigned int cols, unsigned int nw, pt_acc *__restrict__ acc) {
    extern __shared__ unsigned int bits[];  // [8][nw]
    __shared__ unsigned int red[PT_THREADS / 64u];
    __shared__ unsigned int hist[PT_HIST];
    __shared__ unsigned int rp[PT_GROUP_ROWS + 1u];  // the group's row pointers
    const unsigned int gi = blockIdx.x;
    unsigned int lo = 0u, hi = nparts - 1u;  // last partition whose first group <= gi
    while (lo < hi) {
        const unsigned int mid = (lo + hi + 1u) >> 1;
        if (parts[mid].group0 <= gi) lo = mid;
        else hi = mid - 1u;
    }
    const pt_part P = parts[lo];
    pt_acc *A = acc + lo;
    const unsigned int r0 = (gi - P.group0) * PT_GROUP_ROWS;
    if (r0 >= P.rows) return;  // uniform (empty partitions own no groups)
    const unsigned int nr = P.rows - r0 < PT_GROUP_ROWS ? P.rows - r0 : PT_GROUP_ROWS;
    for (unsigned int w = threadIdx.x; w < 8u * nw; w += PT_THREADS) bits[w] = 0u;
    if (threadIdx.x < PT_HIST) hist[threadIdx.x] = 0u;
    __syncthreads();

    // row lengths: slices and histogram (wave 0, one row per lane)
    if (threadIdx.x < 64u) {
        const unsigned int t = threadIdx.x;
        const unsigned int n = t < nr ? P.ptr[r0 + t + 1u] - P.ptr[r0 + t] : 0u;
        if (t < nr) {
            const unsigned int b = n == 0u ? 0u : 32u - (unsigned int) __clz(n);
            atomicAdd(&hist[b < PT_HIST ? b : PT_HIST - 1u], 1u);
        }
        unsigned int m32 = n;  // max over lanes 0-31 / 32-63 (the two 32-row slices)
        for (unsigned int o = 16u; o >= 1u; o >>= 1) m32 = max(m32, (unsigned int) __shfl_xor((int) m32, (int) o));
        const unsigned int m64 = max(m32, (unsigned int) __shfl_xor((int) m32, 32));
        const unsigned int m1 = (unsigned int) __shfl((int) m32, 32);  // all lanes active
        if (t == 0u) {
            const unsigned int rows0 = nr < 32u ? nr : 32u;
            const unsigned int rows1 = nr > 32u ? nr - 32u : 0u;
            atomicAdd(&A->sell_padded[0], (unsigned long long) rows0 * m32 + (unsigned long long) rows1 * m1);
            atomicAdd(&A->sell_padded[1], (unsigned long long) nr * m64);
        }
    }
    // distinct 8-column blocks per 8-row block: one wave per row.  The row is
    // read in 16-byte units (4 indices) from its first 16-byte boundary, with
    // PT_UNROLL units in flight per lane (loads clamped to the row's last
    // unit, never under a branch); lanes 0-5 mark the up to 3 + 3 indices
    // before the first and after the last whole unit.
    if (threadIdx.x <= nr) rp[threadIdx.x] = P.ptr[r0 + threadIdx.x];
    __syncthreads();
    const unsigned int wave = threadIdx.x >> 6, lane = lane_id();
    const unsigned int mis = (unsigned int) (((uintptr_t) P.idx >> 2) & 3u);  // P.idx's distance past a 16-byte boundary, in indices
    for (unsigned int r = wave; r < nr; r += PT_THREADS / 64u) {
        unsigned int *bm = bits + (r >> 3) * nw;
        const unsigned int b = rp[r], e = rp[r + 1u];
        const unsigned int a = min(b + ((4u - ((b + mis) & 3u)) & 3u), e);
        const unsigned int units = (e - a) >> 2;
        const unsigned int ea = a + 4u * units;
        if (lane < 6u) {
            const unsigned int i = lane < 3u ? b + lane : ea + lane - 3u;
            if (lane < 3u ? i < a : i < e) {
                const unsigned int c = P.idx[i];
                if (c < cols) atomicOr(&bm[c >> 8], 1u << ((c >> 3) & 31u));
            }
        }
        const pt_u32x4 *U = (const pt_u32x4 *) (P.idx + a);
        for (unsigned int j0 = lane; j0 < units; j0 += 64u * PT_UNROLL) {
            pt_u32x4 q[PT_UNROLL];
#pragma unroll
            for (unsigned int u = 0; u < PT_UNROLL; ++u) q[u] = U[min(j0 + 64u * u, units - 1u)];
#pragma unroll
            for (unsigned int u = 0; u < PT_UNROLL; ++u) pt_mark4(bm, q[u], j0 + 64u * u < units, cols);
        }
    }
    __syncthreads();
    if (threadIdx.x < PT_HIST && hist[threadIdx.x] != 0u) atomicAdd(&A->hist[threadIdx.x], hist[threadIdx.x]);
    unsigned int c8[8], c16[4], c32[2];
#pragma unroll
    for (unsigned int k = 0; k < 8u; ++k) c8[k] = 0u;
#pragma unroll
    for (unsigned int k = 0; k < 4u; ++k) c16[k] = 0u;
    c32[0] = c32[1] = 0u;
    for (unsigned int w = threadIdx.x; w < nw; w += PT_THREADS) {
        unsigned int x[8];
#pragma unroll
        for (unsigned int k = 0; k < 8u; ++k) {
            x[k] = bits[k * nw + w];
            c8[k] += (unsigned int) __popc(x[k]);
        }
#pragma unroll
        for (unsigned int k = 0; k < 4u; ++k) c16[k] += (unsigned int) __popc(fold2(x[2u * k] | x[2u * k + 1u]));
#pragma unroll
        for (unsigned int k = 0; k < 2u; ++k) c32[k] += (unsigned int) __popc(fold4(x[4u * k] | x[4u * k + 1u] | x[4u * k + 2u] | x[4u * k + 3u]));
    }
    // row blocks present in the group: 8-row ceil(nr / 8), 16-row ceil(nr / 16), 32-row ceil(nr / 32)
    unsigned int w8 = 0u, w16 = 0u, w32 = 0u;
    unsigned long long t8 = 0ull, t16 = 0ull, t32 = 0ull;
#pragma unroll
    for (unsigned int k = 0; k < 8u; ++k) {
        const unsigned int u = block_sum(c8[k], red);
        if (8u * k < nr) {
            w8 = max(w8, u);
            t8 += u;
        }
    }
#pragma unroll
    for (unsigned int k = 0; k < 4u; ++k) {
        const unsigned int u = block_sum(c16[k], red);
        if (16u * k < nr) {
            w16 = max(w16, u);
            t16 += u;
        }
    }
#pragma unroll
    for (unsigned int k = 0; k < 2u; ++k) {
        const unsigned int u = block_sum(c32[k], red);
        if (32u * k < nr) {
            w32 = max(w32, u);
            t32 += u;
        }
    }
    if (threadIdx.x == 0u) {
        atomicMax(&A->bell_width[0], w8);
        atomicMax(&A->bell_width[1], w16);
        atomicMax(&A->bell_width[2], w32);
        atomicAdd(&A->bell_total[0], t8);
        atomicAdd(&A->bell_total[1], t16);
        atomicAdd(&A->bell_total[2], t32);
    }
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

int csamd_tune_partitions(const csamd_compressed_view *partitions, unsigned int partition_count, unsigned int value_type,
                          csamd_convert_workspace *ws, csamd_partition_tune_result *out) {
    CSAMD_REQUIRE(partitions != nullptr && ws != nullptr && out != nullptr && partition_count != 0u,
                  "tune_partitions: bad arguments");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb > 0, "tune_partitions: unsupported value type %u", value_type);
    const unsigned int cols = partitions[0].cols;
    std::vector<pt_part> hp(partition_count);
    unsigned long long groups = 0ull;
    for (unsigned int p = 0; p < partition_count; ++p) {
        const csamd_compressed_view &v = partitions[p];
        CSAMD_REQUIRE(v.axis == CSAMD_COMPRESSED_BY_ROW, "tune_partitions: partition %u is not row-major", p);
        CSAMD_REQUIRE(v.cols == cols, "tune_partitions: partitions must have the same column count");
        CSAMD_REQUIRE(v.rows == 0u || v.majorPtr != nullptr, "tune_partitions: partition %u has no majorPtr", p);
        hp[p] = pt_part{v.majorPtr, v.minorIdx, v.rows, (unsigned int) groups};
        groups += (v.rows + PT_GROUP_ROWS - 1u) / PT_GROUP_ROWS;
    }
    CSAMD_REQUIRE(groups < 0x7fffffffull, "tune_partitions: too many rows");
    const unsigned int nw = ((cols + 7u) / 8u + 31u) / 32u;
    CSAMD_REQUIRE(nw <= PT_MAX_WORDS, "tune_partitions: %u columns exceed the LDS bitmaps", cols);
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const size_t parts_bytes = align_up((size_t) partition_count * sizeof(pt_part), 256u);
    const size_t acc_bytes = (size_t) partition_count * sizeof(pt_acc);
    char *scratch = (char *) convert_scratch(ws, parts_bytes + acc_bytes);
    if (scratch == nullptr) return 0;
    pt_part *d_parts = (pt_part *) scratch;
    pt_acc *d_acc = (pt_acc *) (scratch + parts_bytes);
    CSAMD_HIP(hipMemcpyAsync(d_parts, hp.data(), (size_t) partition_count * sizeof(pt_part), hipMemcpyHostToDevice, s),
              "H2D tune partitions");
    pt_zero_kernel<<<grid_for(partition_count, 256u, 64u), 256, 0, s>>>(d_acc, partition_count);
    CSAMD_HIP(hipGetLastError(), "pt_zero_kernel");
    if (groups != 0u && cols != 0u) {
        const size_t lds = (size_t) 8u * nw * sizeof(unsigned int);
        auto k = pt_tune_kernel;
        CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        k<<<(unsigned int) groups, PT_THREADS, lds, s>>>(d_parts, partition_count, cols, nw, d_acc);
        CSAMD_HIP(hipGetLastError(), "pt_tune_kernel");
    }
    std::vector<pt_acc> acc(partition_count);
    CSAMD_HIP(hipMemcpyAsync(acc.data(), d_acc, acc_bytes, hipMemcpyDeviceToHost, s), "D2H tune results");
    CSAMD_HIP(hipStreamSynchronize(s), "tune sync");

    static const unsigned int bell_cands[3] = {8u, 16u, 32u};
    static const unsigned int sell_cands[2] = {32u, 64u};
    for (unsigned int p = 0; p < partition_count; ++p) {
        const csamd_compressed_view &v = partitions[p];
        const pt_acc &a = acc[p];
        csamd_partition_tune_result &r = out[p];
        std::memset(&r, 0, sizeof(r));
        const unsigned int rows = v.rows;
        const unsigned long long nnz = v.nnz;
        // Blocked-ELL: the reference's choice among 8, 16, 32
        size_t best_pad = 0u;
        double best_fill = 0.0;
        unsigned int best_bs = 0u, best_w = 0u;
        for (unsigned int i = 0; i < 3u; ++i) {
            const unsigned int bs = bell_cands[i];
            const unsigned int row_blocks = (rows + bs - 1u) / bs;
            const unsigned int w = (rows == 0u || cols == 0u) ? 0u : a.bell_width[i];
            double fill = (row_blocks == 0u || w == 0u) ? 1.0 : (double) a.bell_total[i] / ((double) row_blocks * (double) w);
            if (rows == 0u || cols == 0u) fill = 0.0;
            const size_t padded = (size_t) rows * (size_t) w * (size_t) bs * (size_t) vb;
            if (best_bs == 0u || padded + 1u < best_pad || (padded == best_pad && fill > best_fill + 1.0e-9) ||
                (padded == best_pad && fill + 1.0e-9 >= best_fill && bs > best_bs)) {
                best_bs = bs;
                best_fill = fill;
                best_pad = padded;
                best_w = w;
            }
        }
        r.bell_block_size = best_bs;
        r.bell_ell_width = best_w;
        r.bell_fill_ratio = best_fill;
        r.bell_bytes = (unsigned long long) ((rows + best_bs - 1u) / best_bs) * best_w * 4ull + (unsigned long long) best_pad;
        // Sliced-ELL: the reference's choice among 32, 64
        double best_score = __builtin_inf();
        unsigned int best_s = 0u, best_count = 0u;
        unsigned long long best_slots = 0ull, best_pbytes = 0ull;
        for (unsigned int i = 0; i < 2u; ++i) {
            unsigned int S = 0u, count = 0u;
            unsigned long long padded = 0ull, pbytes = 0ull;
            double score = 1.0;
            if (rows != 0u) {
                S = sell_cands[i] > rows ? rows : sell_cands[i];
                count = (rows + S - 1u) / S;
                // S < the candidate only when rows < the candidate: one slice of every row,
                // which is what the group pass measured for either candidate
                padded = a.sell_padded[i];
                pbytes = padded * 6ull + ((unsigned long long) count * 2ull + 1ull) * 4ull;
                score = (nnz == 0ull ? 1.0 : (double) padded / (double) nnz) + 0.25 * ((double) count / (double) rows);
            }
            if (best_s == 0u || score + 1.0e-12 < best_score || (score <= best_score + 1.0e-12 && pbytes < best_pbytes) ||
                (score <= best_score + 1.0e-12 && pbytes == best_pbytes && S > best_s)) {
                best_score = score;
                best_s = S;
                best_count = count;
                best_slots = padded;
                best_pbytes = pbytes;
            }
        }
        r.sell_slice_rows = best_s;
        r.sell_slice_count = best_count;
        r.sell_padded_slots = best_slots;
        r.sell_score = best_score;
        r.sell_bytes = best_slots * (4ull + (unsigned long long) vb) + ((unsigned long long) best_count * 2ull + 1ull) * 4ull;
        r.csr_bytes = ((unsigned long long) rows + 1ull) * 4ull + nnz * (4ull + (unsigned long long) vb);
        // the layout with the fewest format bytes (ties: compressed, then sliced, then blocked)
        r.layout = CSAMD_LAYOUT_COMPRESSED;
        unsigned long long best = r.csr_bytes;
        if (r.sell_bytes < best) {
            r.layout = CSAMD_LAYOUT_SLICED_ELL;
            best = r.sell_bytes;
        }
        if (r.bell_bytes < best) r.layout = CSAMD_LAYOUT_BLOCKED_ELL;
        for (unsigned int b = 0; b < PT_HIST; ++b) r.row_length_log2_hist[b] = a.hist[b];
    }
    return 1;
}

} // extern "C"
