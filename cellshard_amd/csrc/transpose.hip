// Deterministic compressed transpose (CSR -> CSC or CSC -> CSR).
//
// Reference: src/convert/compressed_transpose_raw.cuh:14-122 (cuSPARSE
// csr2cscEx2, else count / scan / atomic scatter with nondeterministic order
// inside a column).  Here the output order is canonical: inside every output
// segment the source major index ascends (what csr2cscEx2 and scipy give).
//
// gfx950 plan (sources whose rows hold strictly ascending minor indices):
//   K1 count: one 1024-thread workgroup per tile of TR_ROWS source rows, a
//      u32 LDS histogram over all output segments (<= 40k fit in LDS), one
//      wave per row; the same pass records where each TR_BAND-wide column
//      band starts inside each row, band-major per tile (rbt[t][band][row]);
//   K2 per segment: exclusive running sums over tiles in place (each
//      tile's first position inside the segment) and the segment totals;
//      a scan of the totals gives out_ptr;
//   K4 one workgroup per (tile, band) chunk: the chunk's entries are sorted
//      in LDS by output segment, stably by source row: every segment of the
//      band keeps a TR_ROWS-bit mask of the tile rows that hold it, an
//      entry's rank is the popcount of the mask below its row; then each
//      segment's run (its entries from this tile, rows ascending) is written
//      to its final place with consecutive lanes on consecutive positions.
//      Chunks larger than the LDS staging go in batches of whole rows.
// Sources with unsorted or repeated minor indices take a fallback scatter
// (K4u: LDS cursors, rows in order, entries of one row in any order).
#include "common.hh"

#include <cstdlib>

namespace csamd {
namespace {

constexpr unsigned int TR_ROWS = 256u;          // source rows per tile (bits per segment mask)
constexpr unsigned int TR_MASK_WORDS = TR_ROWS / 32u;
constexpr unsigned int TR_BAND = 256u;          // output segments per band
constexpr unsigned int TR_MAX_BANDS = 4096u;    // beyond: fallback path (row band table would be too large)
constexpr unsigned int TR_COUNT_THREADS = 1024u;
constexpr unsigned int TR_SORT_THREADS = 256u;
constexpr unsigned int TR_U_BAND = 4096u;       // fallback: segments per band (16 KB of LDS cursors)

struct tr_layout {
    unsigned int tiles;
    unsigned int bands;     // TR_BAND-wide bands (sorted path)
    unsigned int ubands;    // TR_U_BAND-wide bands (fallback)
    bool band_table;        // the sorted path's row band table fits
    size_t off_counts, off_rbt, off_coltot, off_flags, off_scan, total;
};

tr_layout make_layout(unsigned int cDim, unsigned int uDim) {
    tr_layout L;
    L.tiles = cDim == 0u ? 1u : (cDim + TR_ROWS - 1u) / TR_ROWS;
    L.bands = uDim == 0u ? 1u : (uDim + TR_BAND - 1u) / TR_BAND;
    L.ubands = uDim == 0u ? 1u : (uDim + TR_U_BAND - 1u) / TR_U_BAND;
    L.band_table = L.bands <= TR_MAX_BANDS;
    size_t p = 0;
    L.off_flags = p;
    p = align_up(p + 256u, 256u);
    L.off_counts = p;
    p = align_up(p + (size_t) L.tiles * uDim * 4u, 256u);
    L.off_rbt = p;
    if (L.band_table) p = align_up(p + (size_t) L.tiles * (L.bands + 1u) * TR_ROWS * 4u, 256u);
    L.off_coltot = p;
    p = align_up(p + (size_t) (uDim + 1u) * 4u, 256u);
    L.off_scan = p;
    p = align_up(p + exclusive_scan_scratch_bytes(uDim), 256u);
    L.total = p;
    return L;
}

// K1.  flags[0]: some row's minor indices are not strictly ascending;
// flags[1]: a minor index >= uDim.
__global__ void __launch_bounds__(TR_COUNT_THREADS) tr_count_kernel(unsigned int cDim, unsigned int uDim, const unsigned int *__restrict__ ptr,
                                                                    const unsigned int *__restrict__ idx, unsigned int bands,
                                                                    int use_lds, unsigned int *__restrict__ counts,
                                                                    unsigned int *__restrict__ rbt, unsigned int *__restrict__ flags) {
    extern __shared__ unsigned int hist[];
    const unsigned int t = blockIdx.x;
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int r1 = r0 + TR_ROWS < cDim ? r0 + TR_ROWS : cDim;
    unsigned int *tile_counts = counts + (size_t) t * uDim;
    unsigned int *tile_rbt = rbt != nullptr ? rbt + (size_t) t * (bands + 1u) * TR_ROWS : nullptr;
    if (use_lds) {
        for (unsigned int c = threadIdx.x; c < uDim; c += blockDim.x) hist[c] = 0u;
        __syncthreads();
    }
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = lane_id();
    for (unsigned int r = r0 + wave; r < r1; r += nw) {
        const unsigned int rl = r - r0;
        const unsigned int b = ptr[r], e = ptr[r + 1u];
        if (b == e) {
            if (tile_rbt != nullptr)
                for (unsigned int k = lane; k <= bands; k += 64u) tile_rbt[(size_t) k * TR_ROWS + rl] = b;
            continue;
        }
        // 256 entries per pass, four independent loads per lane; the
        // previous entry's column comes from the neighbour lane
        unsigned int carry = CSAMD_INVALID_COL;  // column of the entry before the pass
        for (unsigned int i0 = b; i0 < e; i0 += 256u) {
            unsigned int c[4];
#pragma unroll
            for (unsigned int u = 0; u < 4u; ++u) {
                const unsigned int i = i0 + 64u * u + lane;
                c[u] = i < e ? idx[i] : CSAMD_INVALID_COL;
            }
#pragma unroll
            for (unsigned int u = 0; u < 4u; ++u) {
                const unsigned int i = i0 + 64u * u + lane;
                const unsigned int up = (unsigned int) __shfl_up((int) c[u], 1);
                const unsigned int block_prev = u == 0u ? carry : (unsigned int) __shfl((int) c[u == 0u ? 0u : u - 1u], 63);
                const unsigned int pc = lane == 0u ? block_prev : up;
                if (i >= e) continue;
                const unsigned int cc = c[u];
                if (cc >= uDim) {
                    flags[1] = 1u;
                    continue;
                }
                if (use_lds) atomicAdd(&hist[cc], 1u);
                else atomicAdd(&tile_counts[cc], 1u);
                if (tile_rbt == nullptr) continue;
                // band starts: entry i opens every band in (band of entry i - 1, band of cc]
                const unsigned int band = cc / TR_BAND;
                int prev = -1;
                if (i > b) {
                    prev = pc < uDim ? (int) (pc / TR_BAND) : (int) bands;
                    if (pc >= cc) flags[0] = 1u;
                }
                for (int k = prev + 1; k <= (int) band; ++k) tile_rbt[(size_t) k * TR_ROWS + rl] = i;
                if (i == e - 1u)
                    for (unsigned int k = band + 1u; k <= bands; ++k) tile_rbt[(size_t) k * TR_ROWS + rl] = e;
            }
            carry = (unsigned int) __shfl((int) c[3], 63);
        }
    }
    if (use_lds) {
        __syncthreads();
        for (unsigned int c = threadIdx.x; c < uDim; c += blockDim.x) tile_counts[c] = hist[c];
    }
}

// K2: counts[t][c] <- sum_{t' < t} counts[t'][c] (in place), coltot[c] = total.
__global__ void tr_tile_offsets_kernel(unsigned int *__restrict__ counts, unsigned int tiles, unsigned int uDim,
                                       unsigned int *__restrict__ coltot) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < uDim; c += gridDim.x * blockDim.x) {
        unsigned int run = 0u;
        unsigned int t = 0u;
        for (; t + 8u <= tiles; t += 8u) {
            unsigned int n[8];
#pragma unroll
            for (unsigned int k = 0; k < 8u; ++k) n[k] = counts[(size_t) (t + k) * uDim + c];
#pragma unroll
            for (unsigned int k = 0; k < 8u; ++k) {
                counts[(size_t) (t + k) * uDim + c] = run;
                run += n[k];
            }
        }
        for (; t < tiles; ++t) {
            const size_t k = (size_t) t * uDim + c;
            const unsigned int n = counts[k];
            counts[k] = run;
            run += n;
        }
        coltot[c] = run;
    }
}

// K4 (sorted sources).  LDS: segment masks [TR_BAND][TR_MASK_WORDS], row
// band ranges, run starts and cursors, the source row of each staged entry
// in input order, and the staging of K x 256 entries (source row, segment,
// value) in output order.  Each thread loads all of its entries of a batch
// (entry j = thread + 256 k) into registers at once, so one memory latency
// per batch is exposed, not one per row.
template<typename S, unsigned int K>
__global__ void __launch_bounds__(TR_SORT_THREADS) tr_band_kernel(unsigned int cDim, unsigned int uDim, const unsigned int *__restrict__ idx,
                                                                  const S *__restrict__ val, const unsigned int *__restrict__ tile_off,
                                                                  const unsigned int *__restrict__ out_ptr,
                                                                  const unsigned int *__restrict__ rbt, unsigned int bands,
                                                                  unsigned int *__restrict__ out_idx, S *__restrict__ out_val) {
    static_assert(TR_SORT_THREADS == TR_ROWS && TR_BAND % TR_SORT_THREADS == 0u, "one row / whole segments per thread");
    constexpr unsigned int CPT = TR_BAND / TR_SORT_THREADS;  // segments per thread in the run-start scan
    constexpr unsigned int cap = K * TR_SORT_THREADS;
    static_assert(cap >= TR_BAND, "a batch must hold a full row");
    extern __shared__ __align__(16) unsigned char lds[];
    unsigned int *masks = (unsigned int *) lds;                  // [TR_BAND * TR_MASK_WORDS]
    unsigned int *seg_b = masks + TR_BAND * TR_MASK_WORDS;       // [TR_ROWS]
    unsigned int *seg_o = seg_b + TR_ROWS;                       // [TR_ROWS + 1] chunk offsets of the rows
    unsigned int *lstart = seg_o + TR_ROWS + 1u;                 // [TR_BAND + 1]
    unsigned int *gcur = lstart + TR_BAND + 1u;                  // [TR_BAND]
    unsigned int *ctl = gcur + TR_BAND;                          // [4]
    unsigned short *s_in = (unsigned short *) (ctl + 4u);        // [cap] source row, input order
    unsigned short *s_row = s_in + cap;                          // [cap] source row, output order
    unsigned short *s_seg = s_row + cap;                         // [cap]
    S *s_val = (S *) align_up((size_t) (s_seg + cap), 16u);      // [cap]
    __shared__ unsigned int scan_smem[16];

    const unsigned int t = blockIdx.x / bands;
    const unsigned int band = blockIdx.x - t * bands;
    const unsigned int c0 = band * TR_BAND;
    const unsigned int nb = uDim - c0 < TR_BAND ? uDim - c0 : TR_BAND;
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int nr = cDim - r0 < TR_ROWS ? cDim - r0 : TR_ROWS;
    const unsigned int *rb_lo = rbt + ((size_t) t * (bands + 1u) + band) * TR_ROWS;
    const unsigned int *rb_hi = rb_lo + TR_ROWS;

    for (unsigned int w = threadIdx.x; w < TR_BAND * TR_MASK_WORDS; w += TR_SORT_THREADS) masks[w] = 0u;
    for (unsigned int c = threadIdx.x; c < nb; c += TR_SORT_THREADS)
        gcur[c] = out_ptr[c0 + c] + tile_off[(size_t) t * uDim + c0 + c];
    // row ranges inside the band and their offsets in the chunk
    {
        const unsigned int r = threadIdx.x;
        const unsigned int b = r < nr ? rb_lo[r] : 0u;
        const unsigned int n = r < nr ? rb_hi[r] - b : 0u;
        unsigned int total = 0u;
        const unsigned int o = block_exclusive_scan_u32(n, scan_smem, &total);
        seg_b[r] = b;
        seg_o[r] = o;
        if (r == 0u) seg_o[TR_ROWS] = total;
    }
    __syncthreads();

    unsigned int rb0 = 0u;
    while (rb0 < nr) {
        // batch: rows [rb0, rb1) with at most cap entries (a row has at most
        // TR_BAND <= cap entries in the band, so a batch holds >= 1 row)
        if (threadIdx.x == 0u) {
            unsigned int lo = rb0 + 1u, hi = nr;  // largest rb1 with seg_o[rb1] - seg_o[rb0] <= cap
            while (lo < hi) {
                const unsigned int mid = (lo + hi + 1u) >> 1;
                if (seg_o[mid] - seg_o[rb0] <= cap) lo = mid;
                else hi = mid - 1u;
            }
            ctl[0] = lo;
        }
        __syncthreads();
        const unsigned int rb1 = ctl[0];
        const unsigned int base = seg_o[rb0];
        const unsigned int n = seg_o[rb1] - base;
        // source row of every entry of the batch (LDS only; one thread per
        // row, its stores independent of each other)
        {
            const unsigned int r = threadIdx.x;  // TR_SORT_THREADS == TR_ROWS
            if (r >= rb0 && r < rb1) {
                const unsigned int o = seg_o[r] - base, len = seg_o[r + 1u] - seg_o[r];
                for (unsigned int k = 0; k < len; ++k) s_in[o + k] = (unsigned short) r;
            }
        }
        __syncthreads();
        // A: load this thread's entries (all loads issued before any use),
        // then mark which rows hold each segment
        unsigned int c[K], r[K];
        S v[K];
#pragma unroll
        for (unsigned int k = 0; k < K; ++k) {
            const unsigned int j = threadIdx.x + k * TR_SORT_THREADS;
            const unsigned int jj = j < n ? j : (n != 0u ? n - 1u : 0u);
            r[k] = n != 0u ? s_in[jj] : rb0;
            const unsigned int i = seg_b[r[k]] + (base + jj - seg_o[r[k]]);
            c[k] = n != 0u ? idx[i] : c0;
            v[k] = n != 0u ? val[i] : S(0);
        }
#pragma unroll
        for (unsigned int k = 0; k < K; ++k)
            if (threadIdx.x + k * TR_SORT_THREADS < n)
                atomicOr(&masks[(c[k] - c0) * TR_MASK_WORDS + (r[k] >> 5)], 1u << (r[k] & 31u));
        __syncthreads();
        // B: run lengths -> run starts in the staging
        {
            unsigned int cnt[CPT], sum = 0u;
#pragma unroll
            for (unsigned int k = 0; k < CPT; ++k) {
                const unsigned int cc = CPT * threadIdx.x + k;
                cnt[k] = 0u;
#pragma unroll
                for (unsigned int w = 0; w < TR_MASK_WORDS; ++w) cnt[k] += (unsigned int) __popc(masks[cc * TR_MASK_WORDS + w]);
                sum += cnt[k];
            }
            unsigned int total = 0u;
            unsigned int o = block_exclusive_scan_u32(sum, scan_smem, &total);
#pragma unroll
            for (unsigned int k = 0; k < CPT; ++k) {
                lstart[CPT * threadIdx.x + k] = o;
                o += cnt[k];
            }
            if (threadIdx.x == 0u) lstart[TR_BAND] = total;
        }
        __syncthreads();
        // C: every entry to its run start + rank among earlier rows
#pragma unroll
        for (unsigned int k = 0; k < K; ++k) {
            if (threadIdx.x + k * TR_SORT_THREADS >= n) continue;
            const unsigned int cl = c[k] - c0;
            const unsigned int *m = masks + cl * TR_MASK_WORDS;
            const unsigned int wr = r[k] >> 5;
            unsigned int rank = (unsigned int) __popc(m[wr] & ((1u << (r[k] & 31u)) - 1u));
            for (unsigned int w = 0; w < wr; ++w) rank += (unsigned int) __popc(m[w]);
            const unsigned int d = lstart[cl] + rank;
            s_row[d] = (unsigned short) r[k];
            s_seg[d] = (unsigned short) cl;
            s_val[d] = v[k];
        }
        __syncthreads();
        // D: runs to their final places (consecutive staging slots of one run
        // are consecutive output positions)
        for (unsigned int k = threadIdx.x; k < n; k += TR_SORT_THREADS) {
            const unsigned int cl = s_seg[k];
            const unsigned int pos = gcur[cl] + (k - lstart[cl]);
            out_idx[pos] = r0 + s_row[k];
            out_val[pos] = s_val[k];
        }
        __syncthreads();
        // E: advance the cursors, clear the masks
        for (unsigned int cc = threadIdx.x; cc < TR_BAND; cc += TR_SORT_THREADS) {
            if (cc < nb) gcur[cc] += lstart[cc + 1u] - lstart[cc];
#pragma unroll
            for (unsigned int w = 0; w < TR_MASK_WORDS; ++w) masks[cc * TR_MASK_WORDS + w] = 0u;
        }
        __syncthreads();
        rb0 = rb1;
    }
}

// K4u (fallback): one wave per (tile, band of TR_U_BAND segments); rows of
// the tile in order; LDS cursors; every row is scanned and filtered by band.
template<typename S>
__global__ void __launch_bounds__(64) tr_scatter_kernel(unsigned int cDim, unsigned int uDim, const unsigned int *__restrict__ ptr,
                                                        const unsigned int *__restrict__ idx, const S *__restrict__ val,
                                                        const unsigned int *__restrict__ tile_off, const unsigned int *__restrict__ out_ptr,
                                                        unsigned int ubands, unsigned int *__restrict__ out_idx, S *__restrict__ out_val) {
    __shared__ unsigned int cursor[TR_U_BAND];
    const unsigned int t = blockIdx.x / ubands;
    const unsigned int band = blockIdx.x - t * ubands;
    const unsigned int c0 = band * TR_U_BAND;
    const unsigned int c1 = c0 + TR_U_BAND < uDim ? c0 + TR_U_BAND : uDim;
    const unsigned int lane = lane_id();
    for (unsigned int k = lane; k < c1 - c0; k += 64u) cursor[k] = out_ptr[c0 + k] + tile_off[(size_t) t * uDim + c0 + k];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int r1 = r0 + TR_ROWS < cDim ? r0 + TR_ROWS : cDim;
    for (unsigned int r = r0; r < r1; ++r) {
        const unsigned int b = ptr[r], e = ptr[r + 1u];
        for (unsigned int i = b + lane; i < e; i += 64u) {
            const unsigned int c = idx[i];
            if (c < c0 || c >= c1) continue;
            const unsigned int pos = atomicAdd(&cursor[c - c0], 1u);
            out_idx[pos] = r;
            out_val[pos] = val[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}

size_t band_lds_bytes(unsigned int cap, size_t vb) {
    const size_t fixed = ((size_t) TR_BAND * TR_MASK_WORDS + TR_ROWS + (TR_ROWS + 1u) + (TR_BAND + 1u) + TR_BAND + 4u) * 4u;
    return align_up(fixed + (size_t) cap * 6u, 16u) + (size_t) cap * vb;
}

template<typename S>
int launch_scatter(const tr_layout &L, bool sorted, unsigned int cDim, unsigned int uDim, const unsigned int *ptr,
                   const unsigned int *idx, const void *val, const unsigned int *tile_off, const unsigned int *out_ptr,
                   const unsigned int *rbt, unsigned int *out_idx, void *out_val, hipStream_t s) {
    if (sorted) {
        const unsigned long long blocks = (unsigned long long) L.tiles * L.bands;
        CSAMD_REQUIRE(blocks <= 0x7fffffffull, "transpose: too many chunks");
        // staging of K x 256 entries per batch (CSAMD_TR_K = 8 / 12 / 16 for sweeps)
        const char *env = std::getenv("CSAMD_TR_K");
        const unsigned int want = env != nullptr ? (unsigned int) std::atoi(env) : 12u;
        const unsigned int K = want == 8u ? 8u : want == 16u ? 16u : 12u;
        const size_t lds = band_lds_bytes(K * TR_SORT_THREADS, sizeof(S));
        auto k = K == 8u ? tr_band_kernel<S, 8u> : K == 12u ? tr_band_kernel<S, 12u> : tr_band_kernel<S, 16u>;
        CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        k<<<(unsigned int) blocks, TR_SORT_THREADS, lds, s>>>(cDim, uDim, idx, (const S *) val, tile_off, out_ptr, rbt, L.bands,
                                                              out_idx, (S *) out_val);
        return hip_ok(hipGetLastError(), "tr_band_kernel");
    }
    const unsigned long long blocks = (unsigned long long) L.tiles * L.ubands;
    CSAMD_REQUIRE(blocks <= 0x7fffffffull, "transpose: too many tiles");
    tr_scatter_kernel<S><<<(unsigned int) blocks, 64, 0, s>>>(cDim, uDim, ptr, idx, (const S *) val, tile_off, out_ptr, L.ubands,
                                                                out_idx, (S *) out_val);
    return hip_ok(hipGetLastError(), "tr_scatter_kernel");
}

int transpose_impl(unsigned int cDim, unsigned int uDim, unsigned int nnz, const unsigned int *d_ptr, const unsigned int *d_idx,
                   const void *d_val, unsigned int value_type, unsigned int *d_out_ptr, unsigned int *d_out_idx, void *d_out_val,
                   void *d_scratch, size_t scratch_bytes, hipStream_t s) {
    CSAMD_REQUIRE(d_out_ptr != nullptr, "transpose: null output pointer array");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "transpose: unsupported value type %u", value_type);
    if (nnz == 0u || cDim == 0u) {
        return zero_bytes(d_out_ptr, (size_t) (uDim + 1u) * 4u, s);
    }
    CSAMD_REQUIRE(d_ptr != nullptr && d_idx != nullptr && d_val != nullptr && d_out_idx != nullptr && d_out_val != nullptr,
                  "transpose: null input");
    const tr_layout L = make_layout(cDim, uDim);
    CSAMD_REQUIRE(d_scratch != nullptr && scratch_bytes >= L.total, "transpose: scratch %zu < required %zu", scratch_bytes, L.total);
    char *base = (char *) d_scratch;
    unsigned int *flags = (unsigned int *) (base + L.off_flags);
    unsigned int *counts = (unsigned int *) (base + L.off_counts);
    unsigned int *rbt = L.band_table ? (unsigned int *) (base + L.off_rbt) : nullptr;
    unsigned int *coltot = (unsigned int *) (base + L.off_coltot);
    void *scan_tmp = base + L.off_scan;
    if (!zero_bytes(flags, 256u, s)) return 0;
    int max_lds = 65536;
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) (void) hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    }
    const int use_lds = (size_t) uDim * 4u <= (size_t) max_lds ? 1 : 0;
    if (!use_lds && !zero_bytes(counts, (size_t) L.tiles * uDim * 4u, s)) return 0;
    {
        const size_t lds = use_lds ? (size_t) uDim * 4u : 0u;
        auto k = tr_count_kernel;
        if (use_lds) CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        k<<<L.tiles, TR_COUNT_THREADS, lds, s>>>(cDim, uDim, d_ptr, d_idx, L.bands, use_lds, counts, rbt, flags);
        CSAMD_HIP(hipGetLastError(), "tr_count_kernel");
    }
    tr_tile_offsets_kernel<<<grid_for(uDim, 256u, 4096u), 256, 0, s>>>(counts, L.tiles, uDim, coltot);
    CSAMD_HIP(hipGetLastError(), "tr_tile_offsets_kernel");
    if (!exclusive_scan_u32(coltot, uDim, d_out_ptr, nullptr, scan_tmp, exclusive_scan_scratch_bytes(uDim), s)) return 0;
    unsigned int hflags[2] = {0u, 0u};
    CSAMD_HIP(hipMemcpyAsync(hflags, flags, 8u, hipMemcpyDeviceToHost, s), "D2H transpose flags");
    CSAMD_HIP(hipStreamSynchronize(s), "transpose sync");
    CSAMD_REQUIRE(hflags[1] == 0u, "transpose: minor index out of range (>= %u)", uDim);
    const bool sorted = hflags[0] == 0u && L.band_table;
    switch (vb) {
        case 1: return launch_scatter<uint8_t>(L, sorted, cDim, uDim, d_ptr, d_idx, d_val, counts, d_out_ptr, rbt, d_out_idx, d_out_val, s);
        case 2: return launch_scatter<uint16_t>(L, sorted, cDim, uDim, d_ptr, d_idx, d_val, counts, d_out_ptr, rbt, d_out_idx, d_out_val, s);
        default: return launch_scatter<float>(L, sorted, cDim, uDim, d_ptr, d_idx, d_val, counts, d_out_ptr, rbt, d_out_idx, d_out_val, s);
    }
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

size_t csamd_compressed_transpose_scratch_bytes(unsigned int cDim, unsigned int uDim, unsigned int nnz) {
    (void) nnz;
    return make_layout(cDim, uDim).total;
}

int csamd_build_compressed_transpose_raw(unsigned int cDim, unsigned int uDim, unsigned int nnz, const unsigned int *d_cAxPtr,
                                         const unsigned int *d_uAxIdx, const csamd_half *d_val, unsigned int *d_out_uAxPtr,
                                         unsigned int *d_heads, unsigned int *d_out_cAxIdx, csamd_half *d_out_val, void *d_scan_tmp,
                                         size_t scan_bytes, csamd_stream_t stream) {
    (void) d_heads;
    return transpose_impl(cDim, uDim, nnz, d_cAxPtr, d_uAxIdx, d_val, CSAMD_VALUE_F16, d_out_uAxPtr, d_out_cAxIdx, d_out_val,
                          d_scan_tmp, scan_bytes, (hipStream_t) stream);
}

int csamd_build_compressed_transpose_raw_typed(unsigned int cDim, unsigned int uDim, unsigned int nnz, const unsigned int *d_cAxPtr,
                                               const unsigned int *d_uAxIdx, const void *d_val, unsigned int value_type,
                                               unsigned int *d_out_uAxPtr, unsigned int *d_out_cAxIdx, void *d_out_val,
                                               void *d_scan_tmp, size_t scan_bytes, csamd_stream_t stream) {
    return transpose_impl(cDim, uDim, nnz, d_cAxPtr, d_uAxIdx, d_val, value_type, d_out_uAxPtr, d_out_cAxIdx, d_out_val, d_scan_tmp,
                          scan_bytes, (hipStream_t) stream);
}

} // extern "C"
