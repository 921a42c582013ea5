// Streaming reference-ABI masked grouped row reduction (compute_sparse_group_reduce
// over compressed rows / Sliced-ELL; src/runtime/mask_groups.cu:107-381 for
// the per-row semantics: a row takes part iff row_keep[r] != 0; a stored entry
// (c, v) counts iff c < cols, v != 0 and feature_keep[c] != 0; sum += v,
// detected += 1, max = max(0, v, ...), group b += v for each set bit b < G
// of feature_group_masks[c]; keep = sum >= min_total && detected >=
// min_detected && (group[fi] / sum, or 0 when sum == 0) <= max_fraction;
// percentages 100 * group / sum).
//
// gfx950 plan:
//   * one 1024-thread workgroup per CU over a contiguous row range with about
//     the same number of stored entries (gr_bounds_kernel);
//   * the range is taken in windows of 1,024 rows: the workgroup reads the
//     rows' keep flags and entry ranges once, writes the outputs of masked
//     and empty rows itself, and lists the rest in LDS; its 16 waves then take
//     rows from the list through an LDS counter, so no row's metadata is
//     waited for inside the streaming loop;
//   * a wave streams its rows in chunks of 512 entries (2 units of 4
//     entries per lane: one 16-byte index load and one value load each)
//     through two register buffers in turn, the next chunk (of this row or
//     the next) in flight while this one is consumed;
//   * feature test from LDS: per 32-column word {keep mask, kept before};
//   * with groups, the wave's kept entries of a chunk go to an LDS queue
//     (ballot + mbcnt), so that everything past the feature test scales with
//     the kept entries: one lane per queued entry finds its group mask (the
//     kept rank from the word's popcount, then the masks of the kept
//     features, staged in LDS at 1, 2 or 4 bytes each, the narrowest that
//     holds G bits, or read from memory when they do not fit) and rewrites
//     the queue with the entries that belong to some group; then P =
//     min(2^ceil(log2 G), 16) lanes take each of those together, lane l
//     adding it to group l % P (and l % P + 16 when G > 16) when its bit is
//     set, 128 / P entries per step; one xor-shuffle tree per row folds the
//     copies;
//   * row sums, counts and maxima per lane, folded by wave reductions when the
//     row's last chunk is consumed; the outputs are written then.
// Rows whose last unit would pass the array's last whole 4-entry unit (at
// most three per matrix) are read entry by entry.
#include <mutex>

#include "group_reduce.hh"
#include "stream.hh"

namespace csamd {
namespace {

constexpr unsigned int GR_THREADS = 1024u;
constexpr unsigned int GR_WAVES = GR_THREADS / 64u;
constexpr unsigned int GR_UNITS = 2u;                  // units per lane per chunk
constexpr unsigned int GR_CHUNK = 64u * GR_UNITS;      // units per chunk
constexpr unsigned int GR_WINDOW = 1024u;              // rows per window
constexpr unsigned int GR_RPT = GR_WINDOW / GR_THREADS;  // window rows per thread
constexpr unsigned int GR_QCAP = 64u * 4u * GR_UNITS; // queued entries per wave (one chunk)
constexpr unsigned int GR_QPAD = 128u;                 // zero entries after the queued ones (one loop step)
constexpr unsigned int GR_MAX_WORDS = 6144u;           // feature map words (196,608 columns)
constexpr unsigned int GR_MAX_GRID = 1024u;
constexpr unsigned int GR_TAIL = 4u;                   // rows read entry by entry, per window
constexpr unsigned int GR_ROW_COST = 384u;             // per-row cost in entries, for the workgroup ranges
constexpr size_t GR_LDS_BUDGET = 160u * 1024u - 1024u;

struct gr_params {
    gr_source src;
    const unsigned char *row_keep;
    const uint2 *fmap;        // [words + 1]: {keep mask, kept before}; last {0, F}
    unsigned int words;
    const uint32_t *ktab;     // [F]: group mask of each kept feature by rank
    const uint32_t *fgm;      // [cols]: group masks by column (when F > cap)
    uint32_t gbits;           // bits of the call's groups
    unsigned int cap;         // kept features whose masks fit in the LDS table
    unsigned int G;
    unsigned int logP;        // P = 2^logP lanes per queued entry: P >= G, or 16 (two groups per lane) for G > 16
    csamd_sparse_group_filter_params filter;
    gr_outputs o;
    const unsigned int *bounds;  // [gridDim.x + 1] (compressed rows / Sliced-ELL)
    // Blocked-ELL: block columns [row_blocks x width], block size 2^lbs, values [rows x ell_cols]
    const unsigned int *bcol;
    unsigned int lbs, ell_cols, width;
    unsigned int nrb;  // row blocks per window (block columns staged in LDS)
};

// Entry range [b, e) of row r: compressed rows, or the row's slots of its slice.
__device__ __forceinline__ void gr_range(const gr_source &s, unsigned int r, unsigned int &b, unsigned int &e) {
    if (s.ptr != nullptr) {
        b = gmem(s.ptr)[r];
        e = gmem(s.ptr)[r + 1u];
        return;
    }
    if (s.slice_count == 0u) {
        b = e = 0u;
        return;
    }
    unsigned int sl;
    if (s.slice_rows != 0u) {
        sl = r / s.slice_rows;
        if (sl >= s.slice_count) sl = s.slice_count - 1u;
    } else {  // last slice whose first row is <= r
        unsigned int lo = 0u, hi = s.slice_count - 1u;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (gmem(s.slice_offsets)[mid] <= r) lo = mid;
            else hi = mid - 1u;
        }
        sl = lo;
    }
    const unsigned int w = gmem(s.slice_widths)[sl];
    b = gmem(s.slot_offsets)[sl] + (r - gmem(s.slice_offsets)[sl]) * w;
    e = b + w;
}

// Per 32-column word: keep mask (bit k: column 32 w + k < cols is kept) and
// the number of kept columns before the word; a {0, F} sentinel closes it.
__global__ void __launch_bounds__(1024) gr_fmap_kernel(const unsigned char *__restrict__ fk, unsigned int cols, unsigned int words,
                                                       uint2 *__restrict__ fmap) {
    __shared__ unsigned int smem[16];
    unsigned int carry = 0u;
    for (unsigned int w0 = 0; w0 < words; w0 += blockDim.x) {
        const unsigned int w = w0 + threadIdx.x;
        unsigned int mask = 0u;
        if (w < words) {
            const unsigned int c0 = w * 32u;
            const unsigned int n = cols - c0 < 32u ? cols - c0 : 32u;
            if (fk == nullptr) {
                mask = n == 32u ? 0xffffffffu : (1u << n) - 1u;
            } else if (n == 32u && ((uintptr_t) (fk + c0) & 3u) == 0u) {
                const unsigned int *f4 = (const unsigned int *) (fk + c0);
                unsigned int q[8];
#pragma unroll
                for (unsigned int k = 0; k < 8u; ++k) q[k] = f4[k];
#pragma unroll
                for (unsigned int k = 0; k < 8u; ++k)
#pragma unroll
                    for (unsigned int j = 0; j < 4u; ++j) mask |= (((q[k] >> (8u * j)) & 0xffu) != 0u ? 1u : 0u) << (4u * k + j);
            } else {
                for (unsigned int k = 0; k < n; ++k) mask |= (fk[c0 + k] != 0u ? 1u : 0u) << k;
            }
        }
        unsigned int total = 0u;
        const unsigned int ex = block_exclusive_scan_u32((unsigned int) __popc(mask), smem, &total);
        if (w < words) fmap[w] = make_uint2(mask, carry + ex);
        carry += total;
    }
    if (threadIdx.x == 0u) fmap[words] = make_uint2(0u, carry);
}

// Group mask of every kept feature, by kept rank.
__global__ void gr_ktab_kernel(const uint2 *__restrict__ fmap, const uint32_t *__restrict__ fgm, unsigned int cols, uint32_t gbits,
                               uint32_t *__restrict__ ktab) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
        const uint2 pm = fmap[c >> 5];
        const unsigned int bit = c & 31u;
        if (((pm.x >> bit) & 1u) != 0u) ktab[pm.y + (unsigned int) __popc(pm.x & ((1u << bit) - 1u))] = fgm[c] & gbits;
    }
}

// Workgroup row ranges of about the same cost, a row costing its entries
// plus GR_ROW_COST (its share of a partly filled chunk, the row's window
// bookkeeping and its folds and writes), so that ranges of short rows are
// not overloaded.  With position p(r) = start(r) - B0 + r GR_ROW_COST (B0:
// the first row's start; strictly ascending when row starts ascend) and W =
// E - B0 + rows GR_ROW_COST (E: the last row's end), bounds[k] = the first
// row whose position is at or after T_k = floor(k W / grid): row r + 1 when
// T_k lies in (p(r), p(r + 1)].  The host fills bounds with `rows` first, so
// a target no row takes (row starts that do not ascend) leaves an empty
// range behind, and every bound stays within [0, rows].
__global__ void gr_bounds_kernel(gr_source s, unsigned int grid, unsigned int *__restrict__ bounds) {
    __shared__ unsigned int sB0, sE;
    if (threadIdx.x == 0u) {
        unsigned int b, e;
        gr_range(s, 0u, b, e);
        sB0 = b;
        gr_range(s, s.rows - 1u, b, e);
        sE = e;
    }
    __syncthreads();
    const unsigned long long B0 = sB0;
    const unsigned long long W = (sE > sB0 ? (unsigned long long) sE - sB0 : 0ull) + (unsigned long long) s.rows * GR_ROW_COST;
    const unsigned int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0u) {
        bounds[0] = 0u;
        bounds[grid] = s.rows;
        // targets at position 0 (W < grid) start at row 0
        for (unsigned int k = 1u; k < grid && (unsigned long long) k * W < grid; ++k) bounds[k] = 0u;
    }
    if (r >= s.rows) return;
    unsigned int b, e, b1 = sE;
    gr_range(s, r, b, e);
    if (r + 1u < s.rows) gr_range(s, r + 1u, b1, e);
    if (b1 < b || b < B0) return;
    const unsigned long long p0 = (unsigned long long) b - B0 + (unsigned long long) r * GR_ROW_COST;
    const unsigned long long p1 = (unsigned long long) b1 - B0 + (unsigned long long) (r + 1u) * GR_ROW_COST;
    // every k with p0 < T_k <= p1 starts at row r + 1 (T_0 = 0 is row 0's)
    unsigned int k = (unsigned int) ((p0 * grid) / W);
    if (k < 1u) k = 1u;
    for (; k < grid; ++k) {
        const unsigned long long T = (unsigned long long) k * W / grid;
        if (T > p1) break;
        if (T > p0) bounds[k] = r + 1u;
    }
}

__global__ void gr_fill_kernel(unsigned int *p, unsigned int n, unsigned int v) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// LDS tables of one workgroup.
struct gr_lds {
    const uint2 *fmap;
    const void *ktab;  // kept-feature group masks, TB bytes each
    uint2 *queue;      // this wave's queue [GR_QCAP + GR_QPAD]
};

struct gr_acc {
    float sum;
    float vmax;
    unsigned int det;
    float g0;  // group lane % P (over the entries this lane took)
    float g1;  // group lane % P + 16 (G > 16)
};

// Output pointers and filter, kept in LDS (read where a row ends; the
// streaming loop does not hold them in registers).
struct gr_sink {
    gr_outputs o;
    csamd_sparse_group_filter_params filter;
};

// One row's outputs from values already folded over the wave.
__device__ __forceinline__ void gr_write_row(const gr_params &a, const gr_sink &k, unsigned int row, float rs, float mx, unsigned int det,
                                             float g0, float g1, float gsel, bool kept_row) {
    const unsigned int lane = lane_id();
    const gr_outputs &o = k.o;
    if (lane == 0u) {
        gmem(o.row_totals)[row] = rs;
        gmem(o.max_values)[row] = mx;
        gmem(o.detected_features)[row] = det;
        const float frac = rs > 0.0f ? gsel / rs : 0.0f;
        gmem(o.row_keep)[row] = (unsigned char) (kept_row && rs >= k.filter.min_total && det >= k.filter.min_detected_features &&
                                                 frac <= k.filter.max_group_fraction);
    }
    if (o.group_counts != nullptr) {
        const unsigned long long base = (unsigned long long) row * a.G;
        const unsigned int P = 1u << a.logP;
        if (lane < P && lane < a.G) {
            gmem(o.group_counts)[base + lane] = g0;
            gmem(o.group_percentages)[base + lane] = rs > 0.0f ? 100.0f * g0 / rs : 0.0f;
        }
        if (lane < P && lane + 16u < a.G) {
            gmem(o.group_counts)[base + lane + 16u] = g1;
            gmem(o.group_percentages)[base + lane + 16u] = rs > 0.0f ? 100.0f * g1 / rs : 0.0f;
        }
    }
}

// Thread-level outputs of a row with no entry to read: masked out (keep 0,
// emit_empty_row) or kept with no stored entry (keep from the filter).
__device__ __forceinline__ void gr_write_empty(const gr_params &a, const gr_sink &k, unsigned int row, bool kept_row) {
    const gr_outputs &o = k.o;
    gmem(o.row_totals)[row] = 0.0f;
    gmem(o.max_values)[row] = 0.0f;
    gmem(o.detected_features)[row] = 0u;
    gmem(o.row_keep)[row] = (unsigned char) (kept_row && 0.0f >= k.filter.min_total && 0u >= k.filter.min_detected_features &&
                                             0.0f <= k.filter.max_group_fraction);
    if (o.group_counts != nullptr) {
        const unsigned long long base = (unsigned long long) row * a.G;
        for (unsigned int g = 0; g < a.G; ++g) {
            gmem(o.group_counts)[base + g] = 0.0f;
            gmem(o.group_percentages)[base + g] = 0.0f;
        }
    }
}

// Group mask of the kept feature of rank `rank` from the LDS table of TB-byte
// entries (TB = 1, 2 or 4: the narrowest that holds G bits; rank < F <= cap).
template<unsigned int TB> __device__ __forceinline__ uint32_t gr_tab(const void *t, unsigned int rank) {
    if constexpr (TB == 1u) return ((const uint8_t *) t)[rank];
    else if constexpr (TB == 2u) return ((const uint16_t *) t)[rank];
    else return ((const uint32_t *) t)[rank];
}

// v where bit b of m is set, else +0 (b < 32): a one-bit signed field
// extract gives 0 or all ones.
__device__ __forceinline__ float gr_bit(uint32_t m, unsigned int b, float v) {
    return __uint_as_float(__float_as_uint(v) & (uint32_t) __builtin_amdgcn_sbfe((int) m, b, 1u));
}

// K entries per lane, kept[k] when entry k's column is kept and its value is
// not zero (c[k] < cols then).
//   * G = 0: the kept entries add to the lane's row partials directly.
//   * G > 0: the wave's kept entries go to its LDS queue as {value, column}
//     (ballot + mbcnt); pass 1 takes them one per lane, adds them to the row
//     partials, looks up their group masks (TB = 1, 2, 4: the LDS table by
//     kept rank; TB = 0: the column masks in memory) and rewrites the queue in
//     place as {value, mask} of the entries that belong to some group; pass 2
//     adds those to the group partials, P lanes per entry, lane l taking
//     groups l % P and, when G > 16, l % P + 16.
// The lookups and group work thus scale with the kept entries, not the
// stored ones.
template<unsigned int K, unsigned int TB>
__device__ __forceinline__ void gr_accum(const gr_params &a, const gr_lds &L, const unsigned int (&c)[K], const float (&v)[K],
                                         const bool (&kept)[K], gr_acc &A) {
    if (a.G == 0u) {
#pragma unroll
        for (unsigned int k = 0; k < K; ++k) {
            A.sum += kept[k] ? v[k] : 0.0f;
            A.det += kept[k] ? 1u : 0u;
            A.vmax = kept[k] && v[k] > A.vmax ? v[k] : A.vmax;  // fmaxf over kept values: a NaN never replaces the max
        }
        return;
    }
    const unsigned int lane = lane_id();
    unsigned int n = 0u;
#pragma unroll
    for (unsigned int k = 0; k < K; ++k) {
        const unsigned long long bal = __ballot(kept[k]);
        if (kept[k]) {
            const unsigned int pos = __builtin_amdgcn_mbcnt_hi((unsigned int) (bal >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned int) bal, n));
            L.queue[pos] = make_uint2(__float_as_uint(v[k]), c[k]);
        }
        n += (unsigned int) __popcll(bal);
    }
    if (n == 0u) return;
    A.det += lane == 0u ? n : 0u;
    __builtin_amdgcn_wave_barrier();
    // pass 1: entry i0 + lane; the rewrite goes to a position <= the one read
    unsigned int nq = 0u;
    for (unsigned int i0 = 0; i0 < n; i0 += 64u) {
        const unsigned int i = i0 + lane;
        uint32_t m = 0u;
        float x = 0.0f;
        if (i < n) {
            const uint2 e = L.queue[i];
            x = __uint_as_float(e.x);
            const unsigned int col = e.y;  // kept, so col < cols
            A.sum += x;
            A.vmax = x > A.vmax ? x : A.vmax;
            if constexpr (TB == 0u) {
                m = gmem(a.fgm)[col] & a.gbits;
            } else {
                const uint2 pm = L.fmap[col >> 5];
                m = gr_tab<TB>(L.ktab, pm.y + (unsigned int) __popc(pm.x & ((1u << (col & 31u)) - 1u)));
            }
        }
        const unsigned long long bal = __ballot(m != 0u);
        if (m != 0u) {
            const unsigned int pos = __builtin_amdgcn_mbcnt_hi((unsigned int) (bal >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned int) bal, nq));
            L.queue[pos] = make_uint2(__float_as_uint(x), m);
        }
        nq += (unsigned int) __popcll(bal);
    }
    if (nq != 0u) {
        // zeros after the queued entries, so that a pass-2 step needs no bounds test
        L.queue[nq + lane] = make_uint2(0u, 0u);
        L.queue[nq + 64u + lane] = make_uint2(0u, 0u);
        __builtin_amdgcn_wave_barrier();
        // pass 2: lane l takes entries 2 s and 2 s + 1 of each step of 128 / P, s = l / P
        const unsigned int slot = lane >> a.logP, grp = lane & ((1u << a.logP) - 1u), step = 128u >> a.logP;
        const uint4 *q4 = (const uint4 *) L.queue;
        for (unsigned int i0 = 0; i0 < nq; i0 += step) {
            const uint4 q = q4[(i0 >> 1) + slot];
            const float v0 = __uint_as_float(q.x), v1 = __uint_as_float(q.z);
            A.g0 += gr_bit(q.y, grp, v0) + gr_bit(q.w, grp, v1);
            if (a.G > 16u) A.g1 += gr_bit(q.y, grp + 16u, v0) + gr_bit(q.w, grp + 16u, v1);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// K entries per lane with their columns: the feature test from the LDS
// feature map, then gr_accum.
template<unsigned int K, unsigned int TB>
__device__ __forceinline__ void gr_take(const gr_params &a, const gr_lds &L, const unsigned int (&c)[K], const float (&v)[K],
                                        gr_acc &A) {
    unsigned int pw[K];
#pragma unroll
    for (unsigned int k = 0; k < K; ++k) pw[k] = L.fmap[min(c[k] >> 5, a.words)].x;
    bool kept[K];
#pragma unroll
    for (unsigned int k = 0; k < K; ++k) kept[k] = __builtin_amdgcn_ubfe(pw[k], c[k] & 31u, 1u) != 0u && v[k] != 0.0f;
    gr_accum<K, TB>(a, L, c, v, kept, A);
}

__device__ __forceinline__ void gr_finish(const gr_params &a, const gr_sink &k, unsigned int row, gr_acc &A) {
    const float rs = wave_sum_f32(A.sum);
    const unsigned int det = wave_sum_u32(A.det);
    const float mx = wave_max_f32(A.vmax);
    float g0 = A.g0, g1 = A.g1;
    if (a.G != 0u)
        for (unsigned int o = 1u << a.logP; o < 64u; o <<= 1) {
            g0 += __shfl_xor(g0, (int) o, 64);
            if (a.G > 16u) g1 += __shfl_xor(g1, (int) o, 64);
        }
    // lane l < P now holds group l (g0) and group l + 16 (g1)
    const unsigned int fi = k.filter.fraction_group_index;
    const float gsel = fi >= a.G ? 0.0f : fi < 16u ? __shfl(g0, (int) fi, 64) : __shfl(g1, (int) (fi - 16u), 64);
    gr_write_row(a, k, row, rs, mx, det, g0, g1, gsel, true);
    A.sum = 0.0f;
    A.vmax = 0.0f;
    A.det = 0u;
    A.g0 = 0.0f;
    A.g1 = 0.0f;
}

template<int VT> struct gr_chunk {
    u32x4 c[GR_UNITS];
    typename unit_vals<VT>::raw v[GR_UNITS];
};

struct gr_cursor {
    unsigned int row, b, e;  // the row and its entry range
    unsigned int u, ue;      // first unit of the chunk, end unit of the row
};

template<int VT>
__device__ __forceinline__ void gr_load(const gr_params &a, const gr_cursor &c, gr_chunk<VT> &R) {
    using S = typename value_traits<VT>::storage;
    const unsigned int lane = lane_id();
    const S *val = (const S *) a.src.val;
#pragma unroll
    for (unsigned int j = 0; j < GR_UNITS; ++j) {
        const unsigned int uu = min(c.u + lane + 64u * j, c.ue - 1u);
        R.c[j] = __builtin_nontemporal_load(gmem((const u32x4 *) (a.src.idx + (size_t) uu * 4u)));
        R.v[j] = unit_vals<VT>::vload(val, uu * 4u);
    }
}

// Consume one chunk of row [b, e) starting at unit u.  Only a row's first and
// last chunk hold entries outside the row; those chunks (a wave-uniform test)
// mask them.
template<int VT, unsigned int TB>
__device__ __forceinline__ void gr_consume(const gr_params &a, const gr_lds &L, const gr_cursor &c,
                                           const gr_chunk<VT> &R, gr_acc &A) {
    const unsigned long long first = 4ull * c.u;
    const bool edge = first < c.b || first + 4ull * GR_CHUNK > c.e;
    const unsigned int lane = lane_id();
    unsigned int col[4u * GR_UNITS];
    float v[4u * GR_UNITS];
#pragma unroll
    for (unsigned int j = 0; j < GR_UNITS; ++j)
#pragma unroll
        for (unsigned int k = 0; k < 4u; ++k) {
            col[4u * j + k] = lane_of(R.c[j], k);
            v[4u * j + k] = unit_vals<VT>::get(R.v[j], k);
        }
    if (edge) {
        // entry i = 4 u + k < 2^32 (u < ue <= 2^30) lies in [b, e) iff
        // i - b < e - b in wrapping 32-bit arithmetic
        const unsigned int len = c.e - c.b;
#pragma unroll
        for (unsigned int j = 0; j < GR_UNITS; ++j) {
            const unsigned int t = 4u * (c.u + lane + 64u * j) - c.b;
#pragma unroll
            for (unsigned int k = 0; k < 4u; ++k)
                if (t + k >= len) col[4u * j + k] = CSAMD_INVALID_COL;
        }
    }
    gr_take<4u * GR_UNITS, TB>(a, L, col, v, A);
}

// LDS of the workgroup beyond the feature map and the table.
struct gr_shared {
    unsigned int *wrow, *wb, *we;  // window rows listed for streaming, entry ranges
};

// The workgroup's rows [lo, hi), window by window (TB as in gr_take).
template<int VT, unsigned int TB>
__device__ __forceinline__ void gr_stream_rows(const gr_params &a, const gr_lds &L, const gr_sink &sink, unsigned int *wrow,
                                               unsigned int lo, unsigned int hi) {
    using S = typename value_traits<VT>::storage;
    __shared__ unsigned int s_next, s_nk, s_ntail, s_scan[16];
    __shared__ unsigned int s_tail[GR_TAIL][3];
    const unsigned int tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
    unsigned int *wb = wrow + GR_WINDOW, *we = wb + GR_WINDOW;
    unsigned int Eb, Ee;
    gr_range(a.src, a.src.rows - 1u, Eb, Ee);
    const unsigned int e_main = Ee & ~3u;  // entries of whole units
    gr_acc A{0.0f, 0.0f, 0u, 0.0f, 0.0f};

    for (unsigned int w0 = lo; w0 < hi; w0 += GR_WINDOW) {
        const unsigned int w1 = hi - w0 < GR_WINDOW ? hi : w0 + GR_WINDOW;
        if (tid == 0u) s_ntail = 0u;
        __syncthreads();  // tables staged / the previous window consumed
        // --- window: keep flags and ranges once; outputs of rows with nothing to read
        unsigned int rr[GR_RPT], rb[GR_RPT], re[GR_RPT], take[GR_RPT];
        unsigned int ntake = 0u;
#pragma unroll
        for (unsigned int h = 0; h < GR_RPT; ++h) {
            const unsigned int r = w0 + GR_RPT * tid + h;
            take[h] = 0u;
            rr[h] = r;
            rb[h] = re[h] = 0u;
            if (r < w1) {
                const bool kept = a.row_keep == nullptr || gmem(a.row_keep)[r] != 0u;
                gr_range(a.src, r, rb[h], re[h]);
                if (!kept || rb[h] == re[h]) {
                    gr_write_empty(a, sink, r, kept);
                } else if (re[h] > e_main) {
                    const unsigned int t = atomicAdd(&s_ntail, 1u);
                    if (t < GR_TAIL) {
                        s_tail[t][0] = r;
                        s_tail[t][1] = rb[h];
                        s_tail[t][2] = re[h];
                    }
                } else {
                    take[h] = 1u;
                }
            }
            ntake += take[h];
        }
        unsigned int nk = 0u;
        const unsigned int pos = block_exclusive_scan_u32(ntake, s_scan, &nk);
        unsigned int p = pos;
#pragma unroll
        for (unsigned int h = 0; h < GR_RPT; ++h)
            if (take[h]) {
                wrow[p] = rr[h];
                wb[p] = rb[h];
                we[p] = re[h];
                ++p;
            }
        if (tid == 0u) {
            s_next = 0u;
            s_nk = nk;
        }
        __syncthreads();
        const unsigned int nt = __builtin_amdgcn_readfirstlane(min(s_ntail, GR_TAIL));  // read here: the next window resets it
        // --- stream the listed rows
        {
            const unsigned int n_rows = __builtin_amdgcn_readfirstlane(s_nk);
            auto next_row = [&](gr_cursor &c) -> bool {
                unsigned int i = 0u;
                if (lane == 0u) i = atomicAdd(&s_next, 1u);
                i = __builtin_amdgcn_readfirstlane(i);
                if (i >= n_rows) return false;
                c.row = __builtin_amdgcn_readfirstlane(wrow[i]);  // LDS reads of one address: wave-uniform
                c.b = __builtin_amdgcn_readfirstlane(wb[i]);
                c.e = __builtin_amdgcn_readfirstlane(we[i]);
                c.u = c.b >> 2;
                c.ue = (c.e + 3u) >> 2;
                return true;
            };
            auto advance = [&](gr_cursor &c) -> bool {
                if (c.u + GR_CHUNK < c.ue) {
                    c.u += GR_CHUNK;
                    return true;
                }
                return next_row(c);
            };
            gr_cursor ca, cb;
            if (next_row(ca)) {
                // loads are issued unconditionally (after the last chunk the
                // wave re-loads it, in bounds and unused): a load under a
                // branch makes the compiler wait for it
                gr_chunk<VT> RA, RB;
                gr_load<VT>(a, ca, RA);
                while (true) {
                    cb = ca;
                    bool more = advance(cb);
                    if (!more) cb = ca;
                    gr_load<VT>(a, cb, RB);
                    gr_consume<VT, TB>(a, L, ca, RA, A);
                    if (ca.u + GR_CHUNK >= ca.ue) gr_finish(a, sink, ca.row, A);
                    if (!more) break;
                    ca = cb;
                    more = advance(ca);
                    if (!more) ca = cb;
                    gr_load<VT>(a, ca, RA);
                    gr_consume<VT, TB>(a, L, cb, RB, A);
                    if (cb.u + GR_CHUNK >= cb.ue) gr_finish(a, sink, cb.row, A);
                    if (!more) break;
                }
            }
        }
        __syncthreads();
        // --- rows ending in the array's last partial unit: entry by entry
        if (wave < nt) {
            const unsigned int row = __builtin_amdgcn_readfirstlane(s_tail[wave][0]);
            const unsigned int b = __builtin_amdgcn_readfirstlane(s_tail[wave][1]);
            const unsigned int e = __builtin_amdgcn_readfirstlane(s_tail[wave][2]);
            const S *val = (const S *) a.src.val;
            for (unsigned int i0 = b; i0 < e; i0 += 64u) {
                const unsigned int i = i0 + lane;
                unsigned int col[1] = {i < e ? gmem(a.src.idx)[i] : CSAMD_INVALID_COL};
                float v[1] = {i < e ? value_traits<VT>::decode(gmem(val)[i]) : 0.0f};
                gr_take<1u, TB>(a, L, col, v, A);
            }
            gr_finish(a, sink, row, A);
        }
    }
}

// LDS of one workgroup: the feature map, the kept features' group masks when
// TB != 0 and they fit (F <= cap), the waves' queues, the window lists.
struct gr_staged {
    gr_lds L;
    unsigned int *wrow;
    bool in_lds;
};

template<unsigned int TB>
__device__ __forceinline__ gr_staged gr_stage(const gr_params &a, unsigned char *lds, gr_sink &sink) {
    const unsigned int tid = threadIdx.x, wave = tid >> 6;
    uint2 *fm = (uint2 *) lds;
    unsigned char *kt = lds + align_up((size_t) (a.words + 1u) * 8u, 16u);
    uint2 *queues = (uint2 *) (kt + align_up((size_t) a.cap * (TB != 0u ? TB : 1u), 16u));
    unsigned int *wrow = (unsigned int *) (queues + GR_WAVES * (GR_QCAP + GR_QPAD));
    if (tid == 0u) {
        sink.o = a.o;
        sink.filter = a.filter;
    }
    for (unsigned int w = tid; w <= a.words; w += GR_THREADS) fm[w] = a.fmap[w];
    const unsigned int F = __builtin_amdgcn_readfirstlane(gmem((const unsigned int *) a.fmap)[2u * a.words + 1u]);
    const bool in_lds = TB != 0u && F <= a.cap;
    for (unsigned int k = tid; k < F && in_lds; k += GR_THREADS) {
        const uint32_t m = gmem(a.ktab)[k];
        if (TB == 1u) ((uint8_t *) kt)[k] = (uint8_t) m;
        else if (TB == 2u) ((uint16_t *) kt)[k] = (uint16_t) m;
        else ((uint32_t *) kt)[k] = m;
    }
    return gr_staged{gr_lds{fm, kt, queues + wave * (GR_QCAP + GR_QPAD)}, wrow, in_lds};
}

// TB = 0: group masks from memory (or no groups); TB = 1, 2, 4: the kept
// features' masks staged in LDS at TB bytes each when they fit (F <= cap),
// else from memory.
template<int VT, unsigned int TB>
__global__ void __launch_bounds__(GR_THREADS) gr_stream_kernel(gr_params a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ gr_sink sink;
    const unsigned int lo = min(a.bounds[blockIdx.x], a.src.rows), hi = min(a.bounds[blockIdx.x + 1u], a.src.rows);
    if (lo >= hi) return;  // uniform across the workgroup
    const gr_staged st = gr_stage<TB>(a, lds, sink);
    if constexpr (TB == 0u) {
        gr_stream_rows<VT, 0u>(a, st.L, sink, st.wrow, lo, hi);
    } else {
        if (st.in_lds) gr_stream_rows<VT, TB>(a, st.L, sink, st.wrow, lo, hi);
        else gr_stream_rows<VT, 0u>(a, st.L, sink, st.wrow, lo, hi);
    }
}

// ---------------------------------------------------------------------------
// Blocked-ELL: a lane takes 8 consecutive values of a row per chunk (one
// block holds them: block size 8, 16 or 32), one 16-byte load for 2-byte
// values.  The row block's block column of the lane's slot is loaded one chunk
// ahead of the values; its 8 keep bits come from the LDS feature map, and the
// value load is issued only when one of them is set, so slots of padding
// (CSAMD_INVALID_COL) and blocks of no kept feature are not read.
// ---------------------------------------------------------------------------
template<int VT> struct oct_vals;
// Raw buffer loads (non-temporal): an offset at or past the resource's
// num_records reads zeros and touches no memory.
__device__ __forceinline__ u32x4 gb_load16(__amdgpu_buffer_rsrc_t r, unsigned int off) {
    return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int) off, 0, 2));
}
__device__ __forceinline__ u32x2 gb_load8(__amdgpu_buffer_rsrc_t r, unsigned int off) {
    return __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(r, (int) off, 0, 2));
}

template<> struct oct_vals<CSAMD_VALUE_F16> {
    using raw = u32x4;
    static constexpr unsigned int bytes = 2u;
    static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned int off) { return gb_load16(r, off); }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        const unsigned int w = lane_of(r, k >> 1);
        return f16_bits_to_float((uint16_t) ((k & 1u) ? (w >> 16) : (w & 0xffffu)));
    }
};
template<> struct oct_vals<CSAMD_VALUE_BF16> {
    using raw = u32x4;
    static constexpr unsigned int bytes = 2u;
    static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned int off) { return gb_load16(r, off); }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        const unsigned int w = lane_of(r, k >> 1);
        return __uint_as_float((k & 1u) ? (w & 0xffff0000u) : (w << 16));
    }
};
template<> struct oct_vals<CSAMD_VALUE_F32> {
    struct raw {
        u32x4 a, b;
    };
    static constexpr unsigned int bytes = 4u;
    static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned int off) {
        return raw{gb_load16(r, off), gb_load16(r, off + 16u)};
    }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        return __uint_as_float(k < 4u ? lane_of(r.a, k) : lane_of(r.b, k - 4u));
    }
};
template<> struct oct_vals<CSAMD_VALUE_FP8_E4M3> {
    using raw = u32x2;
    static constexpr unsigned int bytes = 1u;
    static __device__ __forceinline__ raw load(__amdgpu_buffer_rsrc_t r, unsigned int off) { return gb_load8(r, off); }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        return fp8_e4m3_decode_hw(k < 4u ? r.x : r.y, (int) (k & 3u));
    }
};

constexpr unsigned int GB_CHUNK = 64u * 8u;  // entries per wave chunk

struct gb_cursor {
    unsigned int row;
    unsigned int e;  // first entry of the chunk within the row
    bool live;
};

// One chunk's lane state: keep bits of the lane's 8 columns, the first
// column, the values.
template<int VT> struct gb_lane {
    unsigned int kw;
    unsigned int c0;
    typename oct_vals<VT>::raw v;
};

// The workgroup's rows [lo, hi) in windows of a.nrb row blocks: the windows'
// block columns are staged in LDS (bcl), so a chunk's value loads depend on
// LDS reads only and stay in flight while the previous chunk is consumed.
template<int VT, unsigned int TB>
__device__ __forceinline__ void gr_bell_rows(const gr_params &a, const gr_lds &L, const gr_sink &sink, unsigned int *wrow,
                                             unsigned int *bcl, unsigned int lo, unsigned int hi) {
    __shared__ unsigned int s_next, s_nk, s_scan[16];
    using O = oct_vals<VT>;
    const unsigned int tid = threadIdx.x, lane = lane_id();
    const unsigned int lbs = a.lbs, bs = 1u << lbs, ell = a.ell_cols, width = a.width;
    const unsigned int nblocks = (a.src.cols + bs - 1u) >> lbs;  // CSAMD_INVALID_COL is past it
    const unsigned int row_bytes = ell * O::bytes;
    gr_acc A{0.0f, 0.0f, 0u, 0.0f, 0.0f};

    for (unsigned int rb0 = lo >> lbs, rb_end = ((hi - 1u) >> lbs) + 1u; rb0 < rb_end; rb0 += a.nrb) {
        const unsigned int rb1 = rb_end - rb0 < a.nrb ? rb_end : rb0 + a.nrb;
        const unsigned int r0 = max(lo, rb0 << lbs), r1 = min(hi, rb1 << lbs);  // r1 - r0 <= GR_THREADS
        __syncthreads();  // tables staged / the previous window consumed
        const unsigned int nbc = (rb1 - rb0) * width;
        for (unsigned int i = tid; i < nbc; i += GR_THREADS) bcl[i] = gmem(a.bcol)[(size_t) rb0 * width + i];
        const unsigned int r = r0 + tid;
        unsigned int take = 0u;
        if (r < r1) {
            const bool kept = a.row_keep == nullptr || gmem(a.row_keep)[r] != 0u;
            if (!kept || ell == 0u) gr_write_empty(a, sink, r, kept);
            else take = 1u;
        }
        unsigned int nk = 0u;
        const unsigned int pos = block_exclusive_scan_u32(take, s_scan, &nk);
        if (take) wrow[pos] = r;
        if (tid == 0u) {
            s_next = 0u;
            s_nk = nk;
        }
        __syncthreads();
        const unsigned int n_rows = __builtin_amdgcn_readfirstlane(s_nk);
        auto next_row = [&](gb_cursor &c) {
            unsigned int i = 0u;
            if (lane == 0u) i = atomicAdd(&s_next, 1u);
            i = __builtin_amdgcn_readfirstlane(i);
            c.live = i < n_rows;
            c.row = c.live ? __builtin_amdgcn_readfirstlane(wrow[i]) : 0u;
            c.e = 0u;
        };
        auto advance = [&](gb_cursor &c) {
            if (!c.live) return;
            if (c.e + GB_CHUNK < ell) c.e += GB_CHUNK;
            else next_row(c);
        };
        // the lane's block column from LDS, its 8 keep bits from the feature
        // map (8 columns inside one 32-bit word), then the values; a lane with
        // none kept reads past the row's buffer (zeros, no memory traffic)
        auto load_vals = [&](const gb_cursor &c, gb_lane<VT> &s) {
            const unsigned int row = __builtin_amdgcn_readfirstlane(c.row);
            const unsigned int e0 = __builtin_amdgcn_readfirstlane(c.e) + 8u * lane;
            const bool in_row = c.live && e0 < ell;
            const unsigned int bc = bcl[in_row ? ((row >> lbs) - rb0) * width + (e0 >> lbs) : 0u];
            const bool in = in_row && bc < nblocks;
            s.c0 = (bc << lbs) + (e0 & (bs - 1u));
            const unsigned int w = L.fmap[min(s.c0 >> 5, a.words)].x;
            s.kw = in ? (w >> (s.c0 & 31u)) & 0xffu : 0u;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void *) ((const unsigned char *) a.src.val + (size_t) row * row_bytes), (short) 0, (int) row_bytes, 0x00020000);
            s.v = O::load(rs, s.kw != 0u ? e0 * O::bytes : row_bytes);
        };
        auto consume = [&](const gb_cursor &c, const gb_lane<VT> &s) {
            unsigned int col[8];
            float v[8];
            bool kept[8];
#pragma unroll
            for (unsigned int k = 0; k < 8u; ++k) {
                v[k] = O::get(s.v, k);
                col[k] = s.c0 + k;
                kept[k] = ((s.kw >> k) & 1u) != 0u && v[k] != 0.0f;
            }
            gr_accum<8u, TB>(a, L, col, v, kept, A);
            if (c.e + GB_CHUNK >= ell) gr_finish(a, sink, c.row, A);
        };
        gb_cursor ca, cb;
        next_row(ca);
        if (ca.live) {
            gb_lane<VT> RA, RB;
            load_vals(ca, RA);
            while (true) {
                cb = ca;
                advance(cb);
                load_vals(cb, RB);
                consume(ca, RA);
                if (!cb.live) break;
                ca = cb;
                advance(ca);
                load_vals(ca, RA);
                consume(cb, RB);
                if (!ca.live) break;
            }
        }
    }
}

// Blocked-ELL: workgroup b takes rows [rows b / grid, rows (b + 1) / grid)
// (every row holds ell_cols slots' values); TB as in gr_stream_kernel.
template<int VT, unsigned int TB>
__global__ void __launch_bounds__(GR_THREADS) gr_bell_kernel(gr_params a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ gr_sink sink;
    const unsigned int lo = (unsigned int) ((unsigned long long) a.src.rows * blockIdx.x / gridDim.x);
    const unsigned int hi = (unsigned int) ((unsigned long long) a.src.rows * (blockIdx.x + 1u) / gridDim.x);
    if (lo >= hi) return;  // uniform across the workgroup
    const gr_staged st = gr_stage<TB>(a, lds, sink);
    unsigned int *bcl = st.wrow + GR_WINDOW;  // [nrb x width] block columns of a window
    if constexpr (TB == 0u) {
        gr_bell_rows<VT, 0u>(a, st.L, sink, st.wrow, bcl, lo, hi);
    } else {
        if (st.in_lds) gr_bell_rows<VT, TB>(a, st.L, sink, st.wrow, bcl, lo, hi);
        else gr_bell_rows<VT, 0u>(a, st.L, sink, st.wrow, bcl, lo, hi);
    }
}

size_t fixed_lds_bytes() {
    return (size_t) GR_WAVES * (GR_QCAP + GR_QPAD) * 8u + (size_t) 3u * GR_WINDOW * 4u;
}

template<int VT, unsigned int TB, bool BELL>
int launch_stream(const gr_params &p, unsigned int grid, size_t lds, hipStream_t s) {
    const void *fn = BELL ? (const void *) gr_bell_kernel<VT, TB> : (const void *) gr_stream_kernel<VT, TB>;
    static bool attr_set = false;
    if (!attr_set) {
        CSAMD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int) GR_LDS_BUDGET), "group reduce lds");
        attr_set = true;
    }
    if constexpr (BELL) gr_bell_kernel<VT, TB><<<grid, GR_THREADS, lds, s>>>(p);
    else gr_stream_kernel<VT, TB><<<grid, GR_THREADS, lds, s>>>(p);
    return hip_ok(hipGetLastError(), BELL ? "gr_bell_kernel" : "gr_stream_kernel");
}

template<bool BELL> int launch_any(const gr_params &p, unsigned int tb, unsigned int grid, hipStream_t s) {
    // the whole budget, so that exactly one workgroup runs per CU
    const size_t lds = GR_LDS_BUDGET;
#define GR_TB_CASES(VT)                                                                                                          \
    switch (tb) {                                                                                                                \
        case 1u: return launch_stream<VT, 1u, BELL>(p, grid, lds, s);                                                           \
        case 2u: return launch_stream<VT, 2u, BELL>(p, grid, lds, s);                                                           \
        case 4u: return launch_stream<VT, 4u, BELL>(p, grid, lds, s);                                                           \
        default: return launch_stream<VT, 0u, BELL>(p, grid, lds, s);                                                           \
    }
    switch (p.src.value_type) {
        case CSAMD_VALUE_F16: GR_TB_CASES(CSAMD_VALUE_F16)
        case CSAMD_VALUE_BF16: GR_TB_CASES(CSAMD_VALUE_BF16)
        case CSAMD_VALUE_F32: GR_TB_CASES(CSAMD_VALUE_F32)
        case CSAMD_VALUE_FP8_E4M3: GR_TB_CASES(CSAMD_VALUE_FP8_E4M3)
        default: set_error("group reduce: value type %u", p.src.value_type); return 0;
    }
#undef GR_TB_CASES
}

size_t feature_lds_bytes(unsigned int cols) {
    return align_up((((size_t) cols + 31u) / 32u + 1u) * 8u, 16u);
}

// The kernels' common parameters; enqueues the feature map and the kept
// features' group masks.  Returns the LDS table entry bytes (0: no groups)
// through *tb.
int gr_prepare(unsigned int rows, unsigned int cols, const void *val, unsigned int value_type, const uint32_t *fgm,
               unsigned int group_count, const unsigned char *row_keep, const unsigned char *feature_keep,
               const csamd_sparse_group_filter_params &filter, const gr_outputs &out, void *scratch, size_t scratch_bytes,
               hipStream_t s, gr_params *p, unsigned int *tb) {
    CSAMD_REQUIRE(scratch != nullptr && scratch_bytes >= gr_scratch_bytes(cols), "group reduce: scratch %zu < %zu bytes",
                  scratch_bytes, gr_scratch_bytes(cols));
    CSAMD_REQUIRE(group_count <= CSAMD_MAX_MASK_GROUPS, "group reduce: group_count %u > %u", group_count, CSAMD_MAX_MASK_GROUPS);
    const unsigned int G = fgm != nullptr ? group_count : 0u;
    const unsigned int words = (cols + 31u) / 32u;
    char *base = (char *) scratch;
    uint2 *fmap = (uint2 *) (base + align_up(((size_t) GR_MAX_GRID + 1u) * 4u, 256u));
    uint32_t *ktab = (uint32_t *) ((char *) fmap + align_up(((size_t) words + 1u) * 8u, 256u));
    const uint32_t gbits = G >= 32u ? 0xffffffffu : (1u << G) - 1u;
    gr_fmap_kernel<<<1, 1024, 0, s>>>(feature_keep, cols, words, fmap);
    CSAMD_HIP(hipGetLastError(), "gr_fmap_kernel");
    if (G != 0u && cols != 0u) {
        gr_ktab_kernel<<<grid_for(cols, 256u, 1024u), 256, 0, s>>>(fmap, fgm, cols, gbits, ktab);
        CSAMD_HIP(hipGetLastError(), "gr_ktab_kernel");
    }
    std::memset(p, 0, sizeof(*p));
    p->src.rows = rows;
    p->src.cols = cols;
    p->src.val = val;
    p->src.value_type = value_type;
    p->row_keep = row_keep;
    p->fmap = fmap;
    p->words = words;
    p->ktab = ktab;
    p->fgm = fgm;
    p->gbits = gbits;
    p->G = G;
    unsigned int logP = 0u;  // P = min(2^ceil(log2 G), 16) lanes per queued entry
    while ((1u << logP) < G && logP < 4u) ++logP;
    p->logP = logP;
    p->filter = filter;
    p->o = out;
    if (G == 0u) {
        p->o.group_counts = nullptr;
        p->o.group_percentages = nullptr;
    }
    const size_t room = GR_LDS_BUDGET - feature_lds_bytes(cols) - fixed_lds_bytes();
    // LDS table entry bytes: the narrowest that holds G bits
    *tb = G == 0u ? 0u : G <= 8u ? 1u : G <= 16u ? 2u : 4u;
    p->cap = *tb != 0u ? (unsigned int) (room / *tb < cols ? room / *tb : cols) : 0u;
    p->cap &= ~15u;
    return 1;
}

unsigned int gr_grid() {
    unsigned int grid = (unsigned int) device_cu_count();
    if (grid < 1u) grid = 1u;
    return grid > GR_MAX_GRID ? GR_MAX_GRID : grid;
}

} // namespace

size_t gr_scratch_bytes(unsigned int cols) {
    const size_t words = ((size_t) cols + 31u) / 32u;
    return align_up(((size_t) GR_MAX_GRID + 1u) * 4u, 256u) + align_up((words + 1u) * 8u, 256u) + align_up((size_t) cols * 4u + 4u, 256u);
}

bool gr_eligible(const gr_source &src) {
    if (!value_type_decodable(src.value_type) || src.rows == 0u || src.idx == nullptr || src.val == nullptr) return false;
    const size_t words = ((size_t) src.cols + 31u) / 32u;
    if (words > GR_MAX_WORDS) return false;
    if (align_up((words + 1u) * 8u, 16u) + fixed_lds_bytes() > GR_LDS_BUDGET) return false;
    if (src.ptr == nullptr && src.slice_count != 0u &&
        (src.slice_offsets == nullptr || src.slice_widths == nullptr || src.slot_offsets == nullptr))
        return false;
    const size_t vb = (size_t) value_type_bytes(src.value_type);
    return ((uintptr_t) src.idx & 15u) == 0u && ((uintptr_t) src.val % (4u * vb)) == 0u;
}

int gr_reduce_rows(const gr_source &src, const uint32_t *fgm, unsigned int group_count, const unsigned char *row_keep,
                   const unsigned char *feature_keep, const csamd_sparse_group_filter_params &filter, const gr_outputs &out,
                   void *scratch, size_t scratch_bytes, hipStream_t s) {
    CSAMD_REQUIRE(gr_eligible(src), "group reduce: source not eligible for the streaming path");
    gr_params p;
    unsigned int tb = 0u;
    if (!gr_prepare(src.rows, src.cols, src.val, src.value_type, fgm, group_count, row_keep, feature_keep, filter, out, scratch,
                    scratch_bytes, s, &p, &tb))
        return 0;
    p.src = src;
    const unsigned int grid = gr_grid();
    unsigned int *bounds = (unsigned int *) scratch;
    gr_fill_kernel<<<grid_for(grid + 1u, 256u, 16u), 256, 0, s>>>(bounds, grid + 1u, src.rows);
    CSAMD_HIP(hipGetLastError(), "gr_fill_kernel");
    const unsigned int bthreads = src.rows > grid ? src.rows : grid;
    gr_bounds_kernel<<<grid_for(bthreads, 256u, 0x7fffffffu), 256, 0, s>>>(src, grid, bounds);
    CSAMD_HIP(hipGetLastError(), "gr_bounds_kernel");
    p.bounds = bounds;
    return launch_any<false>(p, tb, grid, s);
}

// LDS left for a Blocked-ELL window's block columns and the group-mask table
// (the row list takes GR_WINDOW words; the rest of the window area is reused).
size_t bell_room(unsigned int cols) {
    const size_t used = feature_lds_bytes(cols) + (size_t) GR_WAVES * (GR_QCAP + GR_QPAD) * 8u + (size_t) GR_WINDOW * 4u;
    return used < GR_LDS_BUDGET ? GR_LDS_BUDGET - used : 0u;
}

bool gr_bell_eligible(const gr_bell_source &src) {
    if (!value_type_decodable(src.value_type) || src.rows == 0u || src.val == nullptr) return false;
    const unsigned int bs = src.block_size;
    if (bs != 8u && bs != 16u && bs != 32u) return false;
    if (src.ell_cols % bs != 0u || (src.ell_cols != 0u && src.block_col_idx == nullptr)) return false;
    if (((size_t) src.cols + 31u) / 32u > GR_MAX_WORDS || feature_lds_bytes(src.cols) + fixed_lds_bytes() > GR_LDS_BUDGET) return false;
    if ((size_t) (src.ell_cols / bs) * 4u + 16u > bell_room(src.cols)) return false;  // one row block's block columns
    if ((size_t) src.ell_cols * value_type_bytes(src.value_type) > 0x7fffffffu) return false;  // a row's buffer range
    const size_t align = src.value_type == CSAMD_VALUE_FP8_E4M3 ? 8u : 16u;  // 8-value loads
    return ((uintptr_t) src.val % align) == 0u;
}

int gr_reduce_bell(const gr_bell_source &src, const uint32_t *fgm, unsigned int group_count, const unsigned char *row_keep,
                   const unsigned char *feature_keep, const csamd_sparse_group_filter_params &filter, const gr_outputs &out,
                   void *scratch, size_t scratch_bytes, hipStream_t s) {
    CSAMD_REQUIRE(gr_bell_eligible(src), "group reduce: Blocked-ELL source not eligible for the streaming path");
    gr_params p;
    unsigned int tb = 0u;
    if (!gr_prepare(src.rows, src.cols, src.val, src.value_type, fgm, group_count, row_keep, feature_keep, filter, out, scratch,
                    scratch_bytes, s, &p, &tb))
        return 0;
    p.bcol = src.block_col_idx;
    p.lbs = src.block_size == 8u ? 3u : src.block_size == 16u ? 4u : 5u;
    p.ell_cols = src.ell_cols;
    p.width = src.ell_cols / src.block_size;
    // a window stages nrb row blocks' block columns: up to half of the LDS
    // left (at least one row block), the rest holds the group-mask table
    const size_t room = bell_room(src.cols), one = (size_t) p.width * 4u;
    const unsigned int row_blocks = (src.rows + src.block_size - 1u) >> p.lbs;
    size_t nrb = one != 0u ? (room / 2u) / one : (size_t) row_blocks;
    if (nrb < 1u) nrb = 1u;
    if (nrb > (GR_WINDOW >> p.lbs)) nrb = GR_WINDOW >> p.lbs;
    if (nrb > row_blocks) nrb = row_blocks;
    p.nrb = (unsigned int) nrb;
    const size_t left = room - align_up(nrb * one, 16u);
    p.cap = tb != 0u ? (unsigned int) (left / tb < src.cols ? left / tb : src.cols) & ~15u : 0u;
    unsigned int grid = gr_grid();
    const unsigned int by_rows = (src.rows + GR_WAVES - 1u) / GR_WAVES;  // at least a row per wave
    if (grid > by_rows) grid = by_rows;
    return launch_any<true>(p, tb, grid, s);
}

// ---------------------------------------------------------------------------
// Allocation-free entry point: the streaming path's scratch (workgroup row
// bounds, feature map, kept-feature group masks) in one fixed device array.
// ---------------------------------------------------------------------------
constexpr size_t GR_NOALLOC_SCRATCH = 512u * 1024u;  // every source of up to 122,303 columns
constexpr int GR_NOALLOC_DEVICES = 64;

// Each GPU the library runs on holds its own copy until the process ends.
// Host code never passes this name to a kernel: it is a host-side stand-in;
// the device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char noalloc_scratch[GR_NOALLOC_SCRATCH];

namespace {
std::mutex noalloc_mu;
void *noalloc_ptr[GR_NOALLOC_DEVICES];

// Device address of the current device's noalloc_scratch, looked up on the
// first call on that device (with it current) and kept.
int noalloc_scratch_address(void **out) {
    int dev = -1;
    CSAMD_HIP(hipGetDevice(&dev), "hipGetDevice (allocation-free group reduce)");
    CSAMD_REQUIRE(dev >= 0 && dev < GR_NOALLOC_DEVICES, "allocation-free group reduce: device %d outside [0, %d)", dev,
                  GR_NOALLOC_DEVICES);
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_ptr[dev] == nullptr) {
        void *p = nullptr;
        CSAMD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)), "hipGetSymbolAddress (allocation-free group reduce scratch)");
        CSAMD_REQUIRE(p != nullptr, "allocation-free group reduce: no device address for the scratch array on device %d", dev);
        noalloc_ptr[dev] = p;
    }
    *out = noalloc_ptr[dev];
    return 1;
}
} // namespace

} // namespace csamd

using namespace csamd;

extern "C" int csamd_compute_sparse_group_reduce_compressed_noalloc(const csamd_compressed_view *src, unsigned int value_type,
                                                                    const csamd_group_mask_config_view *groups,
                                                                    const csamd_row_feature_mask_view *masks,
                                                                    const csamd_sparse_group_filter_params *filter,
                                                                    const csamd_sparse_group_reduce_result *out,
                                                                    csamd_stream_t stream) {
    CSAMD_REQUIRE(src != nullptr && filter != nullptr && out != nullptr,
                  "compute_sparse_group_reduce_compressed_noalloc: null argument");
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW, "compute_sparse_group_reduce_compressed_noalloc: axis must be row");
    CSAMD_REQUIRE(value_type_decodable(value_type), "compute_sparse_group_reduce_compressed_noalloc: unsupported value type %u",
                  value_type);
    const unsigned int G = groups != nullptr ? groups->group_count : 0u;
    CSAMD_REQUIRE(G <= CSAMD_MAX_MASK_GROUPS, "compute_sparse_group_reduce_compressed_noalloc: group_count %u > %u", G,
                  CSAMD_MAX_MASK_GROUPS);
    const uint32_t *fgm = groups != nullptr ? groups->feature_group_masks : nullptr;
    CSAMD_REQUIRE(G == 0u || fgm != nullptr,
                  "compute_sparse_group_reduce_compressed_noalloc: %u groups need device feature_group_masks", G);
    CSAMD_REQUIRE(out->rows >= src->rows && out->group_count == G,
                  "compute_sparse_group_reduce_compressed_noalloc: outputs hold %u rows x %u groups, the call needs %u x %u",
                  out->rows, out->group_count, src->rows, G);
    CSAMD_REQUIRE(out->row_totals != nullptr && out->max_values != nullptr && out->detected_features != nullptr &&
                      out->row_keep != nullptr && (G == 0u || (out->group_counts != nullptr && out->group_percentages != nullptr)),
                  "compute_sparse_group_reduce_compressed_noalloc: null output array");
    if (src->rows == 0u) return 1;
    gr_source gs;
    std::memset(&gs, 0, sizeof(gs));
    gs.rows = src->rows;
    gs.cols = src->cols;
    gs.ptr = src->majorPtr;
    gs.idx = src->minorIdx;
    gs.val = src->val;
    gs.value_type = value_type;
    CSAMD_REQUIRE(gr_eligible(gs),
                  "compute_sparse_group_reduce_compressed_noalloc: source not taken by the streaming path (more than %u columns, "
                  "or index / value arrays not aligned to 4-entry units)",
                  GR_MAX_WORDS * 32u);
    const size_t need = gr_scratch_bytes(src->cols);
    CSAMD_REQUIRE(need <= GR_NOALLOC_SCRATCH,
                  "compute_sparse_group_reduce_compressed_noalloc: %u columns need %zu scratch bytes, above the fixed %zu",
                  src->cols, need, GR_NOALLOC_SCRATCH);
    void *scratch = nullptr;
    if (!noalloc_scratch_address(&scratch)) return 0;
    const gr_outputs o{out->row_totals, out->max_values, out->detected_features, out->row_keep, G != 0u ? out->group_counts : nullptr,
                       G != 0u ? out->group_percentages : nullptr};
    return gr_reduce_rows(gs, fgm, G, masks != nullptr ? masks->row_keep : nullptr, masks != nullptr ? masks->feature_keep : nullptr,
                          *filter, o, scratch, GR_NOALLOC_SCRATCH, (hipStream_t) stream);
}
