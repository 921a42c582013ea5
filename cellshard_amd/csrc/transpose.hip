// Deterministic compressed transpose (CSR -> CSC or CSC -> CSR).
//
// Reference: src/convert/compressed_transpose_raw.cuh:14-122 (cuSPARSE
// csr2cscEx2, else count / scan / atomic scatter with nondeterministic order
// inside a column).  Here the output order is canonical: inside every output
// segment the source major index ascends (what csr2cscEx2 and scipy give).
//
// gfx950 plan (sources whose rows hold strictly ascending minor indices):
//   K1 count: one 1024-thread workgroup per tile of TR_ROWS source rows, one
//      wave per row, an LDS histogram over all output segments; the same
//      pass records where each TR_BAND-wide column band starts inside each
//      row, band-major per tile (rbt[t][band][row]), staged in LDS and
//      written whole where it fits (about 40k segments);
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
//      Chunks are ordered so that each XCD's L2 sees neighbouring chunks.
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

// K1 row walk of the general count kernel: one wave per source row, 64 U
// entries per pass with U independent loads per lane (all issued before any
// use); the previous entry's column comes from the neighbour lane.
// count(c) counts an entry of column c; open(k, i) records that band k of
// the row starts at entry i (every band 0..bands of a strictly ascending row
// is opened exactly once; band `bands` is the row end).
// flags[0]: some row's minor indices are not strictly ascending;
// flags[1]: a minor index >= uDim.
template<unsigned int U, typename Count, typename Open>
__device__ __forceinline__ void tr_walk_row(unsigned int b, unsigned int e, unsigned int uDim, unsigned int bands,
                                            const unsigned int *__restrict__ idx, unsigned int lane, unsigned int *__restrict__ flags,
                                            Count count, Open open) {
    if (b == e) {
        for (unsigned int k = lane; k <= bands; k += 64u) open(k, b);
        return;
    }
    unsigned int carry = CSAMD_INVALID_COL;  // column of the entry before the pass
    for (unsigned int i0 = b; i0 < e; i0 += 64u * U) {
        unsigned int c[U];
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) {
            const unsigned int i = i0 + 64u * u + lane;
            c[u] = i < e ? idx[i] : CSAMD_INVALID_COL;
        }
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) {
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
            count(cc);
            // entry i opens every band in (band of entry i - 1, band of cc]
            const unsigned int band = cc / TR_BAND;
            int prev = -1;
            if (i > b) {
                prev = pc < uDim ? (int) (pc / TR_BAND) : (int) bands;
                if (pc >= cc) flags[0] = 1u;
            }
            for (int k = prev + 1; k <= (int) band; ++k) open((unsigned int) k, i);
            if (i == e - 1u)
                for (unsigned int k = band + 1u; k <= bands; ++k) open(k, e);
        }
        carry = (unsigned int) __shfl((int) c[U - 1u], 63);
    }
}

// K1 (general): counts in an LDS u32 histogram over all segments
// (use_lds) or straight into the tile's global counts; band starts
// straight into rbt[t][band][row] (rbt may be null).
template<unsigned int U>
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
        tr_walk_row<U>(
            ptr[r], ptr[r + 1u], uDim, bands, idx, lane, flags,
            [&](unsigned int c) {
                if (use_lds) atomicAdd(&hist[c], 1u);
                else atomicAdd(&tile_counts[c], 1u);
            },
            [&](unsigned int k, unsigned int i) {
                if (tile_rbt != nullptr) tile_rbt[(size_t) k * TR_ROWS + rl] = i;
            });
    }
    if (use_lds) {
        __syncthreads();
        for (unsigned int c = threadIdx.x; c < uDim; c += blockDim.x) tile_counts[c] = hist[c];
    }
}

// K1 (packed, uDim < 65536 and the tile's tables fit in LDS): two u16
// counters per LDS word (a strictly ascending row adds at most 1 per
// segment, so a tile's count stays <= TR_ROWS), and the band starts as u16
// offsets from the row start in LDS: an entry whose band differs from the
// previous entry's (or that starts its row) stores its offset at its band
// (one predicated store, no loop); bands a row skips keep a sentinel and
// take the next band's start in a backward pass per row at the end, when
// the table is written out coalesced (every rbt line once and whole).  A
// tile with a row that is not strictly ascending may overflow a u16
// counter: flags[0] is then set and the host counts again with
// tr_count_kernel.
//
// The tile's entries are walked flat, not row by row, in groups of 4
// consecutive entries (one 16-byte load per lane and group): pass p covers
// 64 U groups from the tile's first group + 64 U p, wave w takes passes w,
// w + 16, ..., and the loads of a wave's next pass are in flight while the
// current one is counted.  Each lane finds its entries' row from the row
// starts in LDS (binary search once per pass, then a forward walk).  The
// group holding the tile's first entry may start in the previous tile and
// the one holding its last may end in the next (those entries are skipped);
// the array's final partial group, if any, is read one entry at a time.
template<unsigned int U>
__global__ void __launch_bounds__(TR_COUNT_THREADS) tr_count_packed_kernel(unsigned int cDim, unsigned int uDim, unsigned int nnz,
                                                                           const unsigned int *__restrict__ ptr,
                                                                           const unsigned int *__restrict__ idx, unsigned int bands,
                                                                           unsigned int *__restrict__ counts, unsigned int *__restrict__ rbt,
                                                                           unsigned int *__restrict__ flags) {
    constexpr unsigned int PG = 64u * U;  // groups per pass
    constexpr unsigned int NW = TR_COUNT_THREADS / 64u;
    extern __shared__ unsigned int lds_words[];
    const unsigned int hw = (uDim + 1u) >> 1;
    constexpr unsigned short NONE = 0xffffu;                          // band without entries in the row (so far)
    unsigned int *hist = lds_words;                                  // [hw] two u16 counters each
    unsigned int *rp = hist + hw;                                    // [TR_ROWS + 1] row starts (+ pad)
    unsigned short *off = (unsigned short *) (rp + TR_ROWS + 4u);    // [bands][TR_ROWS] band starts - row start
    const unsigned int t = blockIdx.x;
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int nr = cDim - r0 < TR_ROWS ? cDim - r0 : TR_ROWS;
    const unsigned int tid = threadIdx.x;
    for (unsigned int w = tid; w < hw; w += blockDim.x) hist[w] = 0u;
    for (unsigned int w = tid; w < bands * (TR_ROWS / 2u); w += blockDim.x) ((unsigned int *) off)[w] = 0xffffffffu;
    for (unsigned int r = tid; r <= nr; r += blockDim.x) rp[r] = ptr[r0 + r];
    __syncthreads();
    const unsigned int E0 = rp[0], E1 = rp[nr];
    const unsigned int wave = tid >> 6, lane = lane_id();
    const unsigned int g0 = E0 >> 2, g1 = (E1 + 3u) >> 2;  // groups [g0, g1) hold the tile
    const unsigned int ng = E1 > E0 ? g1 - g0 : 0u;
    const unsigned int gfull = nnz >> 2;                     // groups wholly inside the array
    const uint4 *idx4 = (const uint4 *) idx;

    // 16-byte loads only (groups past the array's last whole group read that
    // group; their entries are skipped here and visited one by one at the end)
    auto load = [&](unsigned int p, uint4 (&c)[U], unsigned int &before) {
        const unsigned int gp = g0 + p * PG;
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) {
            const unsigned int g = gp + 64u * u + lane;
            const unsigned int gc = g < g1 ? g : g1 - 1u;
            c[u] = idx4[gc < gfull ? gc : gfull - 1u];
        }
        before = idx[gp > 0u ? gp * 4u - 1u : 0u];
    };
    // one entry: i in [E0, E1), its column cc, the previous entry's column pc
    // (used only when i is not its row's first), its row's start rs
    auto visit = [&](unsigned int i, unsigned int cc, unsigned int pc, unsigned int row, unsigned int rs, bool &bad,
                     bool &unsorted) {
        bad |= cc >= uDim;
        const unsigned int cs = min(cc, uDim - 1u);  // out of range: reported, not counted right
        atomicAdd(&hist[cs >> 1], 1u << ((cs & 1u) << 4));
        const bool inner = i > rs;
        unsorted |= inner && pc >= cc;
        const unsigned int band = cs / TR_BAND;
        if (!inner || band != min(pc, uDim - 1u) / TR_BAND) off[band * TR_ROWS + row] = (unsigned short) (i - rs);
    };
    // last row with rp[row] <= i (i in [E0, E1))
    auto row_of = [&](unsigned int i) -> unsigned int {
        unsigned int lo = 0u, hi = nr - 1u;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (rp[mid] <= i) lo = mid;
            else hi = mid - 1u;
        }
        return lo;
    };
    const unsigned int Ev = E1 < gfull * 4u ? E1 : gfull * 4u;  // entries read by the group loads: [E0, Ev)
    const unsigned int cmax = uDim - 1u;                        // uDim > 0 here
    // one lane-private u16 slot per lane (4 bytes apart) takes the band-start
    // stores of entries that open no band, so every entry stores without a branch
    const unsigned int junk = bands * TR_ROWS + 2u * lane;
    auto count = [&](unsigned int p, const uint4 (&c)[U], unsigned int before) {
        const unsigned int ib = (g0 + p * PG) * 4u;  // first entry of the pass
        unsigned int i = ib + 4u * lane;
        unsigned int row = row_of(i < E0 ? E0 : (i < E1 ? i : E1 - 1u));
        unsigned int rs = rp[row], re = rp[row + 1u];
        unsigned int carry = before;  // column of the entry before the group row u
        bool bad = false, unsorted = false;
        unsigned int top = 0u;  // largest column seen on the group path
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) {
            const unsigned int up = (unsigned int) __shfl_up((int) c[u].w, 1);
            const unsigned int first_prev = lane == 0u ? carry : up;
            carry = (unsigned int) __shfl((int) c[u].w, 63);
            const unsigned int col[4] = {c[u].x, c[u].y, c[u].z, c[u].w};
            const unsigned int base = ib + 256u * u + 4u * lane;
            // forward to the row of the group's first entry (empty rows skipped)
            const unsigned int bi = base < E0 ? E0 : (base < E1 ? base : E1 - 1u);
            while (bi >= re) {
                ++row;
                rs = re;
                re = rp[row + 1u];
            }
            const bool whole = base >= E0 && base < Ev && Ev - base >= 4u && re - base >= 4u;
            if (__all(whole)) {
                // every lane's four entries lie inside the read range and one row
                const unsigned int d = base - rs;
                unsigned int cs[4], bd[4];
#pragma unroll
                for (unsigned int e = 0; e < 4u; ++e) {
                    cs[e] = min(col[e], cmax);
                    bd[e] = cs[e] / TR_BAND;
                    atomicAdd(&hist[cs[e] >> 1], 1u << ((cs[e] & 1u) << 4));
                }
                top = max(top, max(max(col[0], col[1]), max(col[2], col[3])));
                unsorted |= (d != 0u && first_prev >= col[0]) | (col[0] >= col[1]) | (col[1] >= col[2]) | (col[2] >= col[3]);
                const bool open0 = d == 0u || bd[0] != min(first_prev, cmax) / TR_BAND;
                off[open0 ? bd[0] * TR_ROWS + row : junk] = (unsigned short) d;
#pragma unroll
                for (unsigned int e = 1; e < 4u; ++e)
                    off[bd[e] != bd[e - 1u] ? bd[e] * TR_ROWS + row : junk] = (unsigned short) (d + e);
                continue;
            }
#pragma unroll
            for (unsigned int e = 0; e < 4u; ++e) {
                i = base + e;
                if (i < E0 || i >= Ev) continue;
                while (i >= re) {  // forward to the entry's row (empty rows skipped)
                    ++row;
                    rs = re;
                    re = rp[row + 1u];
                }
                visit(i, col[e], e == 0u ? first_prev : col[e == 0u ? 0u : e - 1u], row, rs, bad, unsorted);
            }
        }
        bad |= top > cmax;
        if (bad) flags[1] = 1u;
        if (unsorted) flags[0] = 1u;
    };
    // two register sets in turn: the next pass loads while this one counts
    unsigned int p = wave;
    if (p * PG < ng) {
        uint4 ca[U], cb[U];
        unsigned int ba, bb;
        load(p, ca, ba);
        for (;;) {
            unsigned int pn = p + NW;
            bool more = pn * PG < ng;
            load(more ? pn : p, cb, bb);
            count(p, ca, ba);
            if (!more) break;
            p = pn;
            pn = p + NW;
            more = pn * PG < ng;
            load(more ? pn : p, ca, ba);
            count(p, cb, bb);
            if (!more) break;
            p = pn;
        }
    }
    // the array's final partial group (at most 3 entries, last tile only)
    if (tid == 0u && Ev < E1) {
        bool bad = false, unsorted = false;
        for (unsigned int i = Ev > E0 ? Ev : E0; i < E1; ++i) {
            const unsigned int row = row_of(i);
            visit(i, idx[i], idx[i > 0u ? i - 1u : 0u], row, rp[row], bad, unsorted);
        }
        if (bad) flags[1] = 1u;
        if (unsorted) flags[0] = 1u;
    }
    __syncthreads();
    unsigned int *tile_counts = counts + (size_t) t * uDim;
    for (unsigned int c = tid; c < uDim; c += blockDim.x) tile_counts[c] = (hist[c >> 1] >> ((c & 1u) << 4)) & 0xffffu;
    // band starts, last band first: a band the row skips starts where the
    // next one does; band `bands` is the row end
    if (tid < nr) {
        unsigned int *tile_rbt = rbt + (size_t) t * (bands + 1u) * TR_ROWS + tid;
        const unsigned int rs = rp[tid];
        unsigned int next = rp[tid + 1u];
        tile_rbt[(size_t) bands * TR_ROWS] = next;
        for (unsigned int k = bands; k-- > 0u;) {
            const unsigned int o = off[k * TR_ROWS + tid];
            if (o != NONE) next = rs + o;
            tile_rbt[(size_t) k * TR_ROWS] = next;
        }
    }
}

size_t count_packed_lds_bytes(unsigned int uDim, unsigned int bands) {
    return (size_t) ((uDim + 1u) >> 1) * 4u + (size_t) (TR_ROWS + 4u) * 4u + (size_t) bands * TR_ROWS * 2u + 64u * 4u;
}

// K2: counts[t][c] <- sum_{t' < t} counts[t'][c] (in place), coltot[c] = total.
__global__ void tr_tile_offsets_kernel(unsigned int *__restrict__ counts, unsigned int tiles, unsigned int uDim,
                                       unsigned int *__restrict__ coltot) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < uDim; c += gridDim.x * blockDim.x) {
        unsigned int run = 0u;
        unsigned int t = 0u;
        // one thread per segment walks every tile: 32 loads in flight per
        // step (the grid is only uDim threads, so each needs many)
        constexpr unsigned int TU = 32u;
        for (; t + TU <= tiles; t += TU) {
            unsigned int n[TU];
#pragma unroll
            for (unsigned int k = 0; k < TU; ++k) n[k] = counts[(size_t) (t + k) * uDim + c];
#pragma unroll
            for (unsigned int k = 0; k < TU; ++k) {
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

// K4 (sorted sources): one workgroup per (tile, band) chunk; the chunk's
// entries are sorted in LDS by output segment, stably by source row, in
// batches of at most K x 256 entries (whole rows).  Every segment of the
// band keeps a TR_ROWS-bit mask of the batch rows holding it; an entry's
// rank in its segment's run is the popcount of the mask below its row;
// runs are staged in LDS in output order and written with consecutive
// lanes on consecutive positions.  Each thread loads all of its entries of
// a batch at once, so one memory latency per batch is exposed, not one per
// row.
//
// Chunk order.  Workgroups go to the 8 XCDs round robin (workgroup i to XCD
// i mod 8), each XCD with its own L2.  Workgroup i takes logical chunk
// (i mod 8) * per + i / 8, so every XCD walks one contiguous range of
// logical chunks in order.  Logical chunks run over groups of `group`
// consecutive tiles, band-major inside a group: the chunks on one XCD at
// one time are a few neighbouring bands of a few neighbouring tiles.
struct tr_chunk_order {
    unsigned int tiles, bands, per, group;
};

// Logical chunk L (< tiles * bands) -> (tile, band).
__device__ __forceinline__ void tr_chunk(const tr_chunk_order &o, unsigned int L, unsigned int &t, unsigned int &band) {
    const unsigned int full = o.tiles / o.group;  // complete groups
    const unsigned int gsz = o.group * o.bands;   // chunks per complete group
    const unsigned int g = L / gsz;
    if (g < full) {
        const unsigned int rem = L - g * gsz;
        band = rem / o.group;
        t = g * o.group + (rem - band * o.group);
        return;
    }
    const unsigned int last = o.tiles - full * o.group;  // tiles in the final partial group
    const unsigned int rem = L - full * gsz;
    band = rem / last;
    t = full * o.group + (rem - band * last);
}

// Staged entry of the chunk sort: batch row | segment << 8 | value bits,
// one register for 1- and 2-byte values, two for 4-byte values.
template<typename S> struct tr_rec {
    using T = unsigned int;
    static __device__ __forceinline__ T pack(unsigned int r, unsigned int seg, S v) { return r | (seg << 8) | ((unsigned int) v << 16); }
    static __device__ __forceinline__ unsigned int key(T x) { return x; }
    static __device__ __forceinline__ S val(T x) { return (S) (x >> 16); }
};
template<> struct tr_rec<float> {
    using T = uint2;
    static __device__ __forceinline__ T pack(unsigned int r, unsigned int seg, float v) { return make_uint2(r | (seg << 8), __float_as_uint(v)); }
    static __device__ __forceinline__ unsigned int key(T x) { return x.x; }
    static __device__ __forceinline__ float val(T x) { return __uint_as_float(x.y); }
};

// LDS of tr_band_kernel: the rows mask and the run start per (mask word,
// segment), per-row band starts and run offsets, per-segment output deltas,
// then the batch's entry map, which the sorted staging reuses (about 40 KB
// at 6,144 entries of 2-byte values, so four workgroups fit in a CU's LDS;
// a batch that holds most chunks whole matters more than a fifth
// workgroup: 4,096-entry batches with five workgroups per CU were 8 %
// slower).
struct tr_band_lds {
    static constexpr size_t masks = 0;
    static constexpr size_t pre = masks + (size_t) TR_MASK_WORDS * TR_BAND * 4u;
    static constexpr size_t seg_b = pre + (size_t) TR_MASK_WORDS * TR_BAND * 2u;
    static constexpr size_t seg_o = seg_b + TR_ROWS * 4u;
    static constexpr size_t delta = seg_o + (TR_ROWS + 1u) * 4u;
    static constexpr size_t stage = (delta + TR_BAND * 4u + 15u) & ~(size_t) 15u;
};

template<typename S, unsigned int K>
__global__ void __launch_bounds__(TR_SORT_THREADS) tr_band_kernel(unsigned int cDim, unsigned int uDim, const unsigned int *__restrict__ idx,
                                                                  const S *__restrict__ val, const unsigned int *__restrict__ tile_off,
                                                                  const unsigned int *__restrict__ out_ptr,
                                                                  const unsigned int *__restrict__ rbt, tr_chunk_order order,
                                                                  unsigned int *__restrict__ out_idx, S *__restrict__ out_val) {
    static_assert(TR_SORT_THREADS == TR_ROWS && TR_SORT_THREADS == TR_BAND, "one row and one segment per thread");
    static_assert(TR_ROWS <= 256u && TR_BAND <= 256u, "row and segment packed in 8 bits each");
    using R = tr_rec<S>;
    using RT = typename R::T;
    constexpr unsigned int T = TR_SORT_THREADS;
    constexpr unsigned int cap = K * T;
    extern __shared__ __align__(16) unsigned char lds[];
    unsigned int *masks = (unsigned int *) (lds + tr_band_lds::masks);   // [TR_MASK_WORDS][TR_BAND] rows holding each segment,
                                                                         // word-major: lanes on different segments hit
                                                                         // different banks
    unsigned short *pre = (unsigned short *) (lds + tr_band_lds::pre);   // [TR_MASK_WORDS][TR_BAND] run start of each mask word
    unsigned int *seg_b = (unsigned int *) (lds + tr_band_lds::seg_b);   // [TR_ROWS] row's band start (entry index)
    unsigned int *seg_o = (unsigned int *) (lds + tr_band_lds::seg_o);   // [TR_ROWS + 1] chunk offsets of the rows' runs
    unsigned int *delta = (unsigned int *) (lds + tr_band_lds::delta);   // [TR_BAND] output position - staging slot
    unsigned int *s_map = (unsigned int *) (lds + tr_band_lds::stage);   // [cap] entry -> (offset from the batch start) << 8 | row
    RT *s_out = (RT *) (lds + tr_band_lds::stage);                       // [cap] sorted entries (after the map is read)
    __shared__ unsigned int scan_smem[16];

    const unsigned int tid = threadIdx.x;
    const unsigned int L = (blockIdx.x & 7u) * order.per + (blockIdx.x >> 3);
    if (L >= order.tiles * order.bands) return;  // the < 8 spare workgroups of the grid
    unsigned int t, band;
    tr_chunk(order, L, t, band);
    const unsigned int bands = order.bands;
    const unsigned int c0 = band * TR_BAND;
    const unsigned int nb = uDim - c0 < TR_BAND ? uDim - c0 : TR_BAND;
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int nr = cDim - r0 < TR_ROWS ? cDim - r0 : TR_ROWS;
    const unsigned int *rb_lo = rbt + ((size_t) t * (bands + 1u) + band) * TR_ROWS;

    for (unsigned int w = tid; w < TR_BAND * TR_MASK_WORDS; w += T) masks[w] = 0u;
    // segments past uDim start at 0 and hold no entries
    // thread tid owns segment tid's output cursor
    unsigned int gcur = tid < nb ? out_ptr[c0 + tid] + tile_off[(size_t) t * uDim + c0 + tid] : 0u;
    {
        const unsigned int b = tid < nr ? rb_lo[tid] : 0u;
        const unsigned int n = tid < nr ? rb_lo[tid + TR_ROWS] - b : 0u;
        unsigned int total = 0u;
        const unsigned int o = block_exclusive_scan_u32(n, scan_smem, &total);
        seg_b[tid] = b;
        seg_o[tid] = o;
        if (tid == 0u) seg_o[TR_ROWS] = total;
    }
    __syncthreads();

    unsigned int rb0 = 0u;
    while (rb0 < nr) {
        // batch: rows [rb0, rb1), the most whole rows within cap entries
        // whose span in the arrays stays below 2^24 entries (seg_o and the
        // band ends are nondecreasing; one row holds at most TR_BAND <= cap
        // entries of a band and spans at most uDim < 2^24 entries, so a
        // batch holds at least one row).  The barrier also ends the previous
        // batch: its D and E, and its reads of the staging and delta, finish
        // before this batch writes them.
        const unsigned int base = seg_o[rb0];
        const unsigned int b0 = seg_b[rb0];
        const int fits = tid >= rb0 && tid < nr && seg_o[tid + 1u] - base <= cap &&
                         seg_b[tid] + (seg_o[tid + 1u] - seg_o[tid]) - b0 < (1u << 24);
        const unsigned int rb1 = rb0 + (unsigned int) __syncthreads_count(fits);
        if (rb1 == rb0) break;  // unreachable for a valid band table (uniform): never loop without progress
        const unsigned int n = seg_o[rb1] - base;
        if (n == 0u) {  // rows without entries in the band (uniform)
            rb0 = rb1;
            continue;
        }
        // map: the row owner lists its entries' offsets from the batch start
        if (tid >= rb0 && tid < rb1) {
            const unsigned int o = seg_o[tid] - base, len = seg_o[tid + 1u] - seg_o[tid];
            const unsigned int tag = ((seg_b[tid] - b0) << 8) | (tid - rb0);
            for (unsigned int k = 0; k < len; ++k) s_map[o + k] = tag + (k << 8);
        }
        __syncthreads();
        // A: load this thread's entries (all loads issued before any use),
        // pack each as one staged record, mark which rows hold each segment
        // (32-bit byte offsets from the batch start: the offsets stay below
        // 2^26, so each load is one base register pair + one offset)
        const CSAMD_GLOBAL char *ib = (const CSAMD_GLOBAL char *) gmem(idx + b0);
        const CSAMD_GLOBAL char *vbp = (const CSAMD_GLOBAL char *) gmem(val + b0);
        RT x[K];
        {
            unsigned int m[K];
#pragma unroll
            for (unsigned int k = 0; k < K; ++k) {
                const unsigned int j = tid + k * T;
                m[k] = s_map[j < n ? j : n - 1u];
            }
            unsigned int c[K];
            S v[K];
#pragma unroll
            for (unsigned int k = 0; k < K; ++k) {
                c[k] = *(const CSAMD_GLOBAL unsigned int *) (ib + (m[k] >> 8) * 4u);
                v[k] = *(const CSAMD_GLOBAL S *) (vbp + (m[k] >> 8) * (unsigned int) sizeof(S));
            }
#pragma unroll
            for (unsigned int k = 0; k < K; ++k) x[k] = R::pack(m[k] & 255u, c[k] - c0, v[k]);
        }
#pragma unroll
        for (unsigned int k = 0; k < K; ++k) {
            if (tid + k * T >= n) continue;
            const unsigned int key = R::key(x[k]);
            const unsigned int r = key & 255u, seg = (key >> 8) & 255u;
            atomicOr(&masks[(r >> 5) * TR_BAND + seg], 1u << (r & 31u));
        }
        __syncthreads();
        // B: run lengths -> run starts in the staging, each mask word's
        // first slot (run start + rows in the lower words), and each
        // segment's output position minus its run start
        unsigned int cnt = 0u;
        {
            unsigned int cw[TR_MASK_WORDS];
#pragma unroll
            for (unsigned int w = 0; w < TR_MASK_WORDS; ++w) {
                cw[w] = (unsigned int) __popc(masks[w * TR_BAND + tid]);
                cnt += cw[w];
            }
            unsigned int total = 0u;
            unsigned int o = block_exclusive_scan_u32(cnt, scan_smem, &total);
            delta[tid] = gcur - o;
#pragma unroll
            for (unsigned int w = 0; w < TR_MASK_WORDS; ++w) {
                pre[w * TR_BAND + tid] = (unsigned short) o;
                o += cw[w];
            }
        }
        __syncthreads();
        // C: every entry to its run start + rank among earlier rows (the map
        // was last read in A, so the staging may overwrite it).  The key is
        // made opaque so that A's mask addresses and bits are recomputed here
        // rather than held in registers across the barriers.
#pragma unroll
        for (unsigned int k = 0; k < K; ++k) {
            if (tid + k * T >= n) continue;
            unsigned int key = R::key(x[k]);
            asm volatile("" : "+v"(key));
            const unsigned int r = key & 255u, seg = (key >> 8) & 255u;
            const unsigned int wi = (r >> 5) * TR_BAND + seg;
            s_out[pre[wi] + (unsigned int) __popc(masks[wi] & ((1u << (r & 31u)) - 1u))] = x[k];
        }
        __syncthreads();
        // D: runs to their final places (consecutive staging slots of one run
        // are consecutive output positions).  DU slots per thread per step:
        // all their staged records are read, then all their segments'
        // deltas, then the stores, so a step waits for two LDS round trips
        // rather than two per slot (slots past n read slot n - 1, unused).
        constexpr unsigned int DU = 8u;  // the asm below names eight
        for (unsigned int k0 = 0u; k0 * T < n; k0 += DU) {  // uniform
            RT e[DU];
            unsigned int d[DU];
#pragma unroll
            for (unsigned int u = 0; u < DU; ++u) {
                const unsigned int k = tid + (k0 + u) * T;
                e[u] = s_out[k < n ? k : n - 1u];
            }
#pragma unroll
            for (unsigned int u = 0; u < DU; ++u) d[u] = delta[(R::key(e[u]) >> 8) & 255u];
            // all eight read here, not sunk one by one into the guarded stores
            asm volatile("" : "+v"(d[0]), "+v"(d[1]), "+v"(d[2]), "+v"(d[3]), "+v"(d[4]), "+v"(d[5]), "+v"(d[6]), "+v"(d[7]));
#pragma unroll
            for (unsigned int u = 0; u < DU; ++u) {
                const unsigned int k = tid + (k0 + u) * T;
                if (k >= n) continue;
                const unsigned int key = R::key(e[u]);
                const unsigned int pos = d[u] + k;
                out_idx[pos] = r0 + rb0 + (key & 255u);
                out_val[pos] = R::val(e[u]);
            }
        }
        // E: advance the cursors, clear the masks.  No barrier before: D reads
        // neither (gcur was folded into delta in B, the masks were last read
        // in C); the next batch's first barrier orders E before its A.
        gcur += cnt;
#pragma unroll
        for (unsigned int w = 0; w < TR_MASK_WORDS; ++w) masks[w * TR_BAND + tid] = 0u;
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
    return tr_band_lds::stage + (size_t) cap * (vb == 4u ? 8u : 4u);
}

template<typename S>
int launch_scatter(const tr_layout &L, bool sorted, unsigned int cDim, unsigned int uDim, const unsigned int *ptr,
                   const unsigned int *idx, const void *val, const unsigned int *tile_off, const unsigned int *out_ptr,
                   const unsigned int *rbt, unsigned int *out_idx, void *out_val, hipStream_t s) {
    if (sorted) {
        const unsigned long long chunks = (unsigned long long) L.tiles * L.bands;
        CSAMD_REQUIRE(chunks <= 0xffffffffull, "transpose: too many chunks");
        const unsigned long long per = (chunks + 7ull) / 8ull;
        // tiles per group (CSAMD_TR_G for sweeps)
        const char *genv = std::getenv("CSAMD_TR_G");
        const unsigned int gw = genv != nullptr ? (unsigned int) std::atoi(genv) : 8u;
        const tr_chunk_order order{L.tiles, L.bands, (unsigned int) per, gw >= 1u && gw <= 64u ? gw : 8u};
        // staging of K x 256 entries per batch (CSAMD_TR_K = 8 / 12 / 16 / 20 / 24 for sweeps)
        const char *env = std::getenv("CSAMD_TR_K");
        const int kw = env != nullptr ? std::atoi(env) : 24;
        const unsigned int K = kw == 8 ? 8u : kw == 12 ? 12u : kw == 16 ? 16u : kw == 24 ? 24u : 20u;
        const size_t lds = band_lds_bytes(K * TR_SORT_THREADS, sizeof(S));
        auto k = K == 8u ? tr_band_kernel<S, 8u> : K == 12u ? tr_band_kernel<S, 12u> : K == 16u ? tr_band_kernel<S, 16u>
               : K == 24u ? tr_band_kernel<S, 24u> : tr_band_kernel<S, 20u>;
        CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        const unsigned long long blocks = per * 8ull;
        CSAMD_REQUIRE(blocks <= 0x7fffffffull, "transpose: too many chunks");
        k<<<(unsigned int) blocks, TR_SORT_THREADS, lds, s>>>(cDim, uDim, idx, (const S *) val, tile_off, out_ptr, rbt, order,
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
    // packed kernel: U = 2 / 4 / 8 groups of 4 entries per lane and pass (CSAMD_TR_U for sweeps)
    const char *uenv = std::getenv("CSAMD_TR_U");
    const unsigned int uw = uenv != nullptr ? (unsigned int) std::atoi(uenv) : 8u;
    const size_t packed_lds = count_packed_lds_bytes(uDim, L.bands);
    const bool packed = L.band_table && uDim > 0u && uDim < 65536u && nnz >= 4u && nnz < 0xffff0000u && packed_lds <= (size_t) max_lds &&
                        ((uintptr_t) d_idx & 15u) == 0u;  // 16-byte group loads
    // counts (and band starts) -> per-tile running offsets -> out_ptr
    auto count_pass = [&](bool use_packed, bool with_rbt) -> int {
        if (use_packed) {
            auto k = uw == 2u ? tr_count_packed_kernel<2u> : uw == 8u ? tr_count_packed_kernel<8u> : tr_count_packed_kernel<4u>;
            CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) packed_lds), "set lds");
            k<<<L.tiles, TR_COUNT_THREADS, packed_lds, s>>>(cDim, uDim, nnz, d_ptr, d_idx, L.bands, counts, rbt, flags);
            CSAMD_HIP(hipGetLastError(), "tr_count_packed_kernel");
        } else {
            const int use_lds = (size_t) uDim * 4u <= (size_t) max_lds ? 1 : 0;
            if (!use_lds && !zero_bytes(counts, (size_t) L.tiles * uDim * 4u, s)) return 0;
            const size_t lds = use_lds ? (size_t) uDim * 4u : 0u;
            auto k = tr_count_kernel<16u>;
            if (use_lds) CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            k<<<L.tiles, TR_COUNT_THREADS, lds, s>>>(cDim, uDim, d_ptr, d_idx, L.bands, use_lds, counts, with_rbt ? rbt : nullptr, flags);
            CSAMD_HIP(hipGetLastError(), "tr_count_kernel");
        }
        tr_tile_offsets_kernel<<<grid_for(uDim, 256u, 4096u), 256, 0, s>>>(counts, L.tiles, uDim, coltot);
        CSAMD_HIP(hipGetLastError(), "tr_tile_offsets_kernel");
        return exclusive_scan_u32(coltot, uDim, d_out_ptr, nullptr, scan_tmp, exclusive_scan_scratch_bytes(uDim), s);
    };
    if (!count_pass(packed, true)) return 0;
    unsigned int hflags[2] = {0u, 0u};
    CSAMD_HIP(hipMemcpyAsync(hflags, flags, 8u, hipMemcpyDeviceToHost, s), "D2H transpose flags");
    CSAMD_HIP(hipStreamSynchronize(s), "transpose sync");
    CSAMD_REQUIRE(hflags[1] == 0u, "transpose: minor index out of range (>= %u)", uDim);
    // a row that is not strictly ascending can overflow the packed u16
    // counters: count again in u32 for the fallback scatter
    if (hflags[0] != 0u && packed && !count_pass(false, false)) return 0;
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
