// Row-label masked grouped reduction: per (label g, kept feature f)
// {sum, sumsq, nnz} over kept rows, FP64 partials, deterministic combine.
//
// There is no reference kernel for this operation (mask_groups.cu reduces per
// row over <= 32 feature-group bitmasks); its masking rules are the
// reference's: a row takes part iff row_keep[r] != 0 (mask_groups.cu:65-67),
// a stored entry iff col < cols, v != 0 and feature_keep[col] != 0
// (mask_groups.cu:155).
//
// Design (gfx950):
//   1. kept-feature bitmap + per-word prefix (feature -> compact rank), staged
//      in LDS as {mask, prefix} pairs, one 8-byte LDS read per entry;
//   2. kept rows stably sorted by label (radix sort of (label, row));
//   3. the sorted row list split into one contiguous, nnz-balanced range per
//      persistent workgroup (2 per CU for F <= ~3.3k kept features);
//   4. each workgroup walks its range label segment by label segment; one
//      wave64 streams one row at a time with 16-byte index / 8-byte value
//      vector loads; kept entries land in LDS FP64 accumulators
//      (ds_add_f64) and u32 counters; at each label change the LDS partial
//      is written to slot (workgroup + label);
//   5. a combine kernel sums, per (g, f), the slots of the workgroups that
//      touched label g in ascending workgroup order (FP64), then optional
//      mean/variance.
#include "common.hh"

#include <cstdlib>

namespace csamd {
namespace {

constexpr unsigned int LR_ACC_BYTES = 20u;   // f64 sum + f64 sumsq + u32 nnz per kept feature

struct dev_buf {
    void *p;
    size_t bytes;
};

int ensure_buf(dev_buf &b, size_t bytes, const char *label) {
    if (bytes == 0u) bytes = 16u;
    if (b.p != nullptr && b.bytes >= bytes) return 1;
    if (b.p != nullptr) (void) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0u;
    bytes = align_up(bytes, 4096u);
    if (!hip_ok(hipMalloc(&b.p, bytes), label)) return 0;
    b.bytes = bytes;
    return 1;
}

struct label_private {
    dev_buf fpairs;      // uint2 {mask, prefix} per 32-column word
    dev_buf fids;        // kept feature ids
    dev_buf labels;      // label table [(G + 1) x sort workgroups], gstart [G + 1], F
    dev_buf rows;        // sorted kept rows [rows]
    dev_buf work;        // work weights [rows], tile offsets, ranges [rows]
    dev_buf bounds;      // [W + 1]
    dev_buf partial;     // (W + G) x ftile x {f64, f64, u32}
    dev_buf tail;        // tail records
    dev_buf out;         // outputs
    volatile unsigned int *host;  // pinned: [0] kept rows, [1] kept features
    hipEvent_t ev0, ev1, ev_meta;
    int events_ready;
    int timing;
    int last_valid;
};

// One stored entry past the last whole 4-entry unit, resolved for the
// combine: label g (>= G: contributes nothing), global kept rank f, value.
struct tail_record {
    unsigned int g;
    unsigned int f;
    double v;
};
constexpr unsigned int LR_MAX_TAIL = 3u;

__device__ __forceinline__ float decode_any(const void *val, unsigned long long i, unsigned int vt) {
    switch (vt) {
        case CSAMD_VALUE_F16: return f16_bits_to_float(((const uint16_t *) val)[i]);
        case CSAMD_VALUE_BF16: return bf16_bits_to_float(((const uint16_t *) val)[i]);
        case CSAMD_VALUE_F32: return ((const float *) val)[i];
        default: return fp8_e4m3_to_float(((const uint8_t *) val)[i]);
    }
}

// ---------------------------------------------------------------------------
// Rows grouped by label: a counting sort on key = label (kept rows) or G
// (dropped).  Workgroup b histograms its rows (the tiles b, b + grid, ...)
// in LDS and stores the counts to hist[g * grid + b]; one workgroup scans
// that label-major table into each (label, workgroup) pair's first slot;
// the scatter re-reads the same tiles and places rows.  No global atomics:
// device-scope atomics on a few hot label counters serialize.
// ---------------------------------------------------------------------------
constexpr unsigned int LR_SORT_THREADS = 1024u;
constexpr unsigned int LR_SORT_ITEMS = 8u;
constexpr unsigned int LR_SORT_TILE = LR_SORT_THREADS * LR_SORT_ITEMS;

__device__ __forceinline__ unsigned int row_key(const unsigned int *__restrict__ labels, const unsigned char *__restrict__ row_keep,
                                                unsigned int r, unsigned int G) {
    const unsigned int l = labels[r];
    return ((row_keep == nullptr || row_keep[r] != 0u) && l < G) ? l : G;
}

// Also: clears work[0..work_n) (positions past the kept rows keep weight 0
// for the partition scan); builds the kept-feature mask of each 32-column
// word; clears the tail records (label G + 1: contributes nothing) that the
// scatter fills for kept rows.
__global__ void __launch_bounds__(LR_SORT_THREADS) label_count_kernel(const unsigned int *__restrict__ labels,
                                                                      const unsigned char *__restrict__ row_keep, unsigned int rows,
                                                                      unsigned int G, unsigned int *__restrict__ hist_out,
                                                                      unsigned int *__restrict__ work, unsigned int work_n,
                                                                      const unsigned char *__restrict__ feature_keep, unsigned int cols,
                                                                      unsigned int nwords, unsigned int *__restrict__ word_mask,
                                                                      tail_record *__restrict__ tail) {
    extern __shared__ unsigned int hist[];  // [G + 1]
    for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) hist[g] = 0u;
    if (tail != nullptr && blockIdx.x == 0u && threadIdx.x < LR_MAX_TAIL)
        tail[threadIdx.x] = tail_record{G + 1u, CSAMD_INVALID_COL, 0.0};
    // Feature words, 64 per wave: pass j reads 64 consecutive columns (one
    // coalesced load), a ballot turns them into words 2j and 2j + 1, kept
    // by lanes 2j and 2j + 1.
    const unsigned int lane = lane_id();
    const unsigned int nwave = (gridDim.x * blockDim.x) >> 6;
    for (unsigned int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; gw * 64u < nwords; gw += nwave) {
        const unsigned int w0 = gw * 64u;
        unsigned char v[32];
#pragma unroll
        for (unsigned int j = 0; j < 32u; ++j) {
            const unsigned int c = (w0 + 2u * j) * 32u + lane;
            v[j] = feature_keep == nullptr ? (unsigned char) 1u : feature_keep[min(c, cols - 1u)];
        }
        unsigned int m = 0u;
#pragma unroll
        for (unsigned int j = 0; j < 32u; ++j) {
            const unsigned int c = (w0 + 2u * j) * 32u + lane;
            const unsigned long long b = __ballot(c < cols && v[j] != 0u);
            if (lane == 2u * j) m = (unsigned int) b;
            if (lane == 2u * j + 1u) m = (unsigned int) (b >> 32);
        }
        if (w0 + lane < nwords) word_mask[w0 + lane] = m;
    }
    __syncthreads();
    for (unsigned int t0 = blockIdx.x * LR_SORT_TILE; t0 < rows; t0 += gridDim.x * LR_SORT_TILE) {
        unsigned int key[LR_SORT_ITEMS];
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) {
            const unsigned int r = t0 + k * LR_SORT_THREADS + threadIdx.x;
            key[k] = r < rows ? row_key(labels, row_keep, r, G) : G + 1u;
        }
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k)
            if (key[k] <= G) atomicAdd(&hist[key[k]], 1u);
    }
    const unsigned int stride = gridDim.x * blockDim.x;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < work_n; i += stride) work[i] = 0u;
    __syncthreads();
    for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) hist_out[(unsigned long long) g * gridDim.x + blockIdx.x] = hist[g];
}

// One workgroup: exclusive scan of the label-major table hist[(G + 1) x nb]
// in place; gstart[g] = hist[g * nb] (gstart[G] = kept rows).  Then the
// feature map: per word {kept mask, kept features before the word}, kept
// feature ids in column order, F.
constexpr unsigned int LR_OFF_ITEMS = 8u;

__global__ void __launch_bounds__(1024) label_offsets_kernel(unsigned int *__restrict__ hist, unsigned int G, unsigned int nb,
                                                             unsigned int *__restrict__ gstart,
                                                             const unsigned int *__restrict__ word_mask, unsigned int nwords,
                                                             uint2 *__restrict__ pairs, unsigned int *__restrict__ fids,
                                                             unsigned int *__restrict__ f_total) {
    __shared__ unsigned int smem[16];
    unsigned int fcarry = 0u;  // same value in every thread
    for (unsigned int start = 0; start < nwords; start += blockDim.x) {
        const unsigned int w = start + threadIdx.x;
        const unsigned int m = w < nwords ? word_mask[w] : 0u;
        unsigned int total = 0u;
        const unsigned int ex = fcarry + block_exclusive_scan_u32((unsigned int) __popc(m), smem, &total);
        if (w < nwords) {
            pairs[w] = make_uint2(m, ex);
            unsigned int mm = m, k = 0u;
            while (mm != 0u) {
                fids[ex + k++] = w * 32u + (unsigned int) __builtin_ctz(mm);
                mm &= mm - 1u;
            }
        }
        fcarry += total;
    }
    if (threadIdx.x == 0) *f_total = fcarry;
    const unsigned long long n = (unsigned long long) (G + 1u) * nb;
    unsigned int carry = 0u;  // same value in every thread
    for (unsigned long long start = 0; start < n; start += (unsigned long long) blockDim.x * LR_OFF_ITEMS) {
        const unsigned long long i0 = start + (unsigned long long) threadIdx.x * LR_OFF_ITEMS;
        unsigned int v[LR_OFF_ITEMS], own = 0u;
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
            v[k] = i0 + k < n ? hist[i0 + k] : 0u;
            own += v[k];
        }
        unsigned int total = 0u;
        unsigned int o = carry + block_exclusive_scan_u32(own, smem, &total);
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
            if (i0 + k < n) hist[i0 + k] = o;
            o += v[k];
        }
        carry += total;
    }
    __syncthreads();  // the block's table stores are visible to the block
    for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) gstart[g] = hist[(unsigned long long) g * nb];
}

// ---------------------------------------------------------------------------
// Per-layout row geometry.
// ---------------------------------------------------------------------------
struct csr_geom {
    const unsigned int *ptr;
    __device__ __forceinline__ void range(unsigned int row, unsigned long long *b, unsigned long long *e) const {
        *b = ptr[row];
        *e = ptr[row + 1u];
    }
    // Row holding entry i: the last row r with ptr[r] <= i (rows if none).
    __device__ __forceinline__ unsigned int row_of(unsigned long long i, unsigned int rows) const {
        unsigned int lo = 0u, hi = rows;  // first r with ptr[r + 1] > i
        while (lo < hi) {
            const unsigned int mid = lo + ((hi - lo) >> 1);
            if ((unsigned long long) ptr[mid + 1u] <= i) lo = mid + 1u;
            else hi = mid;
        }
        return lo;
    }
};

struct sell_geom {
    csamd_sliced_ell_view s;
    __device__ __forceinline__ void range(unsigned int row, unsigned long long *b, unsigned long long *e) const {
        if (s.slice_count == 0u) {
            *b = *e = 0ull;
            return;
        }
        unsigned int slice;
        if (s.slice_rows == 32u) slice = row >> 5;
        else if (s.slice_rows != 0u) slice = row / s.slice_rows;
        else {
            unsigned int lo = 0u, hi = s.slice_count - 1u;
            while (lo < hi) {
                const unsigned int mid = (lo + hi + 1u) >> 1;
                if (s.slice_row_offsets[mid] <= row) lo = mid;
                else hi = mid - 1u;
            }
            slice = lo;
        }
        if (slice >= s.slice_count) slice = s.slice_count - 1u;
        const unsigned int w = s.slice_widths[slice];
        const unsigned long long base =
            (unsigned long long) s.slice_slot_offsets[slice] + (unsigned long long) (row - s.slice_row_offsets[slice]) * w;
        *b = base;
        *e = base + w;
    }
    // Row holding slot i (rows if none): its slice, then its row in the slice.
    __device__ __forceinline__ unsigned int row_of(unsigned long long i, unsigned int rows) const {
        if (s.slice_count == 0u) return rows;
        unsigned int lo = 0u, hi = s.slice_count - 1u;  // last slice with slot offset <= i
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if ((unsigned long long) s.slice_slot_offsets[mid] <= i) lo = mid;
            else hi = mid - 1u;
        }
        const unsigned int w = s.slice_widths[lo];
        if (w == 0u) return rows;
        const unsigned long long r = (unsigned long long) s.slice_row_offsets[lo] + (i - s.slice_slot_offsets[lo]) / w;
        return r < rows ? (unsigned int) r : rows;
    }
};

// Blocked-ELL rows all hold ell_cols slots (ranges are not used).
struct bell_geom {
    unsigned int ell_cols;
    __device__ __forceinline__ void range(unsigned int, unsigned long long *b, unsigned long long *e) const {
        *b = 0ull;
        *e = ell_cols;
    }
};

// Scatter kept rows to their label's slots: workgroup b's rows of label g
// start at the scanned table entry (g, b); within a tile, ranks come from
// LDS atomics, so the order of rows within a label is not fixed (sums are
// order independent up to FP64 rounding; counts are exact).  For each placed row
// it also writes its entry range (both layouts index entries with u32),
// clipped at entry clip -- the vector-load kernel streams only whole 4-entry
// units; a kept row holding some of the at most 3 entries past the last
// whole unit writes them as tail records -- and its work weight for the
// partition (entries + 8 for the per-row cost).
template<typename Geom>
__global__ void __launch_bounds__(LR_SORT_THREADS) label_scatter_kernel(Geom geom, const unsigned int *__restrict__ labels,
                                                                        const unsigned char *__restrict__ row_keep, unsigned int rows,
                                                                        unsigned int G, const unsigned int *__restrict__ base_table,
                                                                        unsigned long long clip, unsigned int *__restrict__ sorted_rows,
                                                                        uint2 *__restrict__ ranges, unsigned int *__restrict__ work,
                                                                        tail_record *__restrict__ tail, const unsigned int *__restrict__ idx,
                                                                        const void *val, unsigned int vt, unsigned int cols,
                                                                        const uint2 *__restrict__ fpairs) {
    extern __shared__ unsigned int hist[];  // [G + 1] tile counts, then [G + 1] running bases
    unsigned int *base = hist + (G + 1u);
    for (unsigned int g = threadIdx.x; g < G; g += blockDim.x) base[g] = base_table[(unsigned long long) g * gridDim.x + blockIdx.x];
    for (unsigned int t0 = blockIdx.x * LR_SORT_TILE; t0 < rows; t0 += gridDim.x * LR_SORT_TILE) {
        for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) hist[g] = 0u;
        __syncthreads();
        unsigned int key[LR_SORT_ITEMS], rank[LR_SORT_ITEMS];
        unsigned long long rb[LR_SORT_ITEMS], re[LR_SORT_ITEMS];
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) {
            const unsigned int r = t0 + k * LR_SORT_THREADS + threadIdx.x;
            key[k] = r < rows ? row_key(labels, row_keep, r, G) : G;
            rb[k] = re[k] = 0ull;
            if (r < rows) geom.range(r, &rb[k], &re[k]);
        }
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) rank[k] = key[k] < G ? atomicAdd(&hist[key[k]], 1u) : 0u;
        __syncthreads();
        __syncthreads();
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) {
            if (key[k] >= G) continue;
            const unsigned int r = t0 + k * LR_SORT_THREADS + threadIdx.x;
            const unsigned int pos = base[key[k]] + rank[k];
            const unsigned long long b = rb[k], e = re[k];
            const unsigned long long len = e - b;
            sorted_rows[pos] = r;
            ranges[pos] = make_uint2((unsigned int) (b < clip ? b : clip), (unsigned int) (e < clip ? e : clip));
            work[pos] = (unsigned int) (len > 0xfffffff0ull ? 0xfffffff0ull : len) + 8u;
            if (tail != nullptr && e > clip) {  // this row holds tail entries
                for (unsigned long long i = b > clip ? b : clip; i < e; ++i) {
                    const unsigned int c = idx[i];
                    tail_record rec{G + 1u, CSAMD_INVALID_COL, 0.0};
                    if (c < cols) {
                        const uint2 pm = fpairs[c >> 5];
                        const unsigned int bit = c & 31u;
                        if ((pm.x >> bit) & 1u) {
                            rec.g = key[k];
                            rec.f = pm.y + (unsigned int) __popc(pm.x & ((1u << bit) - 1u));
                            rec.v = (double) decode_any(val, i, vt);
                        }
                    }
                    tail[i - clip] = rec;
                }
            }
        }
        __syncthreads();
        for (unsigned int g = threadIdx.x; g < G; g += blockDim.x) base[g] += hist[g];
        __syncthreads();
    }
}

// Partition bounds from the work weights: bounds[w] = first sorted position
// p with prefix(p) >= floor(total * w / W), prefix(p) = sum of work[0..p).
// Each thread scans its SCAN_ITEMS positions of a CSAMD_SCAN_TILE tile (tile
// offsets precomputed) and writes the bounds that fall in its positions:
// position q + 1 takes every w with prefix(q) < target(w) <= prefix(q + 1);
// position 0 takes every w with target(w) == 0.  Positions at or past the
// kept rows have weight 0 and take none, except through the total: the
// last kept row's interval ends at total = target(W).
constexpr unsigned int LR_BOUNDS_ITEMS = CSAMD_SCAN_TILE / 256u;

__global__ void __launch_bounds__(256) work_bounds_kernel(const unsigned int *__restrict__ work, unsigned int n,
                                                         const unsigned long long *__restrict__ tile_offsets, unsigned int tiles,
                                                         unsigned int W, unsigned int *__restrict__ bounds) {
    __shared__ unsigned long long smem[4];
    const unsigned long long total = tile_offsets[tiles];
    const unsigned long long first = (unsigned long long) blockIdx.x * CSAMD_SCAN_TILE + (unsigned long long) threadIdx.x * LR_BOUNDS_ITEMS;
    unsigned int v[LR_BOUNDS_ITEMS];
    unsigned long long own = 0ull;
#pragma unroll
    for (unsigned int i = 0; i < LR_BOUNDS_ITEMS; ++i) {
        v[i] = first + i < n ? work[first + i] : 0u;
        own += v[i];
    }
    unsigned long long blk = 0ull;
    unsigned long long pre = tile_offsets[blockIdx.x] + block_exclusive_scan_u64(own, smem, &blk);
    if (total == 0ull) {  // no kept rows: every range is empty
        if (blockIdx.x == 0u)
            for (unsigned int w = threadIdx.x; w <= W; w += blockDim.x) bounds[w] = 0u;
        return;
    }
    if (first == 0ull) {
        const unsigned long long wmax = (W - 1ull) / total;  // target(w) == 0 for w <= wmax
        for (unsigned long long w = 0; w <= wmax && w <= W; ++w) bounds[w] = 0u;
    }
    // w whose target lies in this thread's span (pre, pre + own]: usually none
    const unsigned long long w0 = ((pre + 1ull) * W + total - 1ull) / total;
    const unsigned long long w1 = ((pre + own + 1ull) * W - 1ull) / total;
    for (unsigned long long w = w0; w <= w1 && w <= W; ++w) {
        const unsigned long long target = total * w / W;
        unsigned long long hi = pre;
        unsigned int i = 0u;
        while (i < LR_BOUNDS_ITEMS && hi + v[i] < target) hi += v[i++];  // first item whose end reaches target
        bounds[w] = (unsigned int) (first + i + 1ull);
    }
}

// ---------------------------------------------------------------------------
// Main kernel.
// ---------------------------------------------------------------------------
struct acc_lds {
    double *sum;
    double *sumsq;
    unsigned int *cnt;
};

struct reduce_args {
    const unsigned int *sorted_rows;
    const unsigned int *gstart;   // [G+1]
    const unsigned int *bounds;   // [W+1]
    const uint2 *fpairs;          // [nwords]
    const unsigned int *f_total;  // kept features F (device)
    double *psum;                 // [tiles x (W+G) x ftile]
    double *psumsq;
    unsigned int *pcnt;
    unsigned int G;
    unsigned int cols;
    unsigned int nwords;
    unsigned int ftile;           // features held in LDS per pass
    unsigned int f0;              // first feature of this pass (set per pass)
    unsigned int f_count;         // features in this pass (set per pass)
};

// The workgroup's sorted-row range [bounds[b], bounds[b + 1]), walked one
// feature tile (LDS pass) at a time and, within it, one label segment at a
// time: zero the LDS accumulators, run body(segment start, end, tile args),
// store the accumulators to partial slot (tile, workgroup + label).  F is
// read on the device, so the host launches one kernel for all tiles.
template<typename Body>
__device__ __forceinline__ void walk_segments(const reduce_args &a, const acc_lds &acc, Body body) {
    const unsigned int lo = a.bounds[blockIdx.x];
    const unsigned int hi = a.bounds[blockIdx.x + 1u];
    if (lo >= hi) return;  // uniform across the block
    unsigned int g0;
    {
        unsigned int l = 0u, h = a.G;
        while (l < h) {  // largest g with gstart[g] <= lo
            const unsigned int mid = (l + h + 1u) >> 1;
            if (a.gstart[mid] <= lo) l = mid;
            else h = mid - 1u;
        }
        g0 = l;
    }
    const unsigned int F = *a.f_total;
    const unsigned long long tile_slots = (unsigned long long) (gridDim.x + a.G) * a.ftile;
    reduce_args t = a;
    for (unsigned int tile = 0u; tile * a.ftile < F; ++tile) {
        t.f0 = tile * a.ftile;
        t.f_count = F - t.f0 < a.ftile ? F - t.f0 : a.ftile;
        for (unsigned int g = g0; g < a.G; ++g) {
            const unsigned int gs = a.gstart[g], ge = a.gstart[g + 1u];
            const unsigned int s0 = gs > lo ? gs : lo;
            const unsigned int s1 = ge < hi ? ge : hi;
            if (s0 >= hi) break;
            if (s0 >= s1) continue;
            for (unsigned int f = threadIdx.x; f < t.f_count; f += blockDim.x) {
                acc.sum[f] = 0.0;
                acc.sumsq[f] = 0.0;
                acc.cnt[f] = 0u;
            }
            body(s0, s1, (const reduce_args &) t);
            __syncthreads();
            const unsigned long long slot = tile * tile_slots + (unsigned long long) (blockIdx.x + g) * a.ftile;
            for (unsigned int f = threadIdx.x; f < t.f_count; f += blockDim.x) {
                a.psum[slot + f] = acc.sum[f];
                a.psumsq[slot + f] = acc.sumsq[f];
                a.pcnt[slot + f] = acc.cnt[f];
            }
            __syncthreads();
        }
    }
}

template<int VT>
__device__ __forceinline__ void consume(unsigned int c, typename value_traits<VT>::storage raw, const uint2 *__restrict__ lpairs,
                                        const reduce_args &a, const acc_lds &acc) {
    if (c >= a.cols) return;
    const uint2 pm = lpairs[c >> 5];
    const unsigned int bit = c & 31u;
    if (((pm.x >> bit) & 1u) == 0u) return;
    const unsigned int f = pm.y + (unsigned int) __popc(pm.x & ((1u << bit) - 1u)) - a.f0;
    if (f >= a.f_count) return;
    const float v = value_traits<VT>::decode(raw);
    if (v != 0.0f) {
        const double d = (double) v;
        atomicAdd(acc.sum + f, d);
        atomicAdd(acc.sumsq + f, d * d);
        atomicAdd(acc.cnt + f, 1u);
    }
}

// A unit is 4 consecutive entries, loaded with one non-temporal vector load
// of indices and one of values.  Loaded words stay packed in registers until
// they are consumed: any instruction that reads them at load time would make
// the wave wait for the load right there and void the prefetch.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template<int VT> struct unit_vals;
template<> struct unit_vals<CSAMD_VALUE_F16> {
    using raw = u32x2;
    static __device__ __forceinline__ raw vload(const uint16_t *v, unsigned int i0) {
        return __builtin_nontemporal_load((const u32x2 *) (v + i0));
    }
    static __device__ __forceinline__ raw sload(const uint16_t *v, unsigned int i0, unsigned int last) {
        raw r;
        r.x = (unsigned int) v[min(i0, last)] | ((unsigned int) v[min(i0 + 1u, last)] << 16);
        r.y = (unsigned int) v[min(i0 + 2u, last)] | ((unsigned int) v[min(i0 + 3u, last)] << 16);
        return r;
    }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        const unsigned int w = k < 2u ? r.x : r.y;
        return f16_bits_to_float((uint16_t) ((k & 1u) ? (w >> 16) : (w & 0xffffu)));
    }
};
template<> struct unit_vals<CSAMD_VALUE_BF16> {
    using raw = u32x2;
    static __device__ __forceinline__ raw vload(const uint16_t *v, unsigned int i0) {
        return __builtin_nontemporal_load((const u32x2 *) (v + i0));
    }
    static __device__ __forceinline__ raw sload(const uint16_t *v, unsigned int i0, unsigned int last) {
        return unit_vals<CSAMD_VALUE_F16>::sload(v, i0, last);
    }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        const unsigned int w = k < 2u ? r.x : r.y;
        return __uint_as_float((k & 1u) ? (w & 0xffff0000u) : (w << 16));
    }
};
template<> struct unit_vals<CSAMD_VALUE_F32> {
    using raw = u32x4;
    static __device__ __forceinline__ raw vload(const float *v, unsigned int i0) {
        return __builtin_nontemporal_load((const u32x4 *) (v + i0));
    }
    static __device__ __forceinline__ raw sload(const float *v, unsigned int i0, unsigned int last) {
        raw r;
        r.x = __float_as_uint(v[min(i0, last)]);
        r.y = __float_as_uint(v[min(i0 + 1u, last)]);
        r.z = __float_as_uint(v[min(i0 + 2u, last)]);
        r.w = __float_as_uint(v[min(i0 + 3u, last)]);
        return r;
    }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        return __uint_as_float(k == 0u ? r.x : k == 1u ? r.y : k == 2u ? r.z : r.w);
    }
};
template<> struct unit_vals<CSAMD_VALUE_FP8_E4M3> {
    using raw = unsigned int;
    static __device__ __forceinline__ raw vload(const uint8_t *v, unsigned int i0) {
        return __builtin_nontemporal_load((const unsigned int *) (v + i0));
    }
    static __device__ __forceinline__ raw sload(const uint8_t *v, unsigned int i0, unsigned int last) {
        return (unsigned int) v[min(i0, last)] | ((unsigned int) v[min(i0 + 1u, last)] << 8) |
               ((unsigned int) v[min(i0 + 2u, last)] << 16) | ((unsigned int) v[min(i0 + 3u, last)] << 24);
    }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) { return fp8_e4m3_decode_hw(r, (int) k); }
};

constexpr unsigned int LR_UNITS = 4u;               // 4-entry units per lane per chunk (even)
constexpr unsigned int LR_CHUNK_UNITS = 64u * LR_UNITS;

template<int VT> struct chunk_regs {
    u32x4 c[LR_UNITS];
    typename unit_vals<VT>::raw v[LR_UNITS];
};

// Issue the loads of one chunk: unit u + lane + 64 j of the current row.
// Lanes past the row's last unit u_end - 1 re-read that unit (same cache
// lines, masked when consumed), so every lane loads and no lane branches.
// VEC: row ranges end at or before the last whole 4-entry unit of the
// arrays, so every unit is one 16-byte index load and one value load, and
// the loop carries a fixed number of loads per chunk (which lets the
// compiler wait for the current chunk only).  Otherwise (unaligned arrays)
// single entries are read, clamped to index last.
template<int VT, bool VEC>
__device__ __forceinline__ void load_chunk(const unsigned int *__restrict__ idx, const typename value_traits<VT>::storage *__restrict__ val,
                                           unsigned int u, unsigned int u_end, unsigned int last, chunk_regs<VT> &r) {
    const unsigned int lane = lane_id();
    if (VEC) {
#pragma unroll
        for (unsigned int j = 0; j < LR_UNITS; ++j) {
            const unsigned int uu = min(u + lane + 64u * j, u_end - 1u);
            r.c[j] = __builtin_nontemporal_load((const u32x4 *) (idx + uu * 4u));
            r.v[j] = unit_vals<VT>::vload(val, uu * 4u);
        }
    } else {
#pragma unroll
        for (unsigned int j = 0; j < LR_UNITS; ++j) {
            const unsigned int i0 = min(u + lane + 64u * j, u_end - 1u) * 4u;
            r.c[j].x = idx[min(i0, last)];
            r.c[j].y = idx[min(i0 + 1u, last)];
            r.c[j].z = idx[min(i0 + 2u, last)];
            r.c[j].w = idx[min(i0 + 3u, last)];
            r.v[j] = unit_vals<VT>::sload(val, i0, last);
        }
    }
}

__device__ __forceinline__ unsigned int lane_of(const u32x4 &w, unsigned int k) {
    return k == 0u ? w.x : k == 1u ? w.y : k == 2u ? w.z : w.w;
}

// Consume one chunk of the row whose entries are [b, e), starting at unit u,
// two units (8 entries per lane) at a time.  Per pair, the 8 bitmap lookups
// are issued first
// (independent LDS reads); keep tests, ranks and values are computed
// without branches; one branch per entry guards the three accumulator
// atomics.  Nothing here reads global memory, so the prefetched chunk stays
// in flight (a feature scale is applied by the combine).  lpairs[nwords] is
// a {0, 0} sentinel, so columns >= cols (and CSAMD_INVALID_COL) are "not
// kept".  Entries outside [b, e) exist only in a row's first and last
// chunk; only those chunks (a wave-uniform test) pay for the range mask.
// Every loaded unit is consumed: a load whose registers are never read is
// waited for wherever the registers are next written.
template<int VT>
__device__ __forceinline__ void consume_chunk(const chunk_regs<VT> &r, unsigned int u, unsigned int b, unsigned int e,
                                              const uint2 *__restrict__ lpairs, const reduce_args &a, const acc_lds &acc) {
    const unsigned long long first = 4ull * u;
    const bool edge = first < b || first + 4ull * LR_CHUNK_UNITS > e;
    const unsigned int lane = lane_id();
    const unsigned int len = e - b;
#pragma unroll
    for (unsigned int p = 0; p < LR_UNITS / 2u; ++p) {
        unsigned int c[8];
#pragma unroll
        for (unsigned int k = 0; k < 8u; ++k) c[k] = lane_of(r.c[2u * p + (k >> 2)], k & 3u);
        if (edge) {
#pragma unroll
            for (unsigned int k = 0; k < 8u; ++k) {
                const unsigned int i = (u + lane + 64u * (2u * p + (k >> 2))) * 4u + (k & 3u);
                if (i - b >= len) c[k] = CSAMD_INVALID_COL;
            }
        }
        uint2 pm[8];
#pragma unroll
        for (unsigned int k = 0; k < 8u; ++k) pm[k] = lpairs[min(c[k] >> 5, a.nwords)];
#pragma unroll
        for (unsigned int k = 0; k < 8u; ++k) {
            const unsigned int bit = c[k] & 31u;
            const unsigned int f = pm[k].y + (unsigned int) __popc(pm[k].x & ((1u << bit) - 1u)) - a.f0;
            const float v = unit_vals<VT>::get(r.v[2u * p + (k >> 2)], k & 3u);
            const bool kept = ((pm[k].x >> bit) & 1u) != 0u && f < a.f_count && v != 0.0f;
            if (kept) {
                const double d = (double) v;
                atomicAdd(acc.sum + f, d);
                atomicAdd(acc.sumsq + f, d * d);
                atomicAdd(acc.cnt + f, 1u);
            }
        }
    }
}

__device__ __forceinline__ void row_units(uint2 rng, unsigned int *u, unsigned int *ue) {
    *u = rng.x >> 2;
    *ue = rng.x == rng.y ? *u : (unsigned int) (((unsigned long long) rng.y + 3ull) >> 2);
}

// Next row of the label segment for this wave: one LDS atomic per row, so a
// wave that runs faster (wave age decides issue priority) takes more rows
// and all waves of the workgroup finish the segment together.
__device__ __forceinline__ unsigned int grab_row(unsigned int *next) {
    unsigned int v = 0u;
    if (lane_id() == 0u) v = atomicAdd(next, 1u);
    return __builtin_amdgcn_readfirstlane(v);
}

// Position of a chunk in the wave's row stream (wave-uniform values).
struct row_cursor {
    uint2 rng;        // entry range [x, y) of the row
    unsigned int u;   // first unit of the chunk
    unsigned int ue;  // end unit of the row
};

// Rows of one label segment [.., s1), taken one at a time from the shared
// row counter.  The chunk stream runs through two fixed register buffers in
// turn (A, B, A, ...): the next chunk, possibly of the next row, is loaded
// before the current one is consumed, with no register copies between
// them, and the next row is taken and its range fetched one row ahead.
struct row_stream {
    const uint2 *ranges;
    unsigned int *next;
    unsigned int s1;
    unsigned int inext;  // row taken ahead (>= s1: none left)
    uint2 nxt;           // its range

    __device__ __forceinline__ void take() {
        inext = grab_row(next);
        nxt = inext < s1 ? ranges[inext] : make_uint2(0u, 0u);
    }
    // Move to the first non-empty row taken ahead; false when none is left.
    __device__ __forceinline__ bool next_row(row_cursor &c) {
        do {
            if (inext >= s1) return false;
            c.rng = nxt;
            take();
            row_units(c.rng, &c.u, &c.ue);
        } while (c.u >= c.ue);
        return true;
    }
    __device__ __forceinline__ bool advance(row_cursor &c) {
        const unsigned int u2 = c.u + LR_CHUNK_UNITS;
        if (u2 < c.ue) {
            c.u = u2;
            return true;
        }
        return next_row(c);
    }
};

template<int VT, bool VEC>
__device__ __forceinline__ void stream_rows(const unsigned int *__restrict__ idx, const typename value_traits<VT>::storage *__restrict__ val,
                                            const uint2 *__restrict__ ranges, unsigned int *next, unsigned int s1,
                                            unsigned int last, const uint2 *__restrict__ lpairs,
                                            const reduce_args &a, const acc_lds &acc) {
    row_stream st{ranges, next, s1, 0u, make_uint2(0u, 0u)};
    st.take();
    row_cursor ca, cb;
    if (!st.next_row(ca)) return;
    // Loads are issued unconditionally (after the last chunk the wave
    // re-loads it, in bounds and unused): a load under a branch makes the
    // compiler merge old and new buffer contents, which waits for the load.
    chunk_regs<VT> A, B;
    load_chunk<VT, VEC>(idx, val, ca.u, ca.ue, last, A);
    while (true) {
        cb = ca;
        bool more = st.advance(cb);
        if (!more) cb = ca;
        load_chunk<VT, VEC>(idx, val, cb.u, cb.ue, last, B);
        consume_chunk<VT>(A, ca.u, ca.rng.x, ca.rng.y, lpairs, a, acc);
        if (!more) break;
        ca = cb;
        more = st.advance(ca);
        if (!more) ca = cb;
        load_chunk<VT, VEC>(idx, val, ca.u, ca.ue, last, A);
        consume_chunk<VT>(B, cb.u, cb.rng.x, cb.rng.y, lpairs, a, acc);
        if (!more) break;
    }
}

template<int VT, bool VEC>
__global__ void label_reduce_rows_kernel(const unsigned int *__restrict__ idx, const typename value_traits<VT>::storage *__restrict__ val,
                                         unsigned long long limit, const uint2 *__restrict__ ranges, reduce_args a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int next_row;
    const unsigned int nthreads = blockDim.x;
    uint2 *lpairs = (uint2 *) lds;
    double *lsum = (double *) (lds + align_up((size_t) (a.nwords + 1u) * sizeof(uint2), 16u));
    double *lsumsq = lsum + a.ftile;
    unsigned int *lcnt = (unsigned int *) (lsumsq + a.ftile);
    const acc_lds acc{lsum, lsumsq, lcnt};
    const unsigned int last = limit != 0ull ? (unsigned int) (limit - 1ull) : 0u;

    for (unsigned int w = threadIdx.x; w <= a.nwords; w += nthreads) lpairs[w] = w < a.nwords ? a.fpairs[w] : make_uint2(0u, 0u);
    walk_segments(a, acc, [&](unsigned int s0, unsigned int s1, const reduce_args &t) {
        if (threadIdx.x == 0u) next_row = s0;
        __syncthreads();  // also publishes the zeroed accumulators
        stream_rows<VT, VEC>(idx, val, ranges, &next_row, s1, last, lpairs, t, acc);
    });
}

// Blocked-ELL rows: entries decoded through blockColIdx (scalar lane stride).
template<int VT>
__global__ void label_reduce_blocked_kernel(csamd_blocked_ell_view src, reduce_args a) {
    using S = typename value_traits<VT>::storage;
    extern __shared__ __align__(16) unsigned char lds[];
    const unsigned int nthreads = blockDim.x;
    const unsigned int nwaves = nthreads >> 6;
    const unsigned int wave = threadIdx.x >> 6;
    uint2 *lpairs = (uint2 *) lds;
    double *lsum = (double *) (lds + align_up((size_t) (a.nwords + 1u) * sizeof(uint2), 16u));
    double *lsumsq = lsum + a.ftile;
    unsigned int *lcnt = (unsigned int *) (lsumsq + a.ftile);
    const acc_lds acc{lsum, lsumsq, lcnt};
    const S *__restrict__ val = (const S *) src.val;
    const unsigned int bs = src.block_size;
    const unsigned int width = bs != 0u ? src.ell_cols / bs : 0u;

    for (unsigned int w = threadIdx.x; w < a.nwords; w += nthreads) lpairs[w] = a.fpairs[w];
    walk_segments(a, acc, [&](unsigned int s0, unsigned int s1, const reduce_args &t) {
        __syncthreads();  // publishes the zeroed accumulators
        for (unsigned int i = s0 + wave; i < s1; i += nwaves) {
            const unsigned int row = t.sorted_rows[i];
            const unsigned int rb = bs != 0u ? row / bs : 0u;
            const unsigned long long base = (unsigned long long) row * src.ell_cols;
            for (unsigned int e = lane_id(); e < src.ell_cols; e += 64u) {
                const unsigned int slot = e / bs;
                const unsigned int bc = src.blockColIdx[(unsigned long long) rb * width + slot];
                const unsigned int c = bc != CSAMD_INVALID_COL ? bc * bs + (e - slot * bs) : CSAMD_INVALID_COL;
                consume<VT>(c, val[base + e], lpairs, t, acc);
            }
        }
    });
}

// Finish: out[g][f] = sum, over the workgroups w whose range holds rows of
// label g (ascending w), of partial slot (tile of f, w + g); then the tail
// records; then the feature's scale s (if any): s * sum, s^2 * sumsq, no
// nonzeros when s == 0; then mean = sum / n_g and var = sumsq / n_g - mean^2
// (clamped at 0), n_g = kept rows of label g.  F is read on the device;
// gstart and bounds are staged in LDS for the per-thread search.
__global__ void __launch_bounds__(256) label_finish_kernel(
    const double *__restrict__ psum, const double *__restrict__ psumsq, const unsigned int *__restrict__ pcnt,
    const unsigned int *__restrict__ gstart_g, const unsigned int *__restrict__ bounds_g, unsigned int W, unsigned int G,
    unsigned int ftile, const unsigned int *__restrict__ f_total, const float *__restrict__ scale,
    const unsigned int *__restrict__ fids, const tail_record *__restrict__ tail, double *__restrict__ out_sum,
    double *__restrict__ out_sumsq, unsigned int *__restrict__ out_cnt, double *__restrict__ mean, double *__restrict__ var,
    unsigned int *__restrict__ group_rows) {
    extern __shared__ unsigned int meta[];  // gstart[G + 1], bounds[W + 1]
    unsigned int *gstart = meta;
    unsigned int *bounds = meta + G + 1u;
    for (unsigned int i = threadIdx.x; i <= G; i += blockDim.x) gstart[i] = gstart_g[i];
    for (unsigned int i = threadIdx.x; i <= W; i += blockDim.x) bounds[i] = bounds_g[i];
    __syncthreads();
    const unsigned int F = *f_total;
    const unsigned long long tile_slots = (unsigned long long) (W + G) * ftile;
    const unsigned long long total = (unsigned long long) G * F;
    for (unsigned long long t = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; t < total;
         t += (unsigned long long) gridDim.x * blockDim.x) {
        const unsigned int g = (unsigned int) (t / F);
        const unsigned int f = (unsigned int) (t - (unsigned long long) g * F);
        const unsigned int tile = f / ftile;
        const unsigned long long fbase = tile * tile_slots + (f - tile * ftile);
        const unsigned int gs = gstart[g], ge = gstart[g + 1u];
        double sm = 0.0, q = 0.0;
        unsigned int n = 0u;
        if (gs < ge) {
            unsigned int l = 0u, h = W - 1u;  // first w with bounds[w + 1] > gs
            while (l < h) {
                const unsigned int mid = (l + h) >> 1;
                if (bounds[mid + 1u] > gs) h = mid;
                else l = mid + 1u;
            }
            unsigned int r = l;  // one past the last w with bounds[w] < ge
            while (r < W && bounds[r] < ge) ++r;
            for (unsigned int w0 = l; w0 < r; w0 += 4u) {
                double ps[4], pq[4];
                unsigned int pc[4];
#pragma unroll
                for (unsigned int k = 0; k < 4u; ++k) {  // loads first, then sums in w order
                    const unsigned int w = w0 + k < r ? w0 + k : l;  // clamped: never reads past the slots
                    const bool used = w0 + k < r && bounds[w] < bounds[w + 1u];
                    const unsigned long long slot = fbase + (unsigned long long) (w + g) * ftile;
                    ps[k] = used ? psum[slot] : 0.0;
                    pq[k] = used ? psumsq[slot] : 0.0;
                    pc[k] = used ? pcnt[slot] : 0u;
                }
#pragma unroll
                for (unsigned int k = 0; k < 4u; ++k) {
                    sm += ps[k];
                    q += pq[k];
                    n += pc[k];
                }
            }
        }
        if (tail != nullptr) {
#pragma unroll
            for (unsigned int k = 0; k < LR_MAX_TAIL; ++k)
                if (tail[k].g == g && tail[k].f == f && tail[k].v != 0.0) {
                    sm += tail[k].v;
                    q += tail[k].v * tail[k].v;
                    n += 1u;
                }
        }
        if (scale != nullptr) {
            const double sc = (double) scale[fids[f]];
            sm *= sc;
            q *= sc * sc;
            if (sc == 0.0) n = 0u;
        }
        out_sum[t] = sm;
        out_sumsq[t] = q;
        out_cnt[t] = n;
        if (mean != nullptr) {
            const unsigned int rows_g = ge - gs;
            if (rows_g == 0u) {
                mean[t] = 0.0;
                var[t] = 0.0;
            } else {
                const double m = sm / (double) rows_g;
                const double v = q / (double) rows_g - m * m;
                mean[t] = m;
                var[t] = v < 0.0 ? 0.0 : v;
            }
        }
    }
    if (blockIdx.x == 0u)
        for (unsigned int g = threadIdx.x; g < G; g += blockDim.x) group_rows[g] = gstart[g + 1u] - gstart[g];
}

// ---------------------------------------------------------------------------
// Host driver.
// ---------------------------------------------------------------------------
label_private *private_of(csamd_label_reduce_workspace *ws) {
    if (ws->private_workspace == nullptr) {
        ws->private_workspace = std::calloc(1u, sizeof(label_private));
    }
    return (label_private *) ws->private_workspace;
}

struct layout_in {
    int kind;  // 0 CSR, 1 SELL, 2 BELL
    unsigned int rows, cols;
    const unsigned int *idx;
    const void *val;
    unsigned long long limit;
    csr_geom csr;
    sell_geom sell;
    csamd_blocked_ell_view bell;
};

template<int VT>
int launch_rows(const layout_in &L, unsigned int grid, unsigned int threads, size_t lds, hipStream_t s, const reduce_args &a,
                bool vec, const uint2 *ranges) {
    using S = typename value_traits<VT>::storage;
    const S *val = (const S *) L.val;
    if (L.kind != 2) {
        if (vec) {
            auto k = label_reduce_rows_kernel<VT, true>;
            CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            k<<<grid, threads, lds, s>>>(L.idx, val, L.limit, ranges, a);
        } else {
            auto k = label_reduce_rows_kernel<VT, false>;
            CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            k<<<grid, threads, lds, s>>>(L.idx, val, L.limit, ranges, a);
        }
    } else {
        auto k = label_reduce_blocked_kernel<VT>;
        CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        k<<<grid, threads, lds, s>>>(L.bell, a);
    }
    return hip_ok(hipGetLastError(), "label reduce kernel");
}

int max_lds_per_block() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
    return v;
}

int run_label_reduce(const layout_in &L, unsigned int value_type, const csamd_label_reduce_query *q,
                     csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(q != nullptr && ws != nullptr && out != nullptr, "label_reduce: null argument");
    CSAMD_REQUIRE(q->group_count >= 1u && q->group_count <= CSAMD_MAX_LABEL_GROUPS, "label_reduce: group_count %u out of range",
                  q->group_count);
    CSAMD_REQUIRE(value_type_decodable(value_type), "label_reduce: unsupported value type %u", value_type);
    CSAMD_REQUIRE(L.rows == 0u || q->row_labels != nullptr, "label_reduce: row_labels is null");
    CSAMD_REQUIRE(L.kind != 2 || (unsigned long long) L.rows * (L.bell.ell_cols + 8ull) < (1ull << 48),
                  "label_reduce: blocked-ELL has too many slots");
    CSAMD_HIP(hipSetDevice(ws->device >= 0 ? ws->device : 0), "hipSetDevice label reduce");
    label_private *P = private_of(ws);
    CSAMD_REQUIRE(P != nullptr, "label_reduce: host allocation failed");
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int G = q->group_count;
    const unsigned int rows = L.rows, cols = L.cols;
    const unsigned int nwords = (cols + 31u) / 32u;

    if (!P->events_ready) {
        CSAMD_HIP(hipEventCreate(&P->ev0), "hipEventCreate");
        CSAMD_HIP(hipEventCreate(&P->ev1), "hipEventCreate");
        CSAMD_HIP(hipEventCreateWithFlags(&P->ev_meta, hipEventDisableTiming), "hipEventCreate");
        P->events_ready = 1;
    }
    if (P->host == nullptr) {
        void *h = nullptr;
        CSAMD_HIP(hipHostMalloc(&h, 64u, hipHostMallocDefault), "hipHostMalloc label reduce");
        std::memset(h, 0, 64u);
        P->host = (volatile unsigned int *) h;
    }

    // Launch shape.  One 1024-thread workgroup per CU: its waves share the
    // CU's row range through an LDS counter.  (Two workgroups per CU run
    // unevenly: the older one gets issue priority and the younger one
    // finishes late at half occupancy.)  CSAMD_LR_THREADS / CSAMD_LR_PER_CU /
    // CSAMD_LR_SPLIT override the shape for tuning sweeps.
    const size_t fixed_lds = align_up((size_t) (nwords + 1u) * sizeof(uint2), 16u);
    const int max_lds = max_lds_per_block();
    const size_t lds_avail = (size_t) max_lds - 16u;  // static LDS: the row counter
    CSAMD_REQUIRE(fixed_lds + 64u * LR_ACC_BYTES <= lds_avail,
                  "label_reduce: %u columns need %zu bytes of LDS for the feature map", cols, fixed_lds);
    auto env_u = [](const char *name, unsigned int dflt) {
        const char *e = std::getenv(name);
        return e != nullptr ? (unsigned int) std::atoi(e) : dflt;
    };
    const unsigned int cu = (unsigned int) device_cu_count();
    unsigned int threads = env_u("CSAMD_LR_THREADS", 1024u);
    if (threads < 256u || threads > 1024u || threads % 64u != 0u) threads = 1024u;
    unsigned int per_cu = env_u("CSAMD_LR_PER_CU", 1u);
    unsigned int split = env_u("CSAMD_LR_SPLIT", 1u);
    if (split < 1u || split > 16u) split = 1u;

    // Buffers sized by what the host knows (rows, cols, G); the kept counts
    // stay on the device until the feature count is needed.
    const unsigned int tiles = rows == 0u ? 1u : (rows + CSAMD_SCAN_TILE - 1u) / CSAMD_SCAN_TILE;
    unsigned int sort_cap = 2u * cu;  // label table (G + 1) x workgroups stays small enough for one scanning workgroup
    if (sort_cap > 262144u / (G + 1u)) sort_cap = 262144u / (G + 1u);
    if (sort_cap < 8u) sort_cap = 8u;
    const unsigned int sort_grid = grid_for(rows == 0u ? 1u : rows, LR_SORT_TILE, sort_cap);
    const size_t table_n = (size_t) (G + 1u) * sort_grid;
    if (!ensure_buf(P->fpairs, (size_t) (nwords + 1u) * sizeof(uint2), "hipMalloc fpairs")) return 0;
    if (!ensure_buf(P->fids, ((size_t) cols + 16u + nwords + 1u) * sizeof(unsigned int), "hipMalloc fids")) return 0;
    if (!ensure_buf(P->labels, (table_n + (G + 1u) + 1u) * sizeof(unsigned int), "hipMalloc label table")) return 0;
    if (!ensure_buf(P->rows, (size_t) (rows + 1u) * sizeof(unsigned int), "hipMalloc sorted rows")) return 0;
    const size_t work_bytes = align_up((size_t) (rows + 1u) * sizeof(unsigned int), 256u);
    const size_t toff_bytes = align_up((size_t) (tiles + 1u) * sizeof(unsigned long long), 256u);
    if (!ensure_buf(P->work, work_bytes + toff_bytes + (size_t) (rows + 1u) * sizeof(uint2), "hipMalloc work")) return 0;
    if (!ensure_buf(P->tail, LR_MAX_TAIL * sizeof(tail_record), "hipMalloc tail records")) return 0;
    unsigned int *table = (unsigned int *) P->labels.p;  // [(G + 1) x sort_grid]
    unsigned int *gstart = table + table_n;               // [G + 1]
    unsigned int *f_dev = gstart + (G + 1u);              // [1]
    unsigned int *sorted = (unsigned int *) P->rows.p;
    unsigned int *work = (unsigned int *) P->work.p;
    unsigned long long *toff = (unsigned long long *) ((char *) P->work.p + work_bytes);
    uint2 *ranges = (uint2 *) ((char *) P->work.p + work_bytes + toff_bytes);
    unsigned int *fids = (unsigned int *) P->fids.p;
    uint2 *fpairs = (uint2 *) P->fpairs.p;

    // Vector loads need 16-byte aligned indices and unit-aligned values; the
    // vector kernel streams whole 4-entry units only (ranges clipped at
    // clip) and the entries past clip become tail records.
    const bool aligned = L.kind != 2 && ((uintptr_t) L.idx % 16u == 0u) &&
                         ((uintptr_t) L.val % (uintptr_t) (value_type == CSAMD_VALUE_F32 ? 16u
                                                           : value_type == CSAMD_VALUE_FP8_E4M3 ? 4u : 8u) == 0u);
    const unsigned long long clip = aligned ? (L.limit & ~3ull) : ~0ull;
    tail_record *tail = aligned && clip < L.limit ? (tail_record *) P->tail.p : nullptr;

    // 1. label histograms and feature words; 2. label offsets and feature map
    const size_t hist_lds = (size_t) (G + 1u) * sizeof(unsigned int);
    unsigned int *word_mask = (unsigned int *) P->fids.p + (cols + 16u);  // past the kept-feature ids
    label_count_kernel<<<sort_grid, LR_SORT_THREADS, hist_lds, s>>>(q->row_labels, q->row_keep, rows, G, table, work, rows + 1u,
                                                                    q->feature_keep, cols, nwords, word_mask, tail);
    CSAMD_HIP(hipGetLastError(), "label_count_kernel");
    label_offsets_kernel<<<1, 1024, 0, s>>>(table, G, sort_grid, gstart, word_mask, nwords, fpairs, fids, f_dev);
    CSAMD_HIP(hipGetLastError(), "label_offsets_kernel");
    // gstart[G] (kept rows) and F sit next to each other: one copy to pinned
    // memory, waited for at the end of the call (long done by then)
    CSAMD_HIP(hipMemcpyAsync((void *) P->host, gstart + G, 2u * sizeof(unsigned int), hipMemcpyDeviceToHost, s), "D2H counts");
    CSAMD_HIP(hipEventRecord(P->ev_meta, s), "hipEventRecord");
    if (rows != 0u) {
        if (L.kind == 0)
            label_scatter_kernel<csr_geom><<<sort_grid, LR_SORT_THREADS, 2u * hist_lds, s>>>(
                L.csr, q->row_labels, q->row_keep, rows, G, table, clip, sorted, ranges, work, tail, L.idx, L.val, value_type,
                cols, fpairs);
        else if (L.kind == 1)
            label_scatter_kernel<sell_geom><<<sort_grid, LR_SORT_THREADS, 2u * hist_lds, s>>>(
                L.sell, q->row_labels, q->row_keep, rows, G, table, clip, sorted, ranges, work, tail, L.idx, L.val, value_type,
                cols, fpairs);
        else
            label_scatter_kernel<bell_geom><<<sort_grid, LR_SORT_THREADS, 2u * hist_lds, s>>>(
                bell_geom{L.bell.ell_cols}, q->row_labels, q->row_keep, rows, G, table, clip, sorted, ranges, work, nullptr,
                nullptr, nullptr, value_type, cols, fpairs);
        CSAMD_HIP(hipGetLastError(), "label_scatter_kernel");
    }

    // 3. launch shape: the LDS feature tile is sized for all columns (at most
    // F of them are kept; the device reads F), so no host wait is needed here
    const unsigned int fmax = cols == 0u ? 1u : cols;
    if (per_cu != 2u || fixed_lds + 64u * LR_ACC_BYTES > lds_avail / 2u) per_cu = 1u;
    const size_t lds_budget = per_cu == 2u ? lds_avail / 2u : lds_avail;
    unsigned int ftile = (unsigned int) ((lds_budget - fixed_lds) / LR_ACC_BYTES);
    if (ftile >= fmax) ftile = fmax;
    else ftile &= ~63u;
    const unsigned int ntiles = (fmax + ftile - 1u) / ftile;
    const unsigned int W = cu * per_cu * split;
    const size_t lds = fixed_lds + (size_t) ftile * LR_ACC_BYTES;
    if (!ensure_buf(P->bounds, (size_t) (W + 1u) * sizeof(unsigned int), "hipMalloc bounds")) return 0;
    unsigned int *bounds = (unsigned int *) P->bounds.p;

    // 4. partition bounds from the work weights (positions past the kept
    // rows weigh 0)
    if (!scan_tile_offsets_u32(work, rows, toff, s)) return 0;
    work_bounds_kernel<<<tiles, 256, 0, s>>>(work, rows, toff, tiles, W, bounds);
    CSAMD_HIP(hipGetLastError(), "work_bounds_kernel");

    // 5. outputs [G x F] (room for F = cols) and partial slots
    const size_t GF = (size_t) G * fmax;
    const size_t o_sum = 0, o_sumsq = align_up(o_sum + GF * 8u, 256u), o_mean = align_up(o_sumsq + GF * 8u, 256u),
                 o_var = align_up(o_mean + GF * 8u, 256u), o_cnt = align_up(o_var + GF * 8u, 256u),
                 o_grows = align_up(o_cnt + GF * 4u, 256u), o_end = align_up(o_grows + (size_t) G * 4u, 256u);
    if (!ensure_buf(P->out, o_end, "hipMalloc label outputs")) return 0;
    char *ob = (char *) P->out.p;
    double *out_sum = (double *) (ob + o_sum), *out_sumsq = (double *) (ob + o_sumsq);
    double *out_mean = (double *) (ob + o_mean), *out_var = (double *) (ob + o_var);
    unsigned int *out_cnt = (unsigned int *) (ob + o_cnt), *out_grows = (unsigned int *) (ob + o_grows);
    const size_t slots = (size_t) ntiles * ((size_t) W + G) * ftile;
    const size_t p_sum = 0, p_sumsq = align_up(slots * 8u, 256u), p_cnt = align_up(p_sumsq + slots * 8u, 256u),
                 p_end = p_cnt + slots * 4u;
    if (!ensure_buf(P->partial, p_end, "hipMalloc label partials")) return 0;
    char *pb = (char *) P->partial.p;

    // 6. one reduce launch for all feature tiles, then the finish
    reduce_args a;
    a.sorted_rows = sorted;
    a.gstart = gstart;
    a.bounds = bounds;
    a.fpairs = fpairs;
    a.f_total = f_dev;
    a.psum = (double *) (pb + p_sum);
    a.psumsq = (double *) (pb + p_sumsq);
    a.pcnt = (unsigned int *) (pb + p_cnt);
    a.G = G;
    a.cols = cols;
    a.nwords = nwords;
    a.ftile = ftile;
    a.f0 = 0u;
    a.f_count = 0u;
    P->last_valid = 0;
    if (rows != 0u && cols != 0u) {
        const bool timed = P->timing != 0;
        if (timed) CSAMD_HIP(hipEventRecord(P->ev0, s), "hipEventRecord");
        int ok = 0;
        switch (value_type) {
            case CSAMD_VALUE_F16: ok = launch_rows<CSAMD_VALUE_F16>(L, W, threads, lds, s, a, aligned, ranges); break;
            case CSAMD_VALUE_BF16: ok = launch_rows<CSAMD_VALUE_BF16>(L, W, threads, lds, s, a, aligned, ranges); break;
            case CSAMD_VALUE_F32: ok = launch_rows<CSAMD_VALUE_F32>(L, W, threads, lds, s, a, aligned, ranges); break;
            case CSAMD_VALUE_FP8_E4M3: ok = launch_rows<CSAMD_VALUE_FP8_E4M3>(L, W, threads, lds, s, a, aligned, ranges); break;
        }
        if (!ok) return 0;
        if (timed) {
            CSAMD_HIP(hipEventRecord(P->ev1, s), "hipEventRecord");
            P->last_valid = 1;
        }
    }
    const size_t meta_lds = (size_t) (G + 1u + W + 1u) * sizeof(unsigned int);
    label_finish_kernel<<<grid_for((unsigned long long) G * fmax, 256u, 4u * cu), 256, meta_lds, s>>>(
        a.psum, a.psumsq, a.pcnt, gstart, bounds, W, G, ftile, f_dev, q->feature_scale, fids, rows != 0u ? tail : nullptr,
        out_sum, out_sumsq, out_cnt, q->compute_moments ? out_mean : nullptr, out_var, out_grows);
    CSAMD_HIP(hipGetLastError(), "label_finish_kernel");

    // kept rows and features, copied right after the label offsets
    CSAMD_HIP(hipEventSynchronize(P->ev_meta), "label reduce sync (counts)");
    const unsigned int K = P->host[0], F = P->host[1];
    out->group_count = G;
    out->kept_features = F;
    out->kept_rows = K;
    out->reserved = 0u;
    out->sum = out_sum;
    out->sumsq = out_sumsq;
    out->nnz = out_cnt;
    out->group_rows = out_grows;
    out->mean = q->compute_moments ? out_mean : nullptr;
    out->var = q->compute_moments ? out_var : nullptr;
    out->kept_feature_ids = fids;
    return 1;
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_label_reduce_init(csamd_label_reduce_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_label_reduce_clear(csamd_label_reduce_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    label_private *P = (label_private *) ws->private_workspace;
    if (P != nullptr) {
        dev_buf *bufs[] = {&P->fpairs, &P->fids, &P->labels, &P->rows, &P->work, &P->bounds, &P->partial, &P->tail, &P->out};
        for (dev_buf *b : bufs)
            if (b->p != nullptr) (void) hipFree(b->p);
        if (P->host != nullptr) {
            (void) hipStreamSynchronize((hipStream_t) ws->stream);  // kernels may still write it
            (void) hipHostFree((void *) P->host);
        }
        if (P->events_ready) {
            (void) hipEventDestroy(P->ev0);
            (void) hipEventDestroy(P->ev1);
            (void) hipEventDestroy(P->ev_meta);
        }
        std::free(P);
    }
    if (ws->owns_stream != 0 && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    csamd_label_reduce_init(ws);
}

int csamd_label_reduce_setup(csamd_label_reduce_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "label_reduce setup: null workspace");
    csamd_label_reduce_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice label reduce setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags label reduce");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    ws->private_workspace = std::calloc(1u, sizeof(label_private));
    CSAMD_REQUIRE(ws->private_workspace != nullptr, "label_reduce setup: host allocation failed");
    return 1;
}

int csamd_label_reduce_compressed(const csamd_compressed_view *src, const csamd_label_reduce_query *query,
                                  csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(src != nullptr && query != nullptr, "label_reduce_compressed: null argument");
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW, "label_reduce_compressed: axis must be row");
    CSAMD_REQUIRE(src->rows == 0u || src->majorPtr != nullptr, "label_reduce_compressed: null majorPtr");
    layout_in L;
    std::memset(&L, 0, sizeof(L));
    L.kind = 0;
    L.rows = src->rows;
    L.cols = src->cols;
    L.idx = src->minorIdx;
    L.val = src->val;
    L.limit = src->nnz;
    L.csr.ptr = src->majorPtr;
    return run_label_reduce(L, query->value_type, query, ws, out);
}

int csamd_label_reduce_sliced_ell(const csamd_sliced_ell_view *src, const csamd_label_reduce_query *query,
                                  csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(src != nullptr && query != nullptr, "label_reduce_sliced_ell: null argument");
    layout_in L;
    std::memset(&L, 0, sizeof(L));
    L.kind = 1;
    L.rows = src->rows;
    L.cols = src->cols;
    L.idx = src->col_idx;
    L.val = src->val;
    // total slots: last slot offset + rows in last slice * width, computed on the host
    unsigned long long total = 0ull;
    if (src->slice_count != 0u) {
        unsigned int last_off = 0u, last_w = 0u, last_r0 = 0u;
        CSAMD_HIP(hipMemcpy(&last_off, src->slice_slot_offsets + (src->slice_count - 1u), 4u, hipMemcpyDeviceToHost), "D2H sell");
        CSAMD_HIP(hipMemcpy(&last_w, src->slice_widths + (src->slice_count - 1u), 4u, hipMemcpyDeviceToHost), "D2H sell");
        CSAMD_HIP(hipMemcpy(&last_r0, src->slice_row_offsets + (src->slice_count - 1u), 4u, hipMemcpyDeviceToHost), "D2H sell");
        total = (unsigned long long) last_off + (unsigned long long) (src->rows - last_r0) * last_w;
    }
    L.limit = total;
    L.sell.s = *src;
    return run_label_reduce(L, query->value_type, query, ws, out);
}

int csamd_label_reduce_blocked_ell(const csamd_blocked_ell_view *src, const csamd_label_reduce_query *query,
                                   csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(src != nullptr && query != nullptr, "label_reduce_blocked_ell: null argument");
    CSAMD_REQUIRE(src->block_size != 0u || src->ell_cols == 0u, "label_reduce_blocked_ell: block_size is 0");
    layout_in L;
    std::memset(&L, 0, sizeof(L));
    L.kind = 2;
    L.rows = src->rows;
    L.cols = src->cols;
    L.bell = *src;
    return run_label_reduce(L, query->value_type, query, ws, out);
}

int csamd_label_reduce_set_timing(csamd_label_reduce_workspace *ws, int enable) {
    CSAMD_REQUIRE(ws != nullptr, "label_reduce set_timing: null workspace");
    label_private *P = private_of(ws);
    CSAMD_REQUIRE(P != nullptr, "label_reduce set_timing: host allocation failed");
    P->timing = enable != 0;
    if (!P->timing) P->last_valid = 0;
    return 1;
}

float csamd_label_reduce_last_kernel_ms(const csamd_label_reduce_workspace *ws) {
    if (ws == nullptr || ws->private_workspace == nullptr) return 0.0f;
    label_private *P = (label_private *) ws->private_workspace;
    if (!P->last_valid) return 0.0f;
    if (hipEventSynchronize(P->ev1) != hipSuccess) return 0.0f;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, P->ev0, P->ev1) != hipSuccess) return 0.0f;
    return ms;
}

} // extern "C"
