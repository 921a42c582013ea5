// Row-label masked grouped reduction: per (label g, kept feature f)
// {sum, sumsq, nnz} over kept rows, FP64 partials, FP64 combine, mean/var.
// One call covers one matrix or a batch of partitions (CSR / Sliced-ELL
// partitions with the same columns; row arrays span the partitions' rows
// in order), with one reduce launch over all of them.
//
// There is no reference kernel for this operation (mask_groups.cu reduces per
// row over <= 32 feature-group bitmasks); its masking rules are the
// reference's: a row takes part iff row_keep[r] != 0 (mask_groups.cu:65-67),
// a stored entry iff col < cols, v != 0 and feature_keep[col] != 0
// (mask_groups.cu:155).
//
// Design (gfx950):
//   1. label histograms per workgroup (LDS) + kept-feature words (ballots);
//      one workgroup scans the label table into slot offsets and builds the
//      feature map {mask, prefix} per 32-column word (feature -> rank);
//   2. kept rows scattered by label (counting sort) with, per placed row, a
//      32-byte range record {index array, value array, entry range} of its
//      partition and a work weight; entries past a partition's last whole
//      4-entry unit become tail records keyed by the row's sorted position;
//   3. nnz-balanced contiguous ranges of the sorted rows, one per workgroup
//      (one 1024-thread workgroup per CU), from a tile scan of the weights;
//   4. each workgroup walks its range label segment by label segment; its
//      waves take rows from an LDS counter and stream them with 16-byte index
//      / 8-byte value non-temporal vector loads through two register buffers;
//      kept entries land in LDS FP64 accumulators (ds_add_f64) and u32
//      counters; tail records of the segment are added; the LDS partial is
//      stored to slot (feature tile, workgroup + label);
//   5. a finish kernel sums, per (g, f), the slots of the workgroups that
//      hold label g, applies the feature scale, and derives mean/variance.
#include "common.hh"
#include "pipeline.hh"
#include "stream.hh"

#include <cstdlib>
#include <vector>

namespace csamd {
namespace {

constexpr unsigned int LR_ACC_BYTES = 20u;   // f64 sum + f64 sumsq + u32 nnz per kept feature
constexpr unsigned int LR_MAX_PARTS = 4096u;
constexpr unsigned int LR_MAX_TAIL = 3u;     // entries past the last whole 4-entry unit, per partition

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
    dev_buf fids;        // kept feature ids, then per-word masks
    dev_buf labels;      // label table [(G + 1) x sort workgroups], gstart [G + 1], F
    dev_buf rows;        // sorted kept rows [rows]
    dev_buf work;        // work weights [rows], tile offsets, row range records [rows]
    dev_buf bounds;      // [W + 1]
    dev_buf partial;     // tiles x (W + G) x ftile x {f64, f64, u32}
    dev_buf tail;        // tail records [3 x partitions]
    dev_buf parts;       // partition descriptors
    dev_buf out;         // outputs
    void *parts_host;    // pinned staging of the descriptors
    size_t parts_host_bytes;
    volatile unsigned int *host;  // pinned: [0] kept rows, [1] kept features
    hipEvent_t ev0, ev1, ev_meta;
    int events_ready;
    int timing;
    int last_valid;
};

// ---------------------------------------------------------------------------
// Partitions.
// ---------------------------------------------------------------------------
enum : unsigned int { PART_CSR = 0u, PART_SELL = 1u, PART_BELL = 2u };

struct part_desc {
    const unsigned int *idx;              // CSR minorIdx / SELL col_idx
    const void *val;
    const unsigned int *ptr;              // CSR majorPtr
    const unsigned int *slice_row_offsets;  // SELL
    const unsigned int *slice_widths;
    const unsigned int *slice_slot_offsets;
    unsigned long long limit;             // stored entries (CSR nnz / SELL slots); SELL: set on the device
    unsigned long long clip;              // vector loads stream entries below clip
    unsigned int kind;
    unsigned int rows;
    unsigned int row0;                    // first row of the partition in the row arrays
    unsigned int slice_count;
    unsigned int slice_rows;
    unsigned int ell_cols;                // BELL (work weights only)
    unsigned int last;                    // limit - 1 (scalar-load clamp)
    unsigned int vec;                     // 1: vector loads (clip = whole units), 0: clip = limit
};

// Entry range [b, e) of local row r, relative to the partition's arrays.
__device__ __forceinline__ void part_range(const part_desc &P, unsigned int r, unsigned long long *b, unsigned long long *e) {
    if (P.kind == PART_CSR) {
        *b = P.ptr[r];
        *e = P.ptr[r + 1u];
    } else if (P.kind == PART_SELL) {
        if (P.slice_count == 0u) {
            *b = *e = 0ull;
            return;
        }
        unsigned int slice;
        if (P.slice_rows == 32u) slice = r >> 5;
        else if (P.slice_rows != 0u) slice = r / P.slice_rows;
        else {
            unsigned int lo = 0u, hi = P.slice_count - 1u;
            while (lo < hi) {
                const unsigned int mid = (lo + hi + 1u) >> 1;
                if (P.slice_row_offsets[mid] <= r) lo = mid;
                else hi = mid - 1u;
            }
            slice = lo;
        }
        if (slice >= P.slice_count) slice = P.slice_count - 1u;
        const unsigned int w = P.slice_widths[slice];
        const unsigned long long base =
            (unsigned long long) P.slice_slot_offsets[slice] + (unsigned long long) (r - P.slice_row_offsets[slice]) * w;
        *b = base;
        *e = base + w;
    } else {
        *b = 0ull;
        *e = P.ell_cols;
    }
}

// Up to LR_ARG_PARTS descriptors travel as kernel arguments (no copy).
constexpr unsigned int LR_ARG_PARTS = 24u;
struct part_args {
    part_desc d[LR_ARG_PARTS];
};

// Descriptor table on the device (from the kernel arguments when inline),
// with SELL slot totals, clips and clamps computed here (no host round
// trip).  Idempotent: a replayed graph recomputes the same table.  Run by
// the label count kernel's threads (no launch of its own).
__device__ __forceinline__ void prepare_parts(part_desc *__restrict__ parts, unsigned int nparts, const part_args &inl,
                                              unsigned int use_inline) {
    for (unsigned int p = blockIdx.x * blockDim.x + threadIdx.x; p < nparts; p += gridDim.x * blockDim.x) {
        part_desc P = use_inline ? inl.d[p] : parts[p];
        if (P.kind == PART_SELL) {
            unsigned long long total = 0ull;
            if (P.slice_count != 0u) {
                const unsigned int s = P.slice_count - 1u;
                total = (unsigned long long) P.slice_slot_offsets[s] +
                        (unsigned long long) (P.rows - P.slice_row_offsets[s]) * P.slice_widths[s];
            }
            P.limit = total;
        }
        P.clip = P.vec ? (P.limit & ~3ull) : ~0ull;
        P.last = P.limit != 0ull ? (unsigned int) (P.limit - 1ull) : 0u;
        parts[p] = P;
    }
}

// Per placed row: where its entries are.  32 bytes, read with one scalar
// load per row by the reduce kernel.
struct row_range {
    const unsigned int *idx;
    const void *val;
    unsigned int b;     // first entry (partition-relative)
    unsigned int e;     // one past the last entry streamed (clipped)
    unsigned int last;  // partition's last entry index (scalar-load clamp)
    unsigned int part;
};

// An entry past its partition's last whole 4-entry unit, for the reduce
// kernel to add when it holds sorted position pos (0xffffffff: none).
struct tail_record {
    unsigned int pos;
    unsigned int col;
    float v;
    unsigned int pad;
};

__device__ __forceinline__ float decode_any(const void *val, unsigned long long i, unsigned int vt) {
    switch (vt) {
        case CSAMD_VALUE_F16: return value_traits<CSAMD_VALUE_F16>::decode(((const uint16_t *) val)[i]);
        case CSAMD_VALUE_BF16: return value_traits<CSAMD_VALUE_BF16>::decode(((const uint16_t *) val)[i]);
        case CSAMD_VALUE_F32: return ((const float *) val)[i];
        default: return value_traits<CSAMD_VALUE_FP8_E4M3>::decode(((const uint8_t *) val)[i]);
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
constexpr unsigned int LR_SORT_ITEMS = 2u;
constexpr unsigned int LR_SORT_TILE = LR_SORT_THREADS * LR_SORT_ITEMS;

__device__ __forceinline__ unsigned int row_key(const unsigned int *__restrict__ labels, const unsigned char *__restrict__ row_keep,
                                                unsigned int r, unsigned int G) {
    const unsigned int l = labels[r];
    return ((row_keep == nullptr || row_keep[r] != 0u) && l < G) ? l : G;
}

// Also: prepares the partition descriptors; clears work[0..work_n)
// (positions past the kept rows keep weight 0 for the partition scan);
// builds the kept-feature mask of each 32-column word; clears the tail
// records.
__global__ void __launch_bounds__(LR_SORT_THREADS) label_count_kernel(const unsigned int *__restrict__ labels,
                                                                      const unsigned char *__restrict__ row_keep, unsigned int rows,
                                                                      unsigned int G, unsigned int *__restrict__ hist_out,
                                                                      unsigned int *__restrict__ work, unsigned int work_n,
                                                                      const unsigned char *__restrict__ feature_keep, unsigned int cols,
                                                                      unsigned int nwords, unsigned int *__restrict__ word_mask,
                                                                      tail_record *__restrict__ tail, unsigned int ntail,
                                                                      part_desc *__restrict__ parts, unsigned int nparts,
                                                                      part_args inl, unsigned int use_inline) {
    extern __shared__ unsigned int hist[];  // [G + 1]
    prepare_parts(parts, nparts, inl, use_inline);
    for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) hist[g] = 0u;
    for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntail; t += gridDim.x * blockDim.x)
        tail[t] = tail_record{0xffffffffu, CSAMD_INVALID_COL, 0.0f, 0u};
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
    __shared__ unsigned int stage[1024u * LR_OFF_ITEMS];
    // the table's first pass is loaded before the feature map is built, and
    // each later pass while the previous one is scanned
    const unsigned long long n = (unsigned long long) (G + 1u) * nb;
    unsigned int nxt[LR_OFF_ITEMS];
#pragma unroll
    for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
        const unsigned long long i = k * 1024ull + threadIdx.x;
        nxt[k] = i < n ? hist[i] : 0u;
    }
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
    // the table in passes of 1024 x LR_OFF_ITEMS entries, loaded and stored
    // coalesced through LDS; each thread scans LR_OFF_ITEMS consecutive ones
    unsigned int carry = 0u;  // same value in every thread
    for (unsigned long long start = 0; start < n; start += 1024ull * LR_OFF_ITEMS) {
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) stage[k * 1024u + threadIdx.x] = nxt[k];
        const unsigned long long next = start + 1024ull * LR_OFF_ITEMS;
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
            const unsigned long long i = next + k * 1024u + threadIdx.x;
            nxt[k] = i < n ? hist[i] : 0u;
        }
        __syncthreads();
        unsigned int v[LR_OFF_ITEMS], own = 0u;
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
            v[k] = stage[threadIdx.x * LR_OFF_ITEMS + k];
            own += v[k];
        }
        unsigned int total = 0u;
        unsigned int o = carry + block_exclusive_scan_u32(own, smem, &total);
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
            stage[threadIdx.x * LR_OFF_ITEMS + k] = o;
            o += v[k];
        }
        __syncthreads();
#pragma unroll
        for (unsigned int k = 0; k < LR_OFF_ITEMS; ++k) {
            const unsigned long long i = start + k * 1024u + threadIdx.x;
            if (i < n) hist[i] = stage[k * 1024u + threadIdx.x];
        }
        __syncthreads();
        carry += total;
    }
    __syncthreads();  // the block's table stores are visible to the block
    for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) gstart[g] = hist[(unsigned long long) g * nb];
}

// Scatter kept rows to their label's slots: workgroup b's rows of label g
// start at the scanned table entry (g, b); within a tile, ranks come from
// LDS atomics, so the order of rows within a label is not fixed (sums are
// order independent up to FP64 rounding; counts are exact).  For each
// placed row it writes the row's range record (entries clipped at its
// partition's clip: the vector-load kernel streams only whole 4-entry
// units), the row's tail records for entries past the clip, and its work
// weight for the partition (entries + 8 for the per-row cost).
__global__ void __launch_bounds__(LR_SORT_THREADS) label_scatter_kernel(const part_desc *__restrict__ parts, unsigned int nparts,
                                                                        const unsigned int *__restrict__ labels,
                                                                        const unsigned char *__restrict__ row_keep, unsigned int rows,
                                                                        unsigned int G, const unsigned int *__restrict__ base_table,
                                                                        unsigned int *__restrict__ sorted_rows, row_range *__restrict__ ranges,
                                                                        unsigned int *__restrict__ work, tail_record *__restrict__ tail,
                                                                        unsigned int vt) {
    extern __shared__ unsigned int hist[];  // [G + 1] tile counts, [G + 1] running bases, [nparts + 1] first rows
    unsigned int *base = hist + (G + 1u);
    unsigned int *row0 = base + (G + 1u);
    for (unsigned int g = threadIdx.x; g < G; g += blockDim.x) base[g] = base_table[(unsigned long long) g * gridDim.x + blockIdx.x];
    for (unsigned int p = threadIdx.x; p <= nparts; p += blockDim.x) row0[p] = p < nparts ? parts[p].row0 : rows;
    for (unsigned int t0 = blockIdx.x * LR_SORT_TILE; t0 < rows; t0 += gridDim.x * LR_SORT_TILE) {
        for (unsigned int g = threadIdx.x; g <= G; g += blockDim.x) hist[g] = 0u;
        __syncthreads();
        unsigned int key[LR_SORT_ITEMS], rank[LR_SORT_ITEMS], part[LR_SORT_ITEMS];
        unsigned long long rb[LR_SORT_ITEMS], re[LR_SORT_ITEMS];
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) {
            const unsigned int r = t0 + k * LR_SORT_THREADS + threadIdx.x;
            key[k] = r < rows ? row_key(labels, row_keep, r, G) : G;
            unsigned int lo = 0u, hi = nparts - 1u;  // last partition with row0 <= r
            while (lo < hi) {
                const unsigned int mid = (lo + hi + 1u) >> 1;
                if (row0[mid] <= r) lo = mid;
                else hi = mid - 1u;
            }
            part[k] = lo;
            rb[k] = re[k] = 0ull;
            if (key[k] < G) part_range(parts[lo], r - row0[lo], &rb[k], &re[k]);
        }
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) rank[k] = key[k] < G ? atomicAdd(&hist[key[k]], 1u) : 0u;
        __syncthreads();
#pragma unroll
        for (unsigned int k = 0; k < LR_SORT_ITEMS; ++k) {
            if (key[k] >= G) continue;
            const unsigned int r = t0 + k * LR_SORT_THREADS + threadIdx.x;
            const unsigned int pos = base[key[k]] + rank[k];
            const part_desc &P = parts[part[k]];
            const unsigned long long b = rb[k], e = re[k], clip = P.clip;
            const unsigned long long len = e - b;
            sorted_rows[pos] = r;
            row_range rr;
            rr.idx = P.idx;
            rr.val = P.val;
            rr.b = (unsigned int) (b < clip ? b : clip);
            rr.e = (unsigned int) (e < clip ? e : clip);
            rr.last = P.last;
            rr.part = part[k];
            ranges[pos] = rr;
            work[pos] = (unsigned int) (len > 0xfffffff0ull ? 0xfffffff0ull : len) + 8u;
            if (tail != nullptr && P.kind != PART_BELL && e > clip) {  // entries past the last whole unit
                for (unsigned long long i = b > clip ? b : clip; i < e; ++i)
                    tail[part[k] * LR_MAX_TAIL + (unsigned int) (i - clip)] =
                        tail_record{pos, P.idx[i], decode_any(P.val, i, vt), 0u};
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
    const uint2 *lpairs;  // [nwords + 1] the feature tile's map, {0, 0} sentinel last
    uint2 *queues;        // [waves x LR_QCAP] kept entries {tile feature, value bits}
};

struct reduce_args {
    const unsigned int *sorted_rows;
    const unsigned int *gstart;   // [G+1]
    const unsigned int *bounds;   // [W+1]
    const uint2 *fpairs;          // [nwords]
    const unsigned int *f_total;  // kept features F (device)
    const tail_record *tail;      // [ntail]
    double *psum;                 // [tiles x (W+G) x ftile]
    double *psumsq;
    unsigned int *pcnt;
    unsigned int ntail;
    unsigned int G;
    unsigned int cols;
    unsigned int nwords;
    unsigned int ftile;           // features held in LDS per pass
    unsigned int f0;              // first feature of this pass (set per pass)
    unsigned int f_count;         // features in this pass (set per pass)
};

__device__ __forceinline__ void acc_add(const acc_lds &acc, unsigned int f, float v) {
    const double d = (double) v;
    atomicAdd(acc.sum + f, d);
    atomicAdd(acc.sumsq + f, d * d);
    atomicAdd(acc.cnt + f, 1u);
}

// Feature map of one LDS feature tile [f0, f1): per 32-column word, the
// kept columns whose feature rank lies in the tile, and the tile-relative
// rank of the first of them.  A column c is then in the tile iff bit c & 31
// of x is set, and its tile feature is y + popcount(x below bit c & 31), so
// the streaming loop needs neither the tile's first feature nor its size.
__device__ __forceinline__ uint2 tile_pairs(uint2 pm, unsigned int f0, unsigned int f1) {
    const unsigned int n = (unsigned int) __popc(pm.x);
    const unsigned int lo = pm.y > f0 ? pm.y : f0;
    const unsigned int hi = pm.y + n < f1 ? pm.y + n : f1;
    if (lo >= hi) return make_uint2(0u, 0u);
    unsigned int x = pm.x;
    for (unsigned int k = pm.y; k < lo; ++k) x &= x - 1u;  // drop the kept columns ranked before the tile
    unsigned int keep = 0u;
    for (unsigned int k = lo; k < hi; ++k) {                // keep the next hi - lo of them
        const unsigned int low = x & (0u - x);
        keep |= low;
        x ^= low;
    }
    return make_uint2(keep, lo - f0);
}

// Column c against its word's tile map pm: true iff c is a kept feature of
// the tile, with f its tile feature.  The bit-field extracts read only the
// low 5 bits of c.
__device__ __forceinline__ bool tile_feature(uint2 pm, unsigned int c, unsigned int &f) {
    f = pm.y + (unsigned int) __popc(__builtin_amdgcn_ubfe(pm.x, 0u, c));
    return __builtin_amdgcn_ubfe(pm.x, c, 1u) != 0u;
}

// The workgroup's sorted-row range [bounds[b], bounds[b + 1]), walked one
// feature tile (LDS pass) at a time and, within it, one label segment at a
// time: zero the LDS accumulators, run body(segment start, end, tile args),
// add the segment's tail records, store the accumulators to partial slot
// (tile, workgroup + label).  F is read on the device, so the host launches
// one kernel for all tiles.
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
        // the tile's feature map (the previous tile's last segment ended
        // with a barrier; the first segment's barrier publishes it)
        uint2 *lp = (uint2 *) acc.lpairs;
        for (unsigned int w = threadIdx.x; w <= a.nwords; w += blockDim.x)
            lp[w] = w < a.nwords ? tile_pairs(a.fpairs[w], t.f0, t.f0 + t.f_count) : make_uint2(0u, 0u);
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
            for (unsigned int k = threadIdx.x; k < a.ntail; k += blockDim.x) {
                const tail_record rec = a.tail[k];
                if (rec.pos < s0 || rec.pos >= s1 || rec.col >= a.cols || rec.v == 0.0f) continue;
                unsigned int f;
                if (tile_feature(acc.lpairs[rec.col >> 5], rec.col, f)) acc_add(acc, f, rec.v);
            }
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
__device__ __forceinline__ void consume(unsigned int c, typename value_traits<VT>::storage raw, const reduce_args &a,
                                        const acc_lds &acc) {
    if (c >= a.cols) return;
    unsigned int f;
    if (!tile_feature(acc.lpairs[c >> 5], c, f)) return;
    const float v = value_traits<VT>::decode(raw);
    if (v != 0.0f) acc_add(acc, f, v);
}

// A wave's kept entries go through a ring of LR_QCAP slots in LDS: each
// entry slot of a chunk appends the lanes whose entry is kept (ballot +
// mbcnt), and whenever 64 are waiting, every lane takes one and adds it to
// the accumulators.  With ~8 % of entries kept, the accumulator atomics then
// run on full waves instead of once per entry slot on a few lanes.  The ring
// holds < 64 waiting entries plus one slot's 64 new ones.  A wave's LDS
// operations complete in order, so a lane reads what another lane of the
// wave appended before it, with no barrier.
constexpr unsigned int LR_QCAP = 128u;

struct kept_queue {
    uint2 *q;        // this wave's ring
    unsigned int n;  // entries appended (wave-uniform)
    unsigned int h;  // entries added to the accumulators (wave-uniform)
};

__device__ __forceinline__ void queue_push(kept_queue &Q, bool in_tile, unsigned int f, float v, const acc_lds &acc) {
    // two ballots straight from the compares, combined on the scalar unit
    const unsigned long long bal = __builtin_amdgcn_ballot_w64(in_tile) & __builtin_amdgcn_ballot_w64(v != 0.0f);
    const bool kept = in_tile && v != 0.0f;
    const unsigned int pos = __builtin_amdgcn_mbcnt_hi((unsigned int) (bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int) bal, Q.n));
    if (kept) Q.q[pos & (LR_QCAP - 1u)] = make_uint2(f, __float_as_uint(v));
    Q.n += (unsigned int) __popcll(bal);
    if (Q.n - Q.h >= 64u) {
        const uint2 e = Q.q[(Q.h + lane_id()) & (LR_QCAP - 1u)];
        acc_add(acc, e.x, __uint_as_float(e.y));
        Q.h += 64u;
    }
}

// The waiting entries (fewer than 64), before the segment's accumulators are read.
__device__ __forceinline__ void queue_drain(kept_queue &Q, const acc_lds &acc) {
    const unsigned int lane = lane_id();
    if (lane < Q.n - Q.h) {
        const uint2 e = Q.q[(Q.h + lane) & (LR_QCAP - 1u)];
        acc_add(acc, e.x, __uint_as_float(e.y));
    }
    Q.h = Q.n;
}

constexpr unsigned int LR_UNITS = 4u;               // 4-entry units per lane per chunk (even)
constexpr unsigned int LR_CHUNK_UNITS = 64u * LR_UNITS;

template<int VT> struct chunk_regs {
    u32x4 c[LR_UNITS];
    typename unit_vals<VT>::raw v[LR_UNITS];
};

// Issue the loads of one chunk: unit u + lane + 64 j of the current row.
// Lanes past the row's last unit u_end - 1 re-read that unit (same cache
// lines, masked when consumed), so every lane loads and no lane branches.
// VEC: row ranges end at or before the last whole 4-entry unit of their
// partition's arrays, so every unit is one 16-byte index load and one value
// load, and the loop carries a fixed number of loads per chunk (which lets
// the compiler wait for the current chunk only).  Otherwise (unaligned
// arrays) single entries are read, clamped to index last.
template<int VT, bool VEC>
__device__ __forceinline__ void load_chunk(const unsigned int *__restrict__ idx, const typename value_traits<VT>::storage *__restrict__ val,
                                           unsigned int u, unsigned int u_end, unsigned int last, chunk_regs<VT> &r) {
    const unsigned int lane = lane_id();
    if (VEC) {
#pragma unroll
        for (unsigned int j = 0; j < LR_UNITS; ++j) {
            const unsigned int uu = min(u + lane + 64u * j, u_end - 1u);
            r.c[j] = __builtin_nontemporal_load(gmem((const u32x4 *) (idx + uu * 4u)));
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


// Consume one chunk of the row whose entries are [b, e), starting at unit u,
// two units (8 entries per lane) at a time.  Per pair, the 8 bitmap lookups
// are issued first (independent LDS reads); keep tests, ranks and values are
// computed without branches; kept entries go to the wave's queue.  Nothing here reads global memory, so the prefetched
// chunk stays in flight (a feature scale is applied by the finish).
// lpairs[nwords] is a {0, 0} sentinel, so columns >= cols (and
// CSAMD_INVALID_COL) are "not kept".  Entries outside [b, e) exist only in a
// row's first and last chunk; only those chunks (a wave-uniform test) pay
// for the range mask.  Every loaded unit is consumed: a load whose
// registers are never read is waited for wherever the registers are next
// written.
template<int VT>
__device__ __forceinline__ void consume_chunk(const chunk_regs<VT> &r, unsigned int u, unsigned int b, unsigned int e,
                                              const reduce_args &a, const acc_lds &acc, kept_queue &Q) {
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
        for (unsigned int k = 0; k < 8u; ++k) pm[k] = acc.lpairs[min(c[k] >> 5, a.nwords)];
#pragma unroll
        for (unsigned int k = 0; k < 8u; ++k) {
            unsigned int f;
            const bool in_tile = tile_feature(pm[k], c[k], f);
            const float v = unit_vals<VT>::get(r.v[2u * p + (k >> 2)], k & 3u);
            queue_push(Q, in_tile, f, v, acc);
        }
    }
}

__device__ __forceinline__ void row_units(unsigned int b, unsigned int e, unsigned int *u, unsigned int *ue) {
    *u = b >> 2;
    *ue = b == e ? *u : (unsigned int) (((unsigned long long) e + 3ull) >> 2);
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
    row_range rng;    // the row's arrays and entry range [rng.b, rng.e)
    unsigned int u;   // first unit of the chunk
    unsigned int ue;  // end unit of the row
};

// Rows of one label segment [.., s1), taken one at a time from the shared
// row counter.  The chunk stream runs through two fixed register buffers in
// turn (A, B, A, ...): the next chunk, possibly of the next row, is loaded
// before the current one is consumed, with no register copies between
// them, and the next row is taken and its range record fetched one row
// ahead.
struct row_stream {
    const row_range *ranges;
    unsigned int *next;
    unsigned int s1;
    unsigned int inext;  // row taken ahead (>= s1: none left)
    row_range nxt;       // its range record

    __device__ __forceinline__ void take() {
        inext = grab_row(next);
        if (inext < s1) nxt = ranges[inext];
    }
    // Move to the first non-empty row taken ahead; false when none is left.
    __device__ __forceinline__ bool next_row(row_cursor &c) {
        do {
            if (inext >= s1) return false;
            c.rng = nxt;
            take();
            row_units(c.rng.b, c.rng.e, &c.u, &c.ue);
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
__device__ __forceinline__ void load_cursor(const row_cursor &c, chunk_regs<VT> &r) {
    using S = typename value_traits<VT>::storage;
    load_chunk<VT, VEC>(c.rng.idx, (const S *) c.rng.val, c.u, c.ue, c.rng.last, r);
}

template<int VT, bool VEC>
__device__ __forceinline__ void stream_rows(const row_range *__restrict__ ranges, unsigned int *next, unsigned int s1,
                                            const reduce_args &a, const acc_lds &acc) {
    row_stream st{ranges, next, s1, 0u, row_range{}};
    st.take();
    row_cursor ca, cb;
    if (!st.next_row(ca)) return;
    kept_queue Q{acc.queues + (threadIdx.x >> 6) * LR_QCAP, 0u, 0u};
    // Loads are issued unconditionally (after the last chunk the wave
    // re-loads it, in bounds and unused): a load under a branch makes the
    // compiler merge old and new buffer contents, which waits for the load.
    chunk_regs<VT> A, B;
    load_cursor<VT, VEC>(ca, A);
    while (true) {
        cb = ca;
        bool more = st.advance(cb);
        if (!more) cb = ca;
        load_cursor<VT, VEC>(cb, B);
        consume_chunk<VT>(A, ca.u, ca.rng.b, ca.rng.e, a, acc, Q);
        if (!more) break;
        ca = cb;
        more = st.advance(ca);
        if (!more) ca = cb;
        load_cursor<VT, VEC>(ca, A);
        consume_chunk<VT>(B, cb.u, cb.rng.b, cb.rng.e, a, acc, Q);
        if (!more) break;
    }
    queue_drain(Q, acc);
}

__device__ __forceinline__ acc_lds lds_layout(unsigned char *lds, const reduce_args &a) {
    uint2 *lpairs = (uint2 *) lds;
    double *lsum = (double *) (lds + align_up((size_t) (a.nwords + 1u) * sizeof(uint2), 16u));
    double *lsumsq = lsum + a.ftile;
    unsigned int *lcnt = (unsigned int *) (lsumsq + a.ftile);
    uint2 *queues = (uint2 *) (lcnt + ((a.ftile + 1u) & ~1u));  // 8-byte aligned
    return acc_lds{lsum, lsumsq, lcnt, lpairs, queues};
}

template<int VT, bool VEC>
__global__ void label_reduce_rows_kernel(const row_range *__restrict__ ranges, reduce_args a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int next_row;
    const acc_lds acc = lds_layout(lds, a);
    walk_segments(a, acc, [&](unsigned int s0, unsigned int s1, const reduce_args &t) {
        if (threadIdx.x == 0u) next_row = s0;
        __syncthreads();  // also publishes the zeroed accumulators
        stream_rows<VT, VEC>(ranges, &next_row, s1, t, acc);
    });
}

// Blocked-ELL rows: entries decoded through blockColIdx (scalar lane stride).
template<int VT>
__global__ void label_reduce_blocked_kernel(csamd_blocked_ell_view src, reduce_args a) {
    using S = typename value_traits<VT>::storage;
    extern __shared__ __align__(16) unsigned char lds[];
    const unsigned int nwaves = blockDim.x >> 6;
    const unsigned int wave = threadIdx.x >> 6;
    const acc_lds acc = lds_layout(lds, a);
    const S *__restrict__ val = (const S *) src.val;
    const unsigned int bs = src.block_size;
    const unsigned int width = bs != 0u ? src.ell_cols / bs : 0u;

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
                consume<VT>(c, val[base + e], t, acc);
            }
        }
    });
}

// Finish: out[g][f] = sum, over the workgroups w whose range holds rows of
// label g (ascending w), of partial slot (tile of f, w + g); then the
// feature's scale s (if any): s * sum, s^2 * sumsq, no nonzeros when s == 0;
// then mean = sum / n_g and var = sumsq / n_g - mean^2 (clamped at 0),
// n_g = kept rows of label g.  F is read on the device; gstart and bounds
// are staged in LDS for the per-thread search.
__global__ void __launch_bounds__(256) label_finish_kernel(
    const double *__restrict__ psum, const double *__restrict__ psumsq, const unsigned int *__restrict__ pcnt,
    const unsigned int *__restrict__ gstart_g, const unsigned int *__restrict__ bounds_g, unsigned int W, unsigned int G,
    unsigned int ftile, const unsigned int *__restrict__ f_total, const float *__restrict__ scale,
    const unsigned int *__restrict__ fids, double *__restrict__ out_sum, double *__restrict__ out_sumsq,
    unsigned int *__restrict__ out_cnt, double *__restrict__ mean, double *__restrict__ var, unsigned int *__restrict__ group_rows) {
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
            // FU workgroups' partials loaded per step (a group spans about W / G
            // workgroups, so most groups take one or two steps), then summed in w order
            constexpr unsigned int FU = 16u;
            for (unsigned int w0 = l; w0 < r; w0 += FU) {
                double ps[FU], pq[FU];
                unsigned int pc[FU];
#pragma unroll
                for (unsigned int k = 0; k < FU; ++k) {  // loads first, then sums in w order
                    const unsigned int w = w0 + k < r ? w0 + k : l;  // clamped: never reads past the slots
                    const bool used = w0 + k < r && bounds[w] < bounds[w + 1u];
                    const unsigned long long slot = fbase + (unsigned long long) (w + g) * ftile;
                    ps[k] = used ? psum[slot] : 0.0;
                    pq[k] = used ? psumsq[slot] : 0.0;
                    pc[k] = used ? pcnt[slot] : 0u;
                }
#pragma unroll
                for (unsigned int k = 0; k < FU; ++k) {
                    sm += ps[k];
                    q += pq[k];
                    n += pc[k];
                }
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

struct part_in {
    unsigned int kind;  // PART_*
    unsigned int rows;
    unsigned int cols;
    const unsigned int *idx;
    const void *val;
    unsigned long long nnz;  // CSR stored entries
    const unsigned int *ptr;
    csamd_sliced_ell_view sell;
    csamd_blocked_ell_view bell;
};

template<int VT>
int launch_rows(bool bell, const csamd_blocked_ell_view &bv, unsigned int grid, unsigned int threads, size_t lds, hipStream_t s,
                const reduce_args &a, bool vec, const row_range *ranges, bool set_attr) {
    if (!bell) {
        if (vec) {
            auto k = label_reduce_rows_kernel<VT, true>;
            if (set_attr) CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            else k<<<grid, threads, lds, s>>>(ranges, a);
        } else {
            auto k = label_reduce_rows_kernel<VT, false>;
            if (set_attr) CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            else k<<<grid, threads, lds, s>>>(ranges, a);
        }
    } else {
        auto k = label_reduce_blocked_kernel<VT>;
        if (set_attr) CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        else k<<<grid, threads, lds, s>>>(bv, a);
    }
    return hip_ok(hipGetLastError(), "label reduce kernel");
}

int max_lds_per_block() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
    return v;
}

// Everything one call launches, decided and allocated before any launch
// (so that the launches can be captured into a graph).
struct lr_plan {
    unsigned int rows, cols, G, nwords, nparts, sort_grid, tiles, ntail, W, threads, fmax, vt, cu;
    bool bell, aligned, inline_parts, moments;
    size_t lds, hist_lds, scatter_lds, meta_lds, parts_bytes;
    const unsigned int *labels;
    const unsigned char *row_keep, *feature_keep;
    const float *scale;
    unsigned int *table, *gstart, *f_dev, *sorted, *work, *word_mask, *fids, *bounds;
    unsigned long long *toff;
    row_range *ranges;
    uint2 *fpairs;
    tail_record *tail;
    part_desc *d_parts;
    double *out_sum, *out_sumsq, *out_mean, *out_var;
    unsigned int *out_cnt, *out_grows;
    reduce_args a;
    csamd_blocked_ell_view bv;
    part_args inl;
};

int plan_label_reduce(const std::vector<part_in> &parts, unsigned int value_type, const csamd_label_reduce_query *q,
                      label_private *P, lr_plan &L) {
    CSAMD_REQUIRE(!parts.empty() && parts.size() <= LR_MAX_PARTS, "label_reduce: 1..%u partitions", LR_MAX_PARTS);
    CSAMD_REQUIRE(q->group_count >= 1u && q->group_count <= CSAMD_MAX_LABEL_GROUPS, "label_reduce: group_count %u out of range",
                  q->group_count);
    CSAMD_REQUIRE(value_type_decodable(value_type), "label_reduce: unsupported value type %u", value_type);
    std::memset(&L, 0, sizeof(L));
    const unsigned int nparts = (unsigned int) parts.size();
    const bool bell = parts[0].kind == PART_BELL;
    CSAMD_REQUIRE(!bell || nparts == 1u, "label_reduce: a blocked-ELL input is a single partition");
    const unsigned int cols = parts[0].cols;
    unsigned long long rows64 = 0ull;
    bool aligned = !bell;
    const uintptr_t val_align = value_type == CSAMD_VALUE_F32 ? 16u : value_type == CSAMD_VALUE_FP8_E4M3 ? 4u : 8u;
    for (const part_in &pi : parts) {
        CSAMD_REQUIRE(pi.cols == cols, "label_reduce: partitions must have the same column count");
        rows64 += pi.rows;
        aligned = aligned && (uintptr_t) pi.idx % 16u == 0u && (uintptr_t) pi.val % val_align == 0u;
    }
    CSAMD_REQUIRE(rows64 < 0xffffffffull, "label_reduce: more than 2^32 - 1 rows");
    CSAMD_REQUIRE(!bell || (unsigned long long) parts[0].rows * (parts[0].bell.ell_cols + 8ull) < (1ull << 48),
                  "label_reduce: blocked-ELL has too many slots");
    const unsigned int rows = (unsigned int) rows64;
    CSAMD_REQUIRE(rows == 0u || q->row_labels != nullptr, "label_reduce: row_labels is null");
    const unsigned int G = q->group_count;
    const unsigned int nwords = (cols + 31u) / 32u;
    L.rows = rows;
    L.cols = cols;
    L.G = G;
    L.nwords = nwords;
    L.nparts = nparts;
    L.vt = value_type;
    L.bell = bell;
    L.aligned = aligned;
    L.labels = q->row_labels;
    L.row_keep = q->row_keep;
    L.feature_keep = q->feature_keep;
    L.scale = q->feature_scale;
    L.moments = q->compute_moments != 0u;
    if (bell) L.bv = parts[0].bell;

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
    // the waves' kept-entry rings, after the accumulators (16-byte aligned)
    const size_t queue_lds = 16u + (size_t) (threads / 64u) * LR_QCAP * sizeof(uint2);
    CSAMD_REQUIRE(fixed_lds + queue_lds + 64u * LR_ACC_BYTES <= lds_avail,
                  "label_reduce: %u columns need %zu bytes of LDS for the feature map", cols, fixed_lds);
    unsigned int per_cu = env_u("CSAMD_LR_PER_CU", 1u);
    unsigned int split = env_u("CSAMD_LR_SPLIT", 1u);
    if (split < 1u || split > 16u) split = 1u;
    L.cu = cu;
    L.threads = threads;

    // Buffers sized by what the host knows (rows, cols, G, partitions); the
    // kept counts stay on the device.
    const unsigned int tiles = rows == 0u ? 1u : (rows + CSAMD_SCAN_TILE - 1u) / CSAMD_SCAN_TILE;
    unsigned int sort_cap = 2u * cu;  // label table (G + 1) x workgroups stays small enough for one scanning workgroup
    if (sort_cap > 262144u / (G + 1u)) sort_cap = 262144u / (G + 1u);
    if (sort_cap < 8u) sort_cap = 8u;
    const unsigned int sort_grid = grid_for(rows == 0u ? 1u : rows, LR_SORT_TILE, sort_cap);
    const size_t table_n = (size_t) (G + 1u) * sort_grid;
    const unsigned int ntail = aligned ? nparts * LR_MAX_TAIL : 0u;
    L.tiles = tiles;
    L.sort_grid = sort_grid;
    L.ntail = ntail;
    L.parts_bytes = (size_t) nparts * sizeof(part_desc);
    if (!ensure_buf(P->fpairs, (size_t) (nwords + 1u) * sizeof(uint2), "hipMalloc fpairs")) return 0;
    if (!ensure_buf(P->fids, ((size_t) cols + 16u + nwords + 1u) * sizeof(unsigned int), "hipMalloc fids")) return 0;
    if (!ensure_buf(P->labels, (table_n + (G + 1u) + 1u) * sizeof(unsigned int), "hipMalloc label table")) return 0;
    if (!ensure_buf(P->rows, (size_t) (rows + 1u) * sizeof(unsigned int), "hipMalloc sorted rows")) return 0;
    const size_t work_bytes = align_up((size_t) (rows + 1u) * sizeof(unsigned int), 256u);
    const size_t toff_bytes = align_up((size_t) (tiles + 1u) * sizeof(unsigned long long), 256u);
    if (!ensure_buf(P->work, work_bytes + toff_bytes + (size_t) (rows + 1u) * sizeof(row_range), "hipMalloc work")) return 0;
    if (!ensure_buf(P->tail, ((size_t) nparts * LR_MAX_TAIL) * sizeof(tail_record), "hipMalloc tail records")) return 0;
    if (!ensure_buf(P->parts, L.parts_bytes, "hipMalloc label partitions")) return 0;
    L.table = (unsigned int *) P->labels.p;  // [(G + 1) x sort_grid]
    L.gstart = L.table + table_n;             // [G + 1]
    L.f_dev = L.gstart + (G + 1u);            // [1], next to gstart[G]
    L.sorted = (unsigned int *) P->rows.p;
    L.work = (unsigned int *) P->work.p;
    L.toff = (unsigned long long *) ((char *) P->work.p + work_bytes);
    L.ranges = (row_range *) ((char *) P->work.p + work_bytes + toff_bytes);
    L.fids = (unsigned int *) P->fids.p;
    L.word_mask = L.fids + (cols + 16u);  // past the kept-feature ids
    L.fpairs = (uint2 *) P->fpairs.p;
    L.tail = ntail != 0u ? (tail_record *) P->tail.p : nullptr;
    L.d_parts = (part_desc *) P->parts.p;

    // partition descriptors: small tables go to the device as kernel
    // arguments, larger ones through pinned staging
    std::vector<part_desc> descs(nparts);
    {
        unsigned int row0 = 0u;
        for (unsigned int p = 0; p < nparts; ++p) {
            const part_in &pi = parts[p];
            part_desc &d = descs[p];
            std::memset(&d, 0, sizeof(d));
            d.idx = pi.idx;
            d.val = pi.val;
            d.ptr = pi.ptr;
            d.kind = pi.kind;
            d.rows = pi.rows;
            d.row0 = row0;
            d.limit = pi.nnz;
            d.vec = aligned ? 1u : 0u;
            if (pi.kind == PART_SELL) {
                d.idx = pi.sell.col_idx;
                d.val = pi.sell.val;
                d.slice_row_offsets = pi.sell.slice_row_offsets;
                d.slice_widths = pi.sell.slice_widths;
                d.slice_slot_offsets = pi.sell.slice_slot_offsets;
                d.slice_count = pi.sell.slice_count;
                d.slice_rows = pi.sell.slice_rows;
            } else if (pi.kind == PART_BELL) {
                d.ell_cols = pi.bell.ell_cols;
            }
            row0 += pi.rows;
        }
    }
    L.inline_parts = nparts <= LR_ARG_PARTS;
    if (L.inline_parts) {
        for (unsigned int p = 0; p < nparts; ++p) L.inl.d[p] = descs[p];
    } else {
        if (P->parts_host_bytes < L.parts_bytes) {
            if (P->parts_host != nullptr) {
                CSAMD_HIP(hipEventSynchronize(P->ev_meta), "label reduce sync (descriptor staging)");
                (void) hipHostFree(P->parts_host);
            }
            P->parts_host = nullptr;
            P->parts_host_bytes = 0u;
            CSAMD_HIP(hipHostMalloc(&P->parts_host, L.parts_bytes, hipHostMallocDefault), "hipHostMalloc label partitions");
            P->parts_host_bytes = L.parts_bytes;
        } else {
            // the previous call's upload has left the staging once its counts copy is done
            CSAMD_HIP(hipEventSynchronize(P->ev_meta), "label reduce sync (descriptor staging)");
        }
        std::memcpy(P->parts_host, descs.data(), L.parts_bytes);
    }

    // the LDS feature tile is sized for all columns (at most F of them are
    // kept; the device reads F), so no host wait is needed
    const unsigned int fmax = cols == 0u ? 1u : cols;
    if (per_cu != 2u || fixed_lds + queue_lds + 64u * LR_ACC_BYTES > lds_avail / 2u) per_cu = 1u;
    const size_t lds_budget = per_cu == 2u ? lds_avail / 2u : lds_avail;
    unsigned int ftile = (unsigned int) ((lds_budget - fixed_lds - queue_lds) / LR_ACC_BYTES);
    if (ftile >= fmax) ftile = fmax;
    else ftile &= ~63u;
    const unsigned int nftiles = (fmax + ftile - 1u) / ftile;
    const unsigned int W = cu * per_cu * split;
    L.fmax = fmax;
    L.W = W;
    L.lds = fixed_lds + (size_t) ftile * LR_ACC_BYTES + queue_lds;
    L.hist_lds = (size_t) (G + 1u) * sizeof(unsigned int);
    L.scatter_lds = (2u * (size_t) (G + 1u) + nparts + 1u) * sizeof(unsigned int);
    L.meta_lds = (size_t) (G + 1u + W + 1u) * sizeof(unsigned int);
    if (!ensure_buf(P->bounds, (size_t) (W + 1u) * sizeof(unsigned int), "hipMalloc bounds")) return 0;
    L.bounds = (unsigned int *) P->bounds.p;

    // outputs [G x F] (room for F = cols) and partial slots
    const size_t GF = (size_t) G * fmax;
    const size_t o_sum = 0, o_sumsq = align_up(o_sum + GF * 8u, 256u), o_mean = align_up(o_sumsq + GF * 8u, 256u),
                 o_var = align_up(o_mean + GF * 8u, 256u), o_cnt = align_up(o_var + GF * 8u, 256u),
                 o_grows = align_up(o_cnt + GF * 4u, 256u), o_end = align_up(o_grows + (size_t) G * 4u, 256u);
    if (!ensure_buf(P->out, o_end, "hipMalloc label outputs")) return 0;
    char *ob = (char *) P->out.p;
    L.out_sum = (double *) (ob + o_sum);
    L.out_sumsq = (double *) (ob + o_sumsq);
    L.out_mean = (double *) (ob + o_mean);
    L.out_var = (double *) (ob + o_var);
    L.out_cnt = (unsigned int *) (ob + o_cnt);
    L.out_grows = (unsigned int *) (ob + o_grows);
    const size_t slots = (size_t) nftiles * ((size_t) W + G) * ftile;
    const size_t p_sum = 0, p_sumsq = align_up(slots * 8u, 256u), p_cnt = align_up(p_sumsq + slots * 8u, 256u),
                 p_end = p_cnt + slots * 4u;
    if (!ensure_buf(P->partial, p_end, "hipMalloc label partials")) return 0;
    char *pb = (char *) P->partial.p;

    reduce_args &a = L.a;
    a.sorted_rows = L.sorted;
    a.gstart = L.gstart;
    a.bounds = L.bounds;
    a.fpairs = L.fpairs;
    a.f_total = L.f_dev;
    a.tail = L.tail;
    a.ntail = L.tail != nullptr ? ntail : 0u;
    a.psum = (double *) (pb + p_sum);
    a.psumsq = (double *) (pb + p_sumsq);
    a.pcnt = (unsigned int *) (pb + p_cnt);
    a.G = G;
    a.cols = cols;
    a.nwords = nwords;
    a.ftile = ftile;
    a.f0 = 0u;
    a.f_count = 0u;

    // kernel attributes are set here, outside any stream capture
    if (rows != 0u && cols != 0u) {
        int ok = 0;
        switch (value_type) {
            case CSAMD_VALUE_F16: ok = launch_rows<CSAMD_VALUE_F16>(bell, L.bv, W, threads, L.lds, nullptr, a, aligned, L.ranges, true); break;
            case CSAMD_VALUE_BF16: ok = launch_rows<CSAMD_VALUE_BF16>(bell, L.bv, W, threads, L.lds, nullptr, a, aligned, L.ranges, true); break;
            case CSAMD_VALUE_F32: ok = launch_rows<CSAMD_VALUE_F32>(bell, L.bv, W, threads, L.lds, nullptr, a, aligned, L.ranges, true); break;
            case CSAMD_VALUE_FP8_E4M3: ok = launch_rows<CSAMD_VALUE_FP8_E4M3>(bell, L.bv, W, threads, L.lds, nullptr, a, aligned, L.ranges, true); break;
        }
        if (!ok) return 0;
    }
    return 1;
}

// The call's launches, in stream order; no allocation, no host wait, no
// memset.  upload: copy the staged descriptor table first (large tables;
// graphs upload theirs once, before capture).  eager: also record the
// counts copy / timing events.
int enqueue_label_reduce(const lr_plan &L, label_private *P, hipStream_t s, bool upload, bool eager) {
    if (!L.inline_parts && upload)
        CSAMD_HIP(hipMemcpyAsync(L.d_parts, P->parts_host, L.parts_bytes, hipMemcpyHostToDevice, s), "H2D label partitions");
    // 1. label histograms and feature words; 2. label offsets and feature map
    label_count_kernel<<<L.sort_grid, LR_SORT_THREADS, L.hist_lds, s>>>(L.labels, L.row_keep, L.rows, L.G, L.table, L.work,
                                                                        L.rows + 1u, L.feature_keep, L.cols, L.nwords, L.word_mask,
                                                                        L.tail, L.ntail, L.d_parts, L.nparts, L.inl,
                                                                        L.inline_parts ? 1u : 0u);
    CSAMD_HIP(hipGetLastError(), "label_count_kernel");
    label_offsets_kernel<<<1, 1024, 0, s>>>(L.table, L.G, L.sort_grid, L.gstart, L.word_mask, L.nwords, L.fpairs, L.fids, L.f_dev);
    CSAMD_HIP(hipGetLastError(), "label_offsets_kernel");
    if (eager) {
        // gstart[G] (kept rows) and F sit next to each other: one copy to
        // pinned memory, waited for at the end of the call (long done by then)
        CSAMD_HIP(hipMemcpyAsync((void *) P->host, L.gstart + L.G, 2u * sizeof(unsigned int), hipMemcpyDeviceToHost, s), "D2H counts");
        CSAMD_HIP(hipEventRecord(P->ev_meta, s), "hipEventRecord");
    }
    // 3. kept rows grouped by label, with range records, tails and work
    if (L.rows != 0u) {
        label_scatter_kernel<<<L.sort_grid, LR_SORT_THREADS, L.scatter_lds, s>>>(L.d_parts, L.nparts, L.labels, L.row_keep, L.rows,
                                                                                 L.G, L.table, L.sorted, L.ranges, L.work, L.tail, L.vt);
        CSAMD_HIP(hipGetLastError(), "label_scatter_kernel");
    }
    // 4. workgroup ranges from the work weights (positions past the kept rows weigh 0)
    if (!scan_tile_offsets_u32(L.work, L.rows, L.toff, s)) return 0;
    work_bounds_kernel<<<L.tiles, 256, 0, s>>>(L.work, L.rows, L.toff, L.tiles, L.W, L.bounds);
    CSAMD_HIP(hipGetLastError(), "work_bounds_kernel");
    // 5. one reduce launch for all feature tiles and partitions, then the finish
    if (eager) P->last_valid = 0;
    if (L.rows != 0u && L.cols != 0u) {
        const bool timed = eager && P->timing != 0;
        if (timed) CSAMD_HIP(hipEventRecord(P->ev0, s), "hipEventRecord");
        int ok = 0;
        switch (L.vt) {
            case CSAMD_VALUE_F16: ok = launch_rows<CSAMD_VALUE_F16>(L.bell, L.bv, L.W, L.threads, L.lds, s, L.a, L.aligned, L.ranges, false); break;
            case CSAMD_VALUE_BF16: ok = launch_rows<CSAMD_VALUE_BF16>(L.bell, L.bv, L.W, L.threads, L.lds, s, L.a, L.aligned, L.ranges, false); break;
            case CSAMD_VALUE_F32: ok = launch_rows<CSAMD_VALUE_F32>(L.bell, L.bv, L.W, L.threads, L.lds, s, L.a, L.aligned, L.ranges, false); break;
            case CSAMD_VALUE_FP8_E4M3: ok = launch_rows<CSAMD_VALUE_FP8_E4M3>(L.bell, L.bv, L.W, L.threads, L.lds, s, L.a, L.aligned, L.ranges, false); break;
        }
        if (!ok) return 0;
        if (timed) {
            CSAMD_HIP(hipEventRecord(P->ev1, s), "hipEventRecord");
            P->last_valid = 1;
        }
    }
    label_finish_kernel<<<grid_for((unsigned long long) L.G * L.fmax, 256u, 4u * L.cu), 256, L.meta_lds, s>>>(
        L.a.psum, L.a.psumsq, L.a.pcnt, L.gstart, L.bounds, L.W, L.G, L.a.ftile, L.f_dev, L.scale, L.fids, L.out_sum, L.out_sumsq,
        L.out_cnt, L.moments ? L.out_mean : nullptr, L.out_var, L.out_grows);
    CSAMD_HIP(hipGetLastError(), "label_finish_kernel");
    return 1;
}

void fill_result(const lr_plan &L, unsigned int kept_rows, unsigned int kept_features, csamd_label_reduce_result *out) {
    out->group_count = L.G;
    out->kept_features = kept_features;
    out->kept_rows = kept_rows;
    out->reserved = 0u;
    out->sum = L.out_sum;
    out->sumsq = L.out_sumsq;
    out->nnz = L.out_cnt;
    out->group_rows = L.out_grows;
    out->mean = L.moments ? L.out_mean : nullptr;
    out->var = L.moments ? L.out_var : nullptr;
    out->kept_feature_ids = L.fids;
}

int run_label_reduce(const std::vector<part_in> &parts, unsigned int value_type, const csamd_label_reduce_query *q,
                     csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(q != nullptr && ws != nullptr && out != nullptr, "label_reduce: null argument");
    CSAMD_HIP(hipSetDevice(ws->device >= 0 ? ws->device : 0), "hipSetDevice label reduce");
    label_private *P = private_of(ws);
    CSAMD_REQUIRE(P != nullptr, "label_reduce: host allocation failed");
    hipStream_t s = (hipStream_t) ws->stream;
    lr_plan *L = (lr_plan *) std::malloc(sizeof(lr_plan));  // ~3 KB of inline descriptors: off the stack
    CSAMD_REQUIRE(L != nullptr, "label_reduce: host allocation failed");
    int ok = plan_label_reduce(parts, value_type, q, P, *L) && enqueue_label_reduce(*L, P, s, true, true);
    if (ok) {
        // kept rows and features, copied right after the label offsets
        ok = hip_ok(hipEventSynchronize(P->ev_meta), "label reduce sync (counts)");
        if (ok) fill_result(*L, P->host[0], P->host[1], out);
    }
    std::free(L);
    return ok;
}

// A captured call: its own buffers, descriptor table and graph.
struct lr_graph_state {
    label_private priv;
    lr_plan plan;
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipEvent_t done;
    int launched;
    int device;
    pl_state pipe;  // the pipeline's conversion tables (zero for a reduce-only graph)
};

void free_private(label_private *P) {
    dev_buf *bufs[] = {&P->fpairs, &P->fids, &P->labels, &P->rows, &P->work, &P->bounds,
                       &P->partial, &P->tail, &P->parts, &P->out};
    for (dev_buf *b : bufs)
        if (b->p != nullptr) (void) hipFree(b->p);
    if (P->host != nullptr) (void) hipHostFree((void *) P->host);
    if (P->parts_host != nullptr) (void) hipHostFree(P->parts_host);
    if (P->events_ready) {
        (void) hipEventDestroy(P->ev0);
        (void) hipEventDestroy(P->ev1);
        (void) hipEventDestroy(P->ev_meta);
    }
    std::memset(P, 0, sizeof(*P));
}

std::vector<part_in> batched_parts(const csamd_label_reduce_partition *parts, unsigned int partition_count, int *ok);

part_in csr_part(const csamd_compressed_view *src) {
    part_in pi;
    std::memset(&pi, 0, sizeof(pi));
    pi.kind = PART_CSR;
    pi.rows = src->rows;
    pi.cols = src->cols;
    pi.idx = src->minorIdx;
    pi.val = src->val;
    pi.nnz = src->nnz;
    pi.ptr = src->majorPtr;
    return pi;
}

part_in sell_part(const csamd_sliced_ell_view *src) {
    part_in pi;
    std::memset(&pi, 0, sizeof(pi));
    pi.kind = PART_SELL;
    pi.rows = src->rows;
    pi.cols = src->cols;
    pi.idx = src->col_idx;
    pi.val = src->val;
    pi.sell = *src;
    return pi;
}

std::vector<part_in> batched_parts(const csamd_label_reduce_partition *parts, unsigned int partition_count, int *ok) {
    std::vector<part_in> v;
    *ok = 0;
    if (parts == nullptr || partition_count < 1u || partition_count > LR_MAX_PARTS) {
        set_error("label_reduce_batched: 1..%u partitions required", LR_MAX_PARTS);
        return v;
    }
    v.reserve(partition_count);
    for (unsigned int p = 0; p < partition_count; ++p) {
        const csamd_label_reduce_partition &lp = parts[p];
        if (lp.layout == CSAMD_LAYOUT_COMPRESSED) {
            if (lp.compressed.axis != CSAMD_COMPRESSED_BY_ROW) {
                set_error("label_reduce_batched: partition %u is not row-major", p);
                return v;
            }
            if (lp.compressed.rows != 0u && lp.compressed.majorPtr == nullptr) {
                set_error("label_reduce_batched: partition %u has no majorPtr", p);
                return v;
            }
            v.push_back(csr_part(&lp.compressed));
        } else if (lp.layout == CSAMD_LAYOUT_SLICED_ELL) {
            v.push_back(sell_part(&lp.sliced_ell));
        } else {
            set_error("label_reduce_batched: partition %u layout %u (compressed or sliced-ELL only)", p, (unsigned) lp.layout);
            return v;
        }
    }
    *ok = 1;
    return v;
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
        if (ws->stream != nullptr) (void) hipStreamSynchronize((hipStream_t) ws->stream);  // kernels may still use the buffers
        free_private(P);
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
    return run_label_reduce(std::vector<part_in>{csr_part(src)}, query->value_type, query, ws, out);
}

int csamd_label_reduce_sliced_ell(const csamd_sliced_ell_view *src, const csamd_label_reduce_query *query,
                                  csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(src != nullptr && query != nullptr, "label_reduce_sliced_ell: null argument");
    return run_label_reduce(std::vector<part_in>{sell_part(src)}, query->value_type, query, ws, out);
}

int csamd_label_reduce_blocked_ell(const csamd_blocked_ell_view *src, const csamd_label_reduce_query *query,
                                   csamd_label_reduce_workspace *ws, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(src != nullptr && query != nullptr, "label_reduce_blocked_ell: null argument");
    CSAMD_REQUIRE(src->block_size != 0u || src->ell_cols == 0u, "label_reduce_blocked_ell: block_size is 0");
    part_in pi;
    std::memset(&pi, 0, sizeof(pi));
    pi.kind = PART_BELL;
    pi.rows = src->rows;
    pi.cols = src->cols;
    pi.bell = *src;
    return run_label_reduce(std::vector<part_in>{pi}, query->value_type, query, ws, out);
}

int csamd_label_reduce_batched(const csamd_label_reduce_partition *parts, unsigned int partition_count,
                               const csamd_label_reduce_query *query, csamd_label_reduce_workspace *ws,
                               csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(query != nullptr, "label_reduce_batched: null query");
    int ok = 0;
    std::vector<part_in> v = batched_parts(parts, partition_count, &ok);
    if (!ok) return 0;
    return run_label_reduce(v, query->value_type, query, ws, out);
}

int csamd_label_reduce_graph_capture(const csamd_label_reduce_partition *parts, unsigned int partition_count,
                                     const csamd_label_reduce_query *query, int device, csamd_label_reduce_graph *graph) {
    CSAMD_REQUIRE(graph != nullptr && query != nullptr, "label_reduce_graph_capture: null argument");
    std::memset(graph, 0, sizeof(*graph));
    int ok = 0;
    std::vector<part_in> v = batched_parts(parts, partition_count, &ok);
    if (!ok) return 0;
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice label reduce graph");
    lr_graph_state *G = (lr_graph_state *) std::calloc(1u, sizeof(lr_graph_state));
    CSAMD_REQUIRE(G != nullptr, "label_reduce_graph_capture: host allocation failed");
    G->device = device;
    graph->device = device;
    graph->state = G;
    hipStream_t cs = nullptr;
    ok = plan_label_reduce(v, query->value_type, query, &G->priv, G->plan) &&
         hip_ok(hipEventCreateWithFlags(&G->done, hipEventDisableTiming), "hipEventCreate graph");
    if (ok && !G->plan.inline_parts)  // the graph's own descriptor table, uploaded once
        ok = hip_ok(hipMemcpy(G->plan.d_parts, G->priv.parts_host, G->plan.parts_bytes, hipMemcpyHostToDevice), "H2D graph partitions");
    if (ok) ok = hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate capture");
    if (ok) ok = hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (ok) {
        const int enq = enqueue_label_reduce(G->plan, &G->priv, cs, false, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);  // always ends the capture
        G->graph = g;
        ok = enq && hip_ok(e, "hipStreamEndCapture");
    }
    if (ok) ok = hip_ok(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (cs != nullptr) (void) hipStreamDestroy(cs);
    if (!ok) {
        csamd_label_reduce_graph_destroy(graph);
        return 0;
    }
    return 1;
}

int csamd_label_reduce_graph_launch(csamd_label_reduce_graph *graph, csamd_stream_t stream) {
    CSAMD_REQUIRE(graph != nullptr && graph->state != nullptr, "label_reduce_graph_launch: no graph");
    lr_graph_state *G = (lr_graph_state *) graph->state;
    CSAMD_HIP(hipSetDevice(G->device), "hipSetDevice label reduce graph");
    CSAMD_HIP(hipGraphLaunch(G->exec, (hipStream_t) stream), "hipGraphLaunch");
    CSAMD_HIP(hipEventRecord(G->done, (hipStream_t) stream), "hipEventRecord graph");
    G->launched = 1;
    return 1;
}

int csamd_label_reduce_graph_result(csamd_label_reduce_graph *graph, csamd_label_reduce_result *out) {
    CSAMD_REQUIRE(graph != nullptr && graph->state != nullptr && out != nullptr, "label_reduce_graph_result: null argument");
    lr_graph_state *G = (lr_graph_state *) graph->state;
    CSAMD_REQUIRE(G->launched, "label_reduce_graph_result: the graph has not been launched");
    CSAMD_HIP(hipSetDevice(G->device), "hipSetDevice label reduce graph");
    CSAMD_HIP(hipEventSynchronize(G->done), "label reduce graph sync");
    unsigned int counts[2] = {0u, 0u};
    CSAMD_HIP(hipMemcpy(counts, G->plan.gstart + G->plan.G, sizeof(counts), hipMemcpyDeviceToHost), "D2H graph counts");
    if (G->pipe.block != nullptr) {
        unsigned int err = 0u;
        CSAMD_HIP(hipMemcpy(&err, G->pipe.error, 4u, hipMemcpyDeviceToHost), "D2H pipeline error word");
        CSAMD_REQUIRE(err == 0u, "label_reduce pipeline: %s",
                      (err & 1u) ? "a row block holds more block columns than its Blocked-ELL ell width"
                                 : "a Sliced-ELL layout needs more slots than its arrays hold");
    }
    fill_result(G->plan, counts[0], counts[1], out);
    return 1;
}

void csamd_label_reduce_graph_destroy(csamd_label_reduce_graph *graph) {
    if (graph == nullptr || graph->state == nullptr) return;
    lr_graph_state *G = (lr_graph_state *) graph->state;
    (void) hipSetDevice(G->device);
    if (G->launched) (void) hipEventSynchronize(G->done);  // replays may still use the buffers
    if (G->exec != nullptr) (void) hipGraphExecDestroy(G->exec);
    if (G->graph != nullptr) (void) hipGraphDestroy(G->graph);
    if (G->done != nullptr) (void) hipEventDestroy(G->done);
    free_private(&G->priv);
    pl_free(&G->pipe);
    std::free(G);
    graph->state = nullptr;
}

int csamd_label_reduce_pipeline_capture(const csamd_pipeline_partition *parts, unsigned int partition_count,
                                        const csamd_label_reduce_query *query, int device, csamd_label_reduce_graph *graph) {
    CSAMD_REQUIRE(graph != nullptr && query != nullptr && parts != nullptr, "label_reduce_pipeline_capture: null argument");
    CSAMD_REQUIRE(partition_count >= 1u && partition_count <= LR_MAX_PARTS, "label_reduce_pipeline_capture: 1..%u partitions",
                  LR_MAX_PARTS);
    std::memset(graph, 0, sizeof(*graph));
    const int vb = value_type_bytes(query->value_type);
    // the reduce's partitions: Sliced-ELL ones after conversion, the rest as CSR
    std::vector<part_in> v;
    v.reserve(partition_count);
    for (unsigned int p = 0; p < partition_count; ++p) {
        const csamd_pipeline_partition &P = parts[p];
        CSAMD_REQUIRE(P.layout == CSAMD_LAYOUT_COMPRESSED || P.layout == CSAMD_LAYOUT_SLICED_ELL || P.layout == CSAMD_LAYOUT_BLOCKED_ELL,
                      "label_reduce_pipeline_capture: partition %u layout %u", p, P.layout);
        CSAMD_REQUIRE(P.compressed.axis == CSAMD_COMPRESSED_BY_ROW && (P.compressed.rows == 0u || P.compressed.majorPtr != nullptr),
                      "label_reduce_pipeline_capture: partition %u needs a row-major device CSR", p);
        if (P.staged_from != nullptr) {
            const csamd_host_compressed &h = *P.staged_from;
            CSAMD_REQUIRE(h.rows == P.compressed.rows && h.nnz == P.compressed.nnz && h.axis == CSAMD_COMPRESSED_BY_ROW &&
                              h.majorPtr != nullptr && (h.nnz == 0u || (h.minorIdx != nullptr && h.val != nullptr)),
                          "label_reduce_pipeline_capture: partition %u staging source does not match its device CSR", p);
        }
        v.push_back(P.layout == CSAMD_LAYOUT_SLICED_ELL ? sell_part(&P.sliced_ell) : csr_part(&P.compressed));
    }
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice label reduce pipeline");
    lr_graph_state *G = (lr_graph_state *) std::calloc(1u, sizeof(lr_graph_state));
    CSAMD_REQUIRE(G != nullptr, "label_reduce_pipeline_capture: host allocation failed");
    G->device = device;
    graph->device = device;
    graph->state = G;
    hipStream_t cs = nullptr;
    int ok = plan_label_reduce(v, query->value_type, query, &G->priv, G->plan) &&
             pl_prepare(parts, partition_count, query->value_type, &G->pipe) &&
             hip_ok(hipEventCreateWithFlags(&G->done, hipEventDisableTiming), "hipEventCreate pipeline");
    if (ok && !G->plan.inline_parts)
        ok = hip_ok(hipMemcpy(G->plan.d_parts, G->priv.parts_host, G->plan.parts_bytes, hipMemcpyHostToDevice), "H2D pipeline partitions");
    if (ok) ok = hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate capture");
    if (ok) ok = hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (ok) {
        int enq = 1;
        // 1. staging: pinned host CSR arrays into the partitions' device arrays
        for (unsigned int p = 0; enq && p < partition_count; ++p) {
            const csamd_pipeline_partition &P = parts[p];
            if (P.staged_from == nullptr) continue;
            const csamd_host_compressed &h = *P.staged_from;
            enq = hip_ok(hipMemcpyAsync(P.compressed.majorPtr, h.majorPtr, ((size_t) h.rows + 1u) * 4u, hipMemcpyHostToDevice, cs),
                         "H2D staged row pointers");
            if (enq && h.nnz != 0u)
                enq = hip_ok(hipMemcpyAsync(P.compressed.minorIdx, h.minorIdx, (size_t) h.nnz * 4u, hipMemcpyHostToDevice, cs),
                             "H2D staged indices") &&
                      hip_ok(hipMemcpyAsync(P.compressed.val, h.val, (size_t) h.nnz * (size_t) vb, hipMemcpyHostToDevice, cs),
                             "H2D staged values");
        }
        // 2. conversion of every Sliced-ELL / Blocked-ELL partition; 3. the reduce
        enq = enq && pl_enqueue(G->pipe, cs) && enqueue_label_reduce(G->plan, &G->priv, cs, false, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);  // always ends the capture
        G->graph = g;
        ok = enq && hip_ok(e, "hipStreamEndCapture");
    }
    if (ok) ok = hip_ok(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (cs != nullptr) (void) hipStreamDestroy(cs);
    if (!ok) {
        csamd_label_reduce_graph_destroy(graph);
        return 0;
    }
    return 1;
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
