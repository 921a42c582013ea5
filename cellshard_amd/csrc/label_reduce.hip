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
    dev_buf fcounts;     // per-word popcounts / scan tmp
    dev_buf fids;        // kept feature ids
    dev_buf keys;        // sort keys in/out
    dev_buf rows;        // sort values in/out
    dev_buf sort_tmp;
    dev_buf gstart;      // [G+1]
    dev_buf work;        // u32 per sorted row + u64 prefix
    dev_buf bounds;      // [W+1]
    dev_buf partial;     // (W+G) x Ftile x {f64, f64, u32}
    dev_buf tail;        // tail records
    dev_buf out;         // outputs
    hipEvent_t ev0, ev1;
    int events_ready;
    float last_ms;
    int last_valid;
};

// ---------------------------------------------------------------------------
// Feature map.
// ---------------------------------------------------------------------------
__global__ void feature_words_kernel(const unsigned char *__restrict__ feature_keep, unsigned int cols, unsigned int nwords,
                                     unsigned int *__restrict__ words, unsigned int *__restrict__ counts) {
    for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x) {
        unsigned int m = 0u;
        const unsigned int c0 = w * 32u;
        for (unsigned int b = 0; b < 32u; ++b) {
            const unsigned int c = c0 + b;
            if (c < cols && (feature_keep == nullptr || feature_keep[c] != 0u)) m |= 1u << b;
        }
        words[w] = m;
        counts[w] = (unsigned int) __popc(m);
    }
}

__global__ void feature_pairs_kernel(const unsigned int *__restrict__ words, const unsigned int *__restrict__ prefix,
                                     unsigned int nwords, uint2 *__restrict__ pairs, unsigned int *__restrict__ fids) {
    for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x) {
        const unsigned int m = words[w];
        const unsigned int p = prefix[w];
        pairs[w] = make_uint2(m, p);
        unsigned int mm = m, k = 0u;
        while (mm != 0u) {
            const unsigned int b = (unsigned int) __builtin_ctz(mm);
            fids[p + k] = w * 32u + b;
            ++k;
            mm &= mm - 1u;
        }
    }
}

// ---------------------------------------------------------------------------
// Row keys: label if kept else G (dropped); identity values.
// ---------------------------------------------------------------------------
__global__ void row_keys_kernel(const unsigned int *__restrict__ labels, const unsigned char *__restrict__ row_keep,
                                unsigned int rows, unsigned int G, unsigned int *__restrict__ keys,
                                unsigned int *__restrict__ ids) {
    for (unsigned int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        const unsigned int l = labels[r];
        const bool keep = (row_keep == nullptr || row_keep[r] != 0u) && l < G;
        keys[r] = keep ? l : G;
        ids[r] = r;
    }
}

// gstart[g] = first sorted position with key >= g, g = 0..G.
__global__ void group_starts_kernel(const unsigned int *__restrict__ sorted_keys, unsigned int n, unsigned int G,
                                    unsigned int *__restrict__ gstart) {
    for (unsigned int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
        unsigned int lo = 0u, hi = n;
        while (lo < hi) {
            const unsigned int mid = lo + ((hi - lo) >> 1);
            if (sorted_keys[mid] < g) lo = mid + 1u;
            else hi = mid;
        }
        gstart[g] = lo;
    }
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

// Per sorted kept row: its entry range (both layouts index entries with u32)
// and its work weight for the partition.  Ranges are clipped at entry clip:
// the vector-load kernel streams only whole 4-entry units, and the at most 3
// entries past the last whole unit are tail records (tail_records_kernel).
template<typename Geom>
__global__ void row_work_kernel(Geom geom, const unsigned int *__restrict__ sorted_rows, unsigned int n,
                                unsigned long long clip, unsigned int *__restrict__ work, uint2 *__restrict__ ranges) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long b, e;
        geom.range(sorted_rows[i], &b, &e);
        const unsigned long long len = e - b;
        work[i] = (unsigned int) (len > 0xfffffff0ull ? 0xfffffff0ull : len) + 8u;  // +8: per-row overhead
        if (ranges != nullptr) ranges[i] = make_uint2((unsigned int) (b < clip ? b : clip), (unsigned int) (e < clip ? e : clip));
    }
}

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

// Entries [clip, limit) (at most 3): find each one's row, apply the row and
// feature masks and write a tail record.  One thread; runs before the main
// kernel and is read by the combine.
template<typename Geom>
__global__ void tail_records_kernel(Geom geom, unsigned int rows, const unsigned int *__restrict__ idx, const void *val,
                                    unsigned int vt, unsigned long long clip, unsigned long long limit, unsigned int cols,
                                    const unsigned int *__restrict__ labels, const unsigned char *__restrict__ row_keep,
                                    unsigned int G, const uint2 *__restrict__ fpairs, tail_record *__restrict__ out) {
    if (blockIdx.x != 0u || threadIdx.x != 0u) return;
    for (unsigned int t = 0; t < LR_MAX_TAIL; ++t) {
        tail_record rec{G, CSAMD_INVALID_COL, 0.0};
        const unsigned long long i = clip + t;
        if (i < limit) {
            const unsigned int r = geom.row_of(i, rows);
            const unsigned int c = idx[i];
            if (r < rows && (row_keep == nullptr || row_keep[r] != 0u) && labels[r] < G && c < cols) {
                const uint2 pm = fpairs[c >> 5];
                const unsigned int bit = c & 31u;
                if ((pm.x >> bit) & 1u) {
                    rec.g = labels[r];
                    rec.f = pm.y + (unsigned int) __popc(pm.x & ((1u << bit) - 1u));
                    rec.v = (double) decode_any(val, i, vt);
                }
            }
        }
        out[t] = rec;
    }
}

__global__ void blocked_work_kernel(unsigned int ell_cols, unsigned int n, unsigned int *__restrict__ work) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) work[i] = ell_cols + 1u;
}

// bounds[w] = first i with prefix[i] >= w * total / W; bounds[W] = n.
__global__ void work_bounds_kernel(const unsigned long long *__restrict__ prefix, unsigned int n, unsigned int W,
                                   unsigned int *__restrict__ bounds) {
    const unsigned long long total = prefix[n];
    for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w <= W; w += gridDim.x * blockDim.x) {
        if (w == W) {
            bounds[w] = n;
            continue;
        }
        const unsigned long long target = (total / W) * w + ((total % W) * w) / W;
        unsigned int lo = 0u, hi = n;
        while (lo < hi) {
            const unsigned int mid = lo + ((hi - lo) >> 1);
            if (prefix[mid] < target) lo = mid + 1u;
            else hi = mid;
        }
        bounds[w] = lo;
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
    double *psum;                 // [(W+G) x ftile]
    double *psumsq;
    unsigned int *pcnt;
    unsigned int G;
    unsigned int cols;
    unsigned int nwords;
    unsigned int ftile;           // features held in LDS per pass
    unsigned int f0;              // first feature of this pass
    unsigned int f_count;         // features in this pass
};

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
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) {
        return fp8_e4m3_to_float((uint8_t) (r >> (8u * k)));
    }
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

    const unsigned int lo = a.bounds[blockIdx.x];
    const unsigned int hi = a.bounds[blockIdx.x + 1u];
    if (lo >= hi) return;  // uniform across the block
    unsigned int g;
    {
        unsigned int l = 0u, h = a.G;
        while (l < h) {  // largest g with gstart[g] <= lo
            const unsigned int mid = (l + h + 1u) >> 1;
            if (a.gstart[mid] <= lo) l = mid;
            else h = mid - 1u;
        }
        g = l;
    }
    for (; g < a.G; ++g) {
        const unsigned int gs = a.gstart[g], ge = a.gstart[g + 1u];
        const unsigned int s0 = gs > lo ? gs : lo;
        const unsigned int s1 = ge < hi ? ge : hi;
        if (s0 >= hi) break;
        if (s0 >= s1) continue;
        for (unsigned int f = threadIdx.x; f < a.f_count; f += nthreads) {
            lsum[f] = 0.0;
            lsumsq[f] = 0.0;
            lcnt[f] = 0u;
        }
        if (threadIdx.x == 0u) next_row = s0;
        __syncthreads();
        stream_rows<VT, VEC>(idx, val, ranges, &next_row, s1, last, lpairs, a, acc);
        __syncthreads();
        const unsigned long long slot = (unsigned long long) (blockIdx.x + g) * a.ftile;
        for (unsigned int f = threadIdx.x; f < a.f_count; f += nthreads) {
            a.psum[slot + f] = lsum[f];
            a.psumsq[slot + f] = lsumsq[f];
            a.pcnt[slot + f] = lcnt[f];
        }
        __syncthreads();
    }
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
    const unsigned int lo = a.bounds[blockIdx.x];
    const unsigned int hi = a.bounds[blockIdx.x + 1u];
    if (lo >= hi) return;
    unsigned int g;
    {
        unsigned int l = 0u, h = a.G;
        while (l < h) {
            const unsigned int mid = (l + h + 1u) >> 1;
            if (a.gstart[mid] <= lo) l = mid;
            else h = mid - 1u;
        }
        g = l;
    }
    for (; g < a.G; ++g) {
        const unsigned int gs = a.gstart[g], ge = a.gstart[g + 1u];
        const unsigned int s0 = gs > lo ? gs : lo;
        const unsigned int s1 = ge < hi ? ge : hi;
        if (s0 >= hi) break;
        if (s0 >= s1) continue;
        for (unsigned int f = threadIdx.x; f < a.f_count; f += nthreads) {
            lsum[f] = 0.0;
            lsumsq[f] = 0.0;
            lcnt[f] = 0u;
        }
        __syncthreads();
        for (unsigned int i = s0 + wave; i < s1; i += nwaves) {
            const unsigned int row = a.sorted_rows[i];
            const unsigned int rb = bs != 0u ? row / bs : 0u;
            const unsigned long long base = (unsigned long long) row * src.ell_cols;
            for (unsigned int e = lane_id(); e < src.ell_cols; e += 64u) {
                const unsigned int slot = e / bs;
                const unsigned int bc = src.blockColIdx[(unsigned long long) rb * width + slot];
                const unsigned int c = bc != CSAMD_INVALID_COL ? bc * bs + (e - slot * bs) : CSAMD_INVALID_COL;
                consume<VT>(c, val[base + e], lpairs, a, acc);
            }
        }
        __syncthreads();
        const unsigned long long slot = (unsigned long long) (blockIdx.x + g) * a.ftile;
        for (unsigned int f = threadIdx.x; f < a.f_count; f += nthreads) {
            a.psum[slot + f] = lsum[f];
            a.psumsq[slot + f] = lsumsq[f];
            a.pcnt[slot + f] = lcnt[f];
        }
        __syncthreads();
    }
}

// Combine: out[g][f0 + f] = sum over touching workgroups w (ascending) of
// slot (w + g); then the feature's scale s (if any): s * sum, s^2 * sumsq,
// and no nonzeros when s == 0.
__global__ void label_combine_kernel(const double *__restrict__ psum, const double *__restrict__ psumsq,
                                     const unsigned int *__restrict__ pcnt, const unsigned int *__restrict__ gstart,
                                     const unsigned int *__restrict__ bounds, unsigned int W, unsigned int G, unsigned int ftile,
                                     unsigned int f0, unsigned int f_count, unsigned int F, const float *__restrict__ scale,
                                     const unsigned int *__restrict__ fids, const tail_record *__restrict__ tail,
                                     double *__restrict__ out_sum, double *__restrict__ out_sumsq,
                                     unsigned int *__restrict__ out_cnt) {
    const unsigned long long total = (unsigned long long) G * f_count;
    for (unsigned long long t = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; t < total;
         t += (unsigned long long) gridDim.x * blockDim.x) {
        const unsigned int g = (unsigned int) (t / f_count);
        const unsigned int f = (unsigned int) (t - (unsigned long long) g * f_count);
        const unsigned int gs = gstart[g], ge = gstart[g + 1u];
        double s = 0.0, q = 0.0;
        unsigned int n = 0u;
        if (gs < ge) {
            // first w with bounds[w+1] > gs
            unsigned int l = 0u, h = W - 1u;
            while (l < h) {
                const unsigned int mid = (l + h) >> 1;
                if (bounds[mid + 1u] > gs) h = mid;
                else l = mid + 1u;
            }
            for (unsigned int w = l; w < W && bounds[w] < ge; ++w) {
                const unsigned int a0 = bounds[w] > gs ? bounds[w] : gs;
                const unsigned int a1 = bounds[w + 1u] < ge ? bounds[w + 1u] : ge;
                if (a0 >= a1) continue;
                const unsigned long long slot = (unsigned long long) (w + g) * ftile + f;
                s += psum[slot];
                q += psumsq[slot];
                n += pcnt[slot];
            }
        }
        if (tail != nullptr) {
#pragma unroll
            for (unsigned int t = 0; t < LR_MAX_TAIL; ++t)
                if (tail[t].g == g && tail[t].f == f0 + f && tail[t].v != 0.0) {
                    s += tail[t].v;
                    q += tail[t].v * tail[t].v;
                    n += 1u;
                }
        }
        if (scale != nullptr) {
            const double sc = (double) scale[fids[f0 + f]];
            s *= sc;
            q *= sc * sc;
            if (sc == 0.0) n = 0u;
        }
        const unsigned long long o = (unsigned long long) g * F + f0 + f;
        out_sum[o] = s;
        out_sumsq[o] = q;
        out_cnt[o] = n;
    }
}

__global__ void label_moments_kernel(const double *__restrict__ sum, const double *__restrict__ sumsq,
                                     const unsigned int *__restrict__ gstart, unsigned int G, unsigned int F,
                                     double *__restrict__ mean, double *__restrict__ var, unsigned int *__restrict__ group_rows) {
    const unsigned long long total = (unsigned long long) G * F;
    for (unsigned long long t = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; t < total;
         t += (unsigned long long) gridDim.x * blockDim.x) {
        const unsigned int g = (unsigned int) (t / F);
        const unsigned int n = gstart[g + 1u] - gstart[g];
        if (mean != nullptr) {
            if (n == 0u) {
                mean[t] = 0.0;
                var[t] = 0.0;
            } else {
                const double m = sum[t] / (double) n;
                double v = sumsq[t] / (double) n - m * m;
                if (v < 0.0) v = 0.0;
                mean[t] = m;
                var[t] = v;
            }
        }
    }
    for (unsigned int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
        group_rows[g] = gstart[g + 1u] - gstart[g];
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

unsigned int bits_for(unsigned int v) {
    unsigned int b = 0u;
    while (b < 32u && (v >> b) != 0u) ++b;
    return b;
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
    CSAMD_HIP(hipSetDevice(ws->device >= 0 ? ws->device : 0), "hipSetDevice label reduce");
    label_private *P = private_of(ws);
    CSAMD_REQUIRE(P != nullptr, "label_reduce: host allocation failed");
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int G = q->group_count;
    const unsigned int rows = L.rows, cols = L.cols;
    const unsigned int nwords = (cols + 31u) / 32u;

    // 1. feature map
    const size_t scan_n_bytes = exclusive_scan_scratch_bytes(nwords);
    if (!ensure_buf(P->fpairs, (size_t) (nwords + 1u) * sizeof(uint2), "hipMalloc fpairs")) return 0;
    if (!ensure_buf(P->fcounts, (size_t) (2u * (nwords + 1u)) * sizeof(unsigned int) + 256u + scan_n_bytes, "hipMalloc fcounts"))
        return 0;
    unsigned int *words = (unsigned int *) P->fcounts.p;
    unsigned int *wprefix = words + (nwords + 1u);
    void *scan_tmp = (char *) P->fcounts.p + align_up((size_t) (2u * (nwords + 1u)) * sizeof(unsigned int), 256u);
    unsigned int *counts = nullptr;
    {
        // per-word popcounts live past the kept-feature ids (at most cols of them)
        if (!ensure_buf(P->fids, (size_t) (nwords + 1u) * sizeof(unsigned int) + (size_t) cols * sizeof(unsigned int) + 64u,
                        "hipMalloc fids"))
            return 0;
        counts = (unsigned int *) P->fids.p + cols + 16u;
        if (nwords != 0u) {
            feature_words_kernel<<<grid_for(nwords, 256u, 4096u), 256, 0, s>>>(q->feature_keep, cols, nwords, words, counts);
            CSAMD_HIP(hipGetLastError(), "feature_words_kernel");
        }
    }
    if (!exclusive_scan_u32(counts, nwords, wprefix, nullptr, scan_tmp, scan_n_bytes, s)) return 0;
    unsigned int *fids = (unsigned int *) P->fids.p;
    if (nwords != 0u) {
        feature_pairs_kernel<<<grid_for(nwords, 256u, 4096u), 256, 0, s>>>(words, wprefix, nwords, (uint2 *) P->fpairs.p, fids);
        CSAMD_HIP(hipGetLastError(), "feature_pairs_kernel");
    }
    unsigned int F = 0u;
    CSAMD_HIP(hipMemcpyAsync(&F, wprefix + nwords, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "D2H kept features");

    // 2. rows sorted by label
    const size_t sort_bytes = radix_sort_pairs_u32_scratch_bytes(rows);
    if (!ensure_buf(P->keys, (size_t) 2u * rows * sizeof(unsigned int) + 64u, "hipMalloc keys")) return 0;
    if (!ensure_buf(P->rows, (size_t) 2u * rows * sizeof(unsigned int) + 64u, "hipMalloc rows")) return 0;
    if (!ensure_buf(P->sort_tmp, sort_bytes, "hipMalloc sort tmp")) return 0;
    if (!ensure_buf(P->gstart, (size_t) (G + 1u) * sizeof(unsigned int), "hipMalloc gstart")) return 0;
    unsigned int *keys_in = (unsigned int *) P->keys.p, *keys_out = keys_in + rows;
    unsigned int *ids_in = (unsigned int *) P->rows.p, *ids_out = ids_in + rows;
    unsigned int *gstart = (unsigned int *) P->gstart.p;
    if (rows != 0u) {
        row_keys_kernel<<<grid_for(rows, 256u, 8192u), 256, 0, s>>>(q->row_labels, q->row_keep, rows, G, keys_in, ids_in);
        CSAMD_HIP(hipGetLastError(), "row_keys_kernel");
        if (!radix_sort_pairs_u32(keys_in, keys_out, ids_in, ids_out, rows, bits_for(G), P->sort_tmp.p, P->sort_tmp.bytes, s))
            return 0;
    }
    group_starts_kernel<<<grid_for(G + 1u, 256u, 64u), 256, 0, s>>>(keys_out, rows, G, gstart);
    CSAMD_HIP(hipGetLastError(), "group_starts_kernel");
    unsigned int K = 0u;
    CSAMD_HIP(hipMemcpyAsync(&K, gstart + G, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "D2H kept rows");
    CSAMD_HIP(hipStreamSynchronize(s), "label reduce sync (sizes)");

    // 3. launch shape
    const size_t fixed_lds = align_up((size_t) (nwords + 1u) * sizeof(uint2), 16u);
    const int max_lds = max_lds_per_block();
    CSAMD_REQUIRE(fixed_lds + 64u * LR_ACC_BYTES <= (size_t) max_lds,
                  "label_reduce: %u columns need %zu bytes of LDS for the feature map", cols, fixed_lds);
    const unsigned int cu = (unsigned int) device_cu_count();
    // One 1024-thread workgroup per CU: its waves share the CU's row range
    // through an LDS counter.  (Two workgroups per CU run unevenly: the
    // older one gets issue priority and the younger one finishes late at
    // half occupancy.)  CSAMD_LR_THREADS / CSAMD_LR_PER_CU / CSAMD_LR_SPLIT
    // override the shape for tuning sweeps.
    const size_t lds_avail = (size_t) max_lds - 16u;  // static LDS: the row counter
    auto env_u = [](const char *name, unsigned int dflt) {
        const char *e = std::getenv(name);
        return e != nullptr ? (unsigned int) std::atoi(e) : dflt;
    };
    unsigned int threads = env_u("CSAMD_LR_THREADS", 1024u);
    if (threads < 256u || threads > 1024u || threads % 64u != 0u) threads = 1024u;
    unsigned int per_cu = env_u("CSAMD_LR_PER_CU", 1u);
    const size_t need_all = fixed_lds + (size_t) (F == 0u ? 1u : F) * LR_ACC_BYTES;
    if (per_cu != 2u || need_all > lds_avail / 2u) per_cu = 1u;
    unsigned int ftile = F == 0u ? 1u : F;
    if (need_all > lds_avail) {
        ftile = (unsigned int) ((lds_avail - fixed_lds) / LR_ACC_BYTES);
        ftile &= ~63u;
        if (ftile > F) ftile = F;
        if (ftile == 0u) ftile = 64u;
    }
    unsigned int split = env_u("CSAMD_LR_SPLIT", 1u);
    if (split < 1u || split > 16u) split = 1u;
    const unsigned int W = cu * per_cu * split;
    const size_t lds = fixed_lds + (size_t) ftile * LR_ACC_BYTES;

    // 4. work bounds
    const size_t work_scan_bytes = exclusive_scan_scratch_bytes(K);
    const size_t work_bytes = align_up((size_t) (K + 1u) * sizeof(unsigned int), 256u);
    const size_t prefix_bytes = align_up((size_t) (K + 1u) * sizeof(unsigned long long), 256u);
    const size_t ranges_bytes = align_up((size_t) (K + 1u) * sizeof(uint2), 256u);
    if (!ensure_buf(P->work, work_bytes + prefix_bytes + ranges_bytes + work_scan_bytes, "hipMalloc work")) return 0;
    unsigned int *work = (unsigned int *) P->work.p;
    unsigned long long *wprefix64 = (unsigned long long *) ((char *) P->work.p + work_bytes);
    uint2 *ranges = (uint2 *) ((char *) P->work.p + work_bytes + prefix_bytes);
    void *wscan = (char *) P->work.p + work_bytes + prefix_bytes + ranges_bytes;
    if (!ensure_buf(P->bounds, (size_t) (W + 1u) * sizeof(unsigned int), "hipMalloc bounds")) return 0;
    unsigned int *bounds = (unsigned int *) P->bounds.p;
    // Vector loads need 16-byte aligned indices and unit-aligned values; the
    // vector kernel streams whole 4-entry units only (ranges clipped at
    // clip) and the entries past clip become tail records.
    const bool aligned = L.kind != 2 && ((uintptr_t) L.idx % 16u == 0u) &&
                         ((uintptr_t) L.val % (uintptr_t) (value_type == CSAMD_VALUE_F32 ? 16u
                                                           : value_type == CSAMD_VALUE_FP8_E4M3 ? 4u : 8u) == 0u);
    const unsigned long long clip = aligned ? (L.limit & ~3ull) : ~0ull;
    const tail_record *tail = nullptr;
    if (aligned && clip < L.limit) {
        if (!ensure_buf(P->tail, LR_MAX_TAIL * sizeof(tail_record), "hipMalloc tail records")) return 0;
        tail_record *t = (tail_record *) P->tail.p;
        if (L.kind == 0)
            tail_records_kernel<csr_geom><<<1, 64, 0, s>>>(L.csr, rows, L.idx, L.val, value_type, clip, L.limit, cols, q->row_labels,
                                                           q->row_keep, G, (const uint2 *) P->fpairs.p, t);
        else
            tail_records_kernel<sell_geom><<<1, 64, 0, s>>>(L.sell, rows, L.idx, L.val, value_type, clip, L.limit, cols,
                                                            q->row_labels, q->row_keep, G, (const uint2 *) P->fpairs.p, t);
        CSAMD_HIP(hipGetLastError(), "tail_records_kernel");
        tail = t;
    }
    if (K != 0u) {
        if (L.kind == 0)
            row_work_kernel<csr_geom><<<grid_for(K, 256u, 8192u), 256, 0, s>>>(L.csr, ids_out, K, clip, work, ranges);
        else if (L.kind == 1)
            row_work_kernel<sell_geom><<<grid_for(K, 256u, 8192u), 256, 0, s>>>(L.sell, ids_out, K, clip, work, ranges);
        else blocked_work_kernel<<<grid_for(K, 256u, 8192u), 256, 0, s>>>(L.bell.ell_cols, K, work);
        CSAMD_HIP(hipGetLastError(), "row_work_kernel");
    }
    if (!exclusive_scan_u32_to_u64(work, K, wprefix64, wscan, work_scan_bytes, s)) return 0;
    work_bounds_kernel<<<grid_for(W + 1u, 256u, 64u), 256, 0, s>>>(wprefix64, K, W, bounds);
    CSAMD_HIP(hipGetLastError(), "work_bounds_kernel");

    // 5. outputs
    const size_t GF = (size_t) G * F;
    const size_t o_sum = 0, o_sumsq = align_up(o_sum + GF * 8u, 256u), o_mean = align_up(o_sumsq + GF * 8u, 256u),
                 o_var = align_up(o_mean + GF * 8u, 256u), o_cnt = align_up(o_var + GF * 8u, 256u),
                 o_grows = align_up(o_cnt + GF * 4u, 256u), o_end = align_up(o_grows + (size_t) G * 4u, 256u);
    if (!ensure_buf(P->out, o_end, "hipMalloc label outputs")) return 0;
    char *ob = (char *) P->out.p;
    double *out_sum = (double *) (ob + o_sum), *out_sumsq = (double *) (ob + o_sumsq);
    double *out_mean = (double *) (ob + o_mean), *out_var = (double *) (ob + o_var);
    unsigned int *out_cnt = (unsigned int *) (ob + o_cnt), *out_grows = (unsigned int *) (ob + o_grows);

    const size_t slots = (size_t) W + G;
    const size_t p_sum = 0, p_sumsq = align_up(slots * ftile * 8u, 256u), p_cnt = align_up(p_sumsq + slots * ftile * 8u, 256u),
                 p_end = p_cnt + slots * ftile * 4u;
    if (!ensure_buf(P->partial, p_end, "hipMalloc label partials")) return 0;
    char *pb = (char *) P->partial.p;

    if (!P->events_ready) {
        CSAMD_HIP(hipEventCreate(&P->ev0), "hipEventCreate");
        CSAMD_HIP(hipEventCreate(&P->ev1), "hipEventCreate");
        P->events_ready = 1;
    }

    // 6. main passes over feature tiles
    reduce_args a;
    a.sorted_rows = ids_out;
    a.gstart = gstart;
    a.bounds = bounds;
    a.fpairs = (const uint2 *) P->fpairs.p;
    a.psum = (double *) (pb + p_sum);
    a.psumsq = (double *) (pb + p_sumsq);
    a.pcnt = (unsigned int *) (pb + p_cnt);
    a.G = G;
    a.cols = cols;
    a.nwords = nwords;
    a.ftile = ftile;
    P->last_valid = 0;
    bool first = true;
    for (unsigned int f0 = 0u; f0 < F || (F == 0u && first); f0 += ftile) {
        a.f0 = f0;
        a.f_count = F == 0u ? 0u : (F - f0 < ftile ? F - f0 : ftile);
        if (K != 0u && a.f_count != 0u) {
            if (first) CSAMD_HIP(hipEventRecord(P->ev0, s), "hipEventRecord");
            int ok = 0;
            switch (value_type) {
                case CSAMD_VALUE_F16: ok = launch_rows<CSAMD_VALUE_F16>(L, W, threads, lds, s, a, aligned, ranges); break;
                case CSAMD_VALUE_BF16: ok = launch_rows<CSAMD_VALUE_BF16>(L, W, threads, lds, s, a, aligned, ranges); break;
                case CSAMD_VALUE_F32: ok = launch_rows<CSAMD_VALUE_F32>(L, W, threads, lds, s, a, aligned, ranges); break;
                case CSAMD_VALUE_FP8_E4M3: ok = launch_rows<CSAMD_VALUE_FP8_E4M3>(L, W, threads, lds, s, a, aligned, ranges); break;
            }
            if (!ok) return 0;
            if (first) {
                CSAMD_HIP(hipEventRecord(P->ev1, s), "hipEventRecord");
                P->last_valid = 1;
            }
        }
        if (a.f_count != 0u) {
            label_combine_kernel<<<grid_for((unsigned long long) G * a.f_count, 256u, 8192u), 256, 0, s>>>(
                a.psum, a.psumsq, a.pcnt, gstart, bounds, W, G, ftile, f0, a.f_count, F, q->feature_scale, fids, tail, out_sum,
                out_sumsq, out_cnt);
            CSAMD_HIP(hipGetLastError(), "label_combine_kernel");
        }
        first = false;
        if (F == 0u) break;
    }
    label_moments_kernel<<<grid_for(GF == 0u ? (unsigned long long) G : GF, 256u, 8192u), 256, 0, s>>>(
        out_sum, out_sumsq, gstart, G, F, q->compute_moments ? out_mean : nullptr, out_var, out_grows);
    CSAMD_HIP(hipGetLastError(), "label_moments_kernel");

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
        dev_buf *bufs[] = {&P->fpairs, &P->fcounts, &P->fids, &P->keys, &P->rows, &P->sort_tmp, &P->gstart,
                           &P->work, &P->bounds, &P->partial, &P->out, &P->tail};
        for (dev_buf *b : bufs)
            if (b->p != nullptr) (void) hipFree(b->p);
        if (P->events_ready) {
            (void) hipEventDestroy(P->ev0);
            (void) hipEventDestroy(P->ev1);
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
