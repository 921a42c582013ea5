// Reference-ABI masked grouped row reductions and masked Blocked-ELL -> CSR
// compaction.  Semantics follow src/runtime/mask_groups.cu (CellShard):
//   per row r: masked-out row -> zeros, keep = 0 (emit_empty_row, :81-105);
//   otherwise over stored (c, v) with c < cols, v != 0, feature_keep[c]:
//   sum += v, detected += 1, max = max(0, v...), group[b] += v for each set
//   bit b < G of feature_group_masks[c]; keep = sum >= min_total &&
//   detected >= min_detected && fraction <= max_fraction (:171-193).
// Work mapping is gfx950-native: one wave64 per row, lane-strided coalesced
// streams, per-wave group bins in LDS, wave64 butterfly reductions.
#include "common.hh"
#include "group_reduce.hh"
#include "stream.hh"

#include <cstdlib>

namespace csamd {
namespace {

constexpr unsigned int MG_THREADS = 256u;
constexpr unsigned int MG_WAVES = MG_THREADS / 64u;

struct reduce_outputs {
    float *row_totals;
    float *max_values;
    unsigned int *detected_features;
    unsigned char *row_keep;
    float *group_counts;
    float *group_percentages;
};

__device__ __forceinline__ bool row_participates(const csamd_row_feature_mask_view &m, unsigned int row) {
    return m.row_keep == nullptr || m.row_keep[row] != 0u;
}
__device__ __forceinline__ bool feature_participates(const csamd_row_feature_mask_view &m, unsigned int col) {
    return m.feature_keep == nullptr || m.feature_keep[col] != 0u;
}

__device__ __forceinline__ void lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// Per-element accumulation shared by all three layouts.
// Group sums live in each lane's registers (one per group, combined by
// wave reductions at the row's end): an entry adds its value to the groups
// of its feature mask without LDS atomics, which serialise when many lanes
// add to one group.
struct row_acc {
    float sum;
    float vmax;
    unsigned int detected;
    float g[CSAMD_MAX_MASK_GROUPS];
};

__device__ __forceinline__ row_acc row_acc_zero() {
    row_acc a;
    a.sum = 0.0f;
    a.vmax = 0.0f;
    a.detected = 0u;
#pragma unroll
    for (unsigned int k = 0; k < CSAMD_MAX_MASK_GROUPS; ++k) a.g[k] = 0.0f;
    return a;
}

__device__ __forceinline__ void add_groups(row_acc &a, uint32_t m, float v) {
#pragma unroll
    for (unsigned int k = 0; k < CSAMD_MAX_MASK_GROUPS; ++k) a.g[k] += ((m >> k) & 1u) != 0u ? v : 0.0f;
}

__device__ __forceinline__ uint32_t group_bits(unsigned int group_count) {
    return group_count >= 32u ? 0xffffffffu : (1u << group_count) - 1u;
}

// Feature keep flags: a bitmap staged in LDS (kl) when the call provides
// one, else the mask bytes.  col must be < cols.
__device__ __forceinline__ bool keep_of(const unsigned int *kl, const csamd_row_feature_mask_view &masks, unsigned int col) {
    if (kl != nullptr) return ((kl[col >> 5] >> (col & 31u)) & 1u) != 0u;
    return feature_participates(masks, col);
}

// An entry whose column test (col < cols, feature kept) is known; gbits
// selects the call's groups (0: no group masks).
__device__ __forceinline__ void accumulate_kept(row_acc &a, bool kept, unsigned int col, float v, const uint32_t *__restrict__ fgm,
                                                uint32_t gbits) {
    if (kept && v != 0.0f) {
        a.sum += v;
        if (gbits != 0u) add_groups(a, fgm[col] & gbits, v);
        a.vmax = fmaxf(a.vmax, v);
        a.detected += 1u;
    }
}

__device__ __forceinline__ void accumulate(row_acc &a, unsigned int col, float v, unsigned int cols,
                                           const csamd_row_feature_mask_view &masks, const uint32_t *__restrict__ fgm,
                                           unsigned int group_count, float *wave_groups) {
    if (col < cols && v != 0.0f && feature_participates(masks, col)) {
        a.sum += v;
        if (fgm != nullptr && group_count != 0u) add_groups(a, fgm[col] & group_bits(group_count), v);
        a.vmax = fmaxf(a.vmax, v);
        a.detected += 1u;
    }
}

__device__ __forceinline__ void emit_empty_row(unsigned int row, unsigned int group_count, const reduce_outputs &o) {
    const unsigned int lane = lane_id();
    if (lane == 0u) {
        o.row_totals[row] = 0.0f;
        o.detected_features[row] = 0u;
        o.max_values[row] = 0.0f;
        o.row_keep[row] = 0u;
    }
    if (o.group_counts != nullptr && o.group_percentages != nullptr) {
        for (unsigned int g = lane; g < group_count; g += 64u) {
            const unsigned long long off = (unsigned long long) row * group_count + g;
            o.group_counts[off] = 0.0f;
            o.group_percentages[off] = 0.0f;
        }
    }
}

constexpr unsigned int MG_TSTRIDE = CSAMD_MAX_MASK_GROUPS + 1u;  // padded row of the group transpose (no bank conflicts)

__device__ __forceinline__ void finish_row(unsigned int row, row_acc a, unsigned int group_count,
                                           const csamd_sparse_group_filter_params &filter, const float *wave_groups,
                                           const reduce_outputs &o, float *tbuf) {
    const unsigned int lane = lane_id();
    const float row_sum = wave_sum_f32(a.sum);
    const float vmax = wave_max_f32(a.vmax);
    const unsigned int detected = wave_sum_u32(a.detected);
    if (group_count != 0u) {
        // the lanes' group sums to the wave's bins through an LDS transpose:
        // lane l writes its row of partials, lane k < group_count adds up
        // column k in lane order
#pragma unroll
        for (unsigned int k = 0; k < CSAMD_MAX_MASK_GROUPS; ++k) tbuf[lane * MG_TSTRIDE + k] = a.g[k];
        lds_fence();
        if (lane < group_count) {
            float t = 0.0f;
            for (unsigned int l = 0; l < 64u; ++l) t += tbuf[l * MG_TSTRIDE + lane];
            ((float *) wave_groups)[lane] = t;
        }
    }
    lds_fence();
    if (lane == 0u) {
        o.row_totals[row] = row_sum;
        o.max_values[row] = vmax;
        o.detected_features[row] = detected;
        const float gsel = filter.fraction_group_index < group_count ? wave_groups[filter.fraction_group_index] : 0.0f;
        const float frac = row_sum > 0.0f ? gsel / row_sum : 0.0f;
        o.row_keep[row] = (unsigned char) (row_sum >= filter.min_total && detected >= filter.min_detected_features &&
                                           frac <= filter.max_group_fraction);
    }
    if (o.group_counts != nullptr && o.group_percentages != nullptr && lane < group_count) {
        const float gs = wave_groups[lane];
        const unsigned long long off = (unsigned long long) row * group_count + lane;
        o.group_counts[off] = gs;
        o.group_percentages[off] = row_sum > 0.0f ? 100.0f * gs / row_sum : 0.0f;
    }
    lds_fence();
}

// Units of 4 consecutive values: one 4 / 8 / 16-byte load, decoded after
// all of a pass's loads are issued.
template<int VT> struct mg_unit;
template<> struct mg_unit<CSAMD_VALUE_F16> {
    using raw = uint2;
    static __device__ __forceinline__ float get(raw r, unsigned int k) {
        const unsigned int w = k < 2u ? r.x : r.y;
        return f16_bits_to_float((uint16_t) (w >> (16u * (k & 1u))));
    }
};
template<> struct mg_unit<CSAMD_VALUE_BF16> {
    using raw = uint2;
    static __device__ __forceinline__ float get(raw r, unsigned int k) {
        const unsigned int w = k < 2u ? r.x : r.y;
        return bf16_bits_to_float((uint16_t) (w >> (16u * (k & 1u))));
    }
};
template<> struct mg_unit<CSAMD_VALUE_F32> {
    using raw = float4;
    static __device__ __forceinline__ float get(raw r, unsigned int k) { return k == 0u ? r.x : k == 1u ? r.y : k == 2u ? r.z : r.w; }
};
template<> struct mg_unit<CSAMD_VALUE_FP8_E4M3> {
    using raw = unsigned int;
    static __device__ __forceinline__ float get(raw r, unsigned int k) { return fp8_e4m3_decode_hw(r, (int) k); }
};

constexpr unsigned int MG_UNROLL = 4u;  // units in flight per lane

// Whether entry arrays idx (u32) and val (VT) can be read in units: both
// 4-byte / value aligned and the same number of entries past a unit
// boundary (then unit-aligned entries are aligned in both).
template<int VT>
__device__ __forceinline__ bool mg_units_ok(const void *idx, const void *val, unsigned int *mis) {
    constexpr unsigned int VB = sizeof(typename value_traits<VT>::storage);
    const uintptr_t ia = (uintptr_t) idx, va = (uintptr_t) val;
    *mis = (unsigned int) ((ia >> 2) & 3u);
    return (ia & 3u) == 0u && (va % VB) == 0u && (unsigned int) ((va / VB) & 3u) == *mis;
}

// Entries [b, e) of paired arrays (column, value), accumulated: the
// unit-aligned middle by 16-byte index loads and 4-value loads, MG_UNROLL
// units in flight per lane (loads clamped to the last unit, never under a
// branch); the up to 3 + 3 entries outside whole units by lanes 0-5.
template<int VT>
__device__ __forceinline__ void mg_stream(row_acc &a, const unsigned int *__restrict__ idx,
                                          const typename value_traits<VT>::storage *__restrict__ val, unsigned long long b,
                                          unsigned long long e, bool units_ok, unsigned int mis, unsigned int cols,
                                          const csamd_row_feature_mask_view &masks, const uint32_t *__restrict__ fgm,
                                          unsigned int group_count, float *wg, const unsigned int *kl) {
    const unsigned int lane = lane_id();
    if (!units_ok) {
        for (unsigned long long i = b + lane; i < e; i += 64u)
            accumulate(a, idx[i], value_traits<VT>::decode(val[i]), cols, masks, fgm, group_count, wg);
        return;
    }
    const unsigned long long ua = b + ((4u - (unsigned int) ((b + mis) & 3u)) & 3u);
    const unsigned long long a0 = ua < e ? ua : e;
    const unsigned long long units = (e - a0) >> 2;
    const unsigned long long ea = a0 + 4u * units;
    if (lane < 6u) {
        const unsigned long long i = lane < 3u ? b + lane : ea + lane - 3u;
        if (lane < 3u ? i < a0 : i < e) accumulate(a, idx[i], value_traits<VT>::decode(val[i]), cols, masks, fgm, group_count, wg);
    }
    using raw = typename mg_unit<VT>::raw;
    const uint32_t gbits = fgm != nullptr && group_count != 0u ? group_bits(group_count) : 0u;
    const uint4 *U = (const uint4 *) (idx + a0);
    const raw *V = (const raw *) (val + a0);
    for (unsigned long long j0 = lane; j0 < units; j0 += 64u * MG_UNROLL) {
        uint4 q[MG_UNROLL];
        raw r[MG_UNROLL];
#pragma unroll
        for (unsigned int u = 0; u < MG_UNROLL; ++u) {
            const unsigned long long j = j0 + 64u * u < units ? j0 + 64u * u : units - 1u;
            q[u] = U[j];
            r[u] = V[j];
        }
        // every keep flag of the pass read before any is used
        bool kp[MG_UNROLL][4];
#pragma unroll
        for (unsigned int u = 0; u < MG_UNROLL; ++u) {
            const unsigned int c[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
            for (unsigned int k = 0; k < 4u; ++k) kp[u][k] = c[k] < cols && keep_of(kl, masks, c[k] < cols ? c[k] : 0u);
        }
#pragma unroll
        for (unsigned int u = 0; u < MG_UNROLL; ++u) {
            if (j0 + 64u * u >= units) continue;
            const unsigned int c[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
            for (unsigned int k = 0; k < 4u; ++k) accumulate_kept(a, kp[u][k], c[k], mg_unit<VT>::get(r[u], k), fgm, gbits);
        }
    }
}

template<int VT>
__global__ void __launch_bounds__(MG_THREADS) group_reduce_compressed_kernel(csamd_compressed_view src, const uint32_t *__restrict__ fgm,
                                                                             csamd_row_feature_mask_view masks, unsigned int group_count,
                                                                             csamd_sparse_group_filter_params filter, reduce_outputs o, const unsigned int *__restrict__ keep_bits,
    unsigned int keep_words) {
    using S = typename value_traits<VT>::storage;
    __shared__ float groups[MG_WAVES][CSAMD_MAX_MASK_GROUPS];
    __shared__ float transpose[MG_WAVES][64u * MG_TSTRIDE];
    extern __shared__ unsigned int keep_lds[];
    const unsigned int *kl = nullptr;
    if (keep_bits != nullptr) {  // the packed feature keep flags, staged once per workgroup
        for (unsigned int w = threadIdx.x; w < keep_words; w += MG_THREADS) keep_lds[w] = keep_bits[w];
        __syncthreads();
        kl = keep_lds;
    }
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int lane = lane_id();
    float *wg = groups[wave];
    float *tb = transpose[wave];
    const S *__restrict__ val = (const S *) src.val;
    unsigned int mis = 0u;
    const bool units_ok = mg_units_ok<VT>(src.minorIdx, src.val, &mis);
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (lane < CSAMD_MAX_MASK_GROUPS) wg[lane] = 0.0f;
        lds_fence();
        if (!row_participates(masks, row)) {
            emit_empty_row(row, group_count, o);
            continue;
        }
        const unsigned int begin = src.majorPtr[row];
        const unsigned int end = src.majorPtr[row + 1u];
        row_acc a = row_acc_zero();
        mg_stream<VT>(a, src.minorIdx, val, begin, end, units_ok, mis, src.cols, masks, fgm, group_count, wg, kl);
        finish_row(row, a, group_count, filter, wg, o, tb);
    }
}

__device__ __forceinline__ unsigned int find_slice_of_row(const csamd_sliced_ell_view &s, unsigned int row) {
    unsigned int slice;
    if (s.slice_rows == 32u) {
        slice = row >> 5;
    } else if (s.slice_rows != 0u) {
        slice = row / s.slice_rows;
    } else {
        // offsets[slice] <= row < offsets[slice+1]; binary search (the
        // reference walks linearly, same answer).
        unsigned int lo = 0u, hi = s.slice_count - 1u;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (s.slice_row_offsets[mid] <= row) lo = mid;
            else hi = mid - 1u;
        }
        slice = lo;
    }
    if (slice >= s.slice_count) slice = s.slice_count - 1u;
    return slice;
}

template<int VT>
__global__ void __launch_bounds__(MG_THREADS) group_reduce_sliced_kernel(csamd_sliced_ell_view src, const uint32_t *__restrict__ fgm,
                                                                         csamd_row_feature_mask_view masks, unsigned int group_count,
                                                                         csamd_sparse_group_filter_params filter, reduce_outputs o, const unsigned int *__restrict__ keep_bits,
    unsigned int keep_words) {
    using S = typename value_traits<VT>::storage;
    __shared__ float groups[MG_WAVES][CSAMD_MAX_MASK_GROUPS];
    __shared__ float transpose[MG_WAVES][64u * MG_TSTRIDE];
    extern __shared__ unsigned int keep_lds[];
    const unsigned int *kl = nullptr;
    if (keep_bits != nullptr) {  // the packed feature keep flags, staged once per workgroup
        for (unsigned int w = threadIdx.x; w < keep_words; w += MG_THREADS) keep_lds[w] = keep_bits[w];
        __syncthreads();
        kl = keep_lds;
    }
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int lane = lane_id();
    float *wg = groups[wave];
    float *tb = transpose[wave];
    const S *__restrict__ val = (const S *) src.val;
    unsigned int mis = 0u;
    const bool units_ok = mg_units_ok<VT>(src.col_idx, src.val, &mis);
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (lane < CSAMD_MAX_MASK_GROUPS) wg[lane] = 0.0f;
        lds_fence();
        if (!row_participates(masks, row)) {
            emit_empty_row(row, group_count, o);
            continue;
        }
        unsigned int width = 0u;
        unsigned long long slot_base = 0ull;
        if (src.slice_count != 0u) {
            const unsigned int slice = find_slice_of_row(src, row);
            const unsigned int row_begin = src.slice_row_offsets[slice];
            width = src.slice_widths[slice];
            slot_base = (unsigned long long) src.slice_slot_offsets[slice] + (unsigned long long) (row - row_begin) * width;
        }
        row_acc a = row_acc_zero();
        mg_stream<VT>(a, src.col_idx, val, slot_base, slot_base + width, units_ok, mis, src.cols, masks, fgm, group_count, wg, kl);
        finish_row(row, a, group_count, filter, wg, o, tb);
    }
}

template<int VT>
__global__ void __launch_bounds__(MG_THREADS) group_reduce_blocked_kernel(csamd_blocked_ell_view src, const uint32_t *__restrict__ fgm,
                                                                          csamd_row_feature_mask_view masks, unsigned int group_count,
                                                                          csamd_sparse_group_filter_params filter, reduce_outputs o, const unsigned int *__restrict__ keep_bits,
    unsigned int keep_words) {
    using S = typename value_traits<VT>::storage;
    __shared__ float groups[MG_WAVES][CSAMD_MAX_MASK_GROUPS];
    __shared__ float transpose[MG_WAVES][64u * MG_TSTRIDE];
    extern __shared__ unsigned int keep_lds[];
    const unsigned int *kl = nullptr;
    if (keep_bits != nullptr) {  // the packed feature keep flags, staged once per workgroup
        for (unsigned int w = threadIdx.x; w < keep_words; w += MG_THREADS) keep_lds[w] = keep_bits[w];
        __syncthreads();
        kl = keep_lds;
    }
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int lane = lane_id();
    float *wg = groups[wave];
    float *tb = transpose[wave];
    const S *__restrict__ val = (const S *) src.val;
    const unsigned int bs = src.block_size;
    const unsigned int width = bs != 0u ? src.ell_cols / bs : 0u;
    constexpr unsigned int VB = sizeof(S);
    const bool units_ok = bs != 0u && bs % 4u == 0u && src.ell_cols % 4u == 0u && ((uintptr_t) src.val % (4u * VB)) == 0u;
    const uint32_t gbits = fgm != nullptr && group_count != 0u ? group_bits(group_count) : 0u;
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (lane < CSAMD_MAX_MASK_GROUPS) wg[lane] = 0.0f;
        lds_fence();
        if (!row_participates(masks, row)) {
            emit_empty_row(row, group_count, o);
            continue;
        }
        const unsigned int rb = bs != 0u ? row / bs : 0u;
        const unsigned long long row_base = (unsigned long long) row * src.ell_cols;
        row_acc a = row_acc_zero();
        if (units_ok) {
            // 4 values of one block per unit (bs % 4 == 0), MG_UNROLL units in flight per lane
            using raw = typename mg_unit<VT>::raw;
            const raw *V = (const raw *) (val + row_base);
            const unsigned int units = src.ell_cols >> 2;
            const unsigned int *bcr = src.blockColIdx + (unsigned long long) rb * width;
            for (unsigned int j0 = lane; j0 < units; j0 += 64u * MG_UNROLL) {
                raw r[MG_UNROLL];
                unsigned int bc[MG_UNROLL];
#pragma unroll
                for (unsigned int u = 0; u < MG_UNROLL; ++u) {
                    const unsigned int j = j0 + 64u * u < units ? j0 + 64u * u : units - 1u;
                    r[u] = V[j];
                    bc[u] = bcr[(4u * j) / bs];
                }
                bool kp[MG_UNROLL][4];
                unsigned int c0[MG_UNROLL];
#pragma unroll
                for (unsigned int u = 0; u < MG_UNROLL; ++u) {
                    const unsigned int j = j0 + 64u * u < units ? j0 + 64u * u : units - 1u;
                    c0[u] = bc[u] != CSAMD_INVALID_COL ? bc[u] * bs + (4u * j) % bs : src.cols;
#pragma unroll
                    for (unsigned int k = 0; k < 4u; ++k)
                        kp[u][k] = c0[u] + k < src.cols && keep_of(kl, masks, c0[u] + k < src.cols ? c0[u] + k : 0u);
                }
#pragma unroll
                for (unsigned int u = 0; u < MG_UNROLL; ++u) {
                    if (j0 + 64u * u >= units) continue;
#pragma unroll
                    for (unsigned int k = 0; k < 4u; ++k) accumulate_kept(a, kp[u][k], c0[u] + k, mg_unit<VT>::get(r[u], k), fgm, gbits);
                }
            }
            finish_row(row, a, group_count, filter, wg, o, tb);
            continue;
        }
        for (unsigned int e = lane; e < src.ell_cols; e += 64u) {
            const unsigned int slot = bs != 0u ? e / bs : 0u;
            const unsigned int lib = bs != 0u ? e % bs : 0u;
            const unsigned int bc = width != 0u ? src.blockColIdx[(unsigned long long) rb * width + slot] : CSAMD_INVALID_COL;
            const unsigned int gene = bc != CSAMD_INVALID_COL ? bc * bs + lib : src.cols;
            accumulate(a, gene, value_traits<VT>::decode(val[row_base + e]), src.cols, masks, fgm, group_count, wg);
        }
        finish_row(row, a, group_count, filter, wg, o, tb);
    }
}

unsigned int reduce_grid(unsigned int rows) {
    unsigned int blocks = (rows + MG_WAVES - 1u) / MG_WAVES;
    const unsigned int cap = (unsigned int) device_cu_count() * 16u;
    if (blocks < 1u) blocks = 1u;
    if (blocks > cap) blocks = cap;
    return blocks;
}

void bind_result(csamd_sparse_group_reduce_workspace *ws, unsigned int rows, unsigned int group_count,
                 csamd_sparse_group_reduce_result *out) {
    if (out == nullptr) return;
    out->rows = rows;
    out->group_count = group_count;
    out->row_totals = ws->row_totals;
    out->detected_features = ws->detected_features;
    out->max_values = ws->max_values;
    out->row_keep = ws->row_keep;
    out->group_counts = ws->group_counts;
    out->group_percentages = ws->group_percentages;
}

reduce_outputs outputs_of(const csamd_sparse_group_reduce_workspace *ws, unsigned int group_count) {
    reduce_outputs o;
    o.row_totals = ws->row_totals;
    o.max_values = ws->max_values;
    o.detected_features = ws->detected_features;
    o.row_keep = ws->row_keep;
    o.group_counts = group_count != 0u ? ws->group_counts : nullptr;
    o.group_percentages = group_count != 0u ? ws->group_percentages : nullptr;
    return o;
}

unsigned int sat_u32(unsigned long long v) { return v > 0xffffffffull ? 0xffffffffu : (unsigned int) v; }

struct reduce_prelude {
    unsigned int group_count;
    csamd_row_feature_mask_view masks;
    const uint32_t *fgm;
    const unsigned int *keep_bits;  // packed feature_keep (null: none, or too many columns for LDS)
    unsigned int keep_words;
};

constexpr unsigned int MG_KEEP_WORDS_MAX = 6144u;  // 24 KB of LDS (with the 34 KB group transpose): up to 196,608 columns

// feature_keep bytes -> bits (bit c of word c / 32 set when column c is kept).
__global__ void mg_pack_keep_kernel(const unsigned char *__restrict__ keep, unsigned int cols, unsigned int *__restrict__ bits,
                                    unsigned int words) {
    for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gridDim.x * blockDim.x) {
        unsigned int b = 0u;
        const unsigned int c0 = w * 32u;
        for (unsigned int k = 0; k < 32u && c0 + k < cols; ++k) b |= (keep[c0 + k] != 0u ? 1u : 0u) << k;
        bits[w] = b;
    }
}

// Grow-only workspace scratch (an earlier call's kernels may still read the old block).
int ensure_scratch(csamd_sparse_group_reduce_workspace *ws, size_t need) {
    if (need <= ws->scratch_bytes) return 1;
    CSAMD_HIP(hipStreamSynchronize((hipStream_t) ws->stream), "sparse group scratch sync");
    if (ws->scratch_block != nullptr) (void) hipFree(ws->scratch_block);
    ws->scratch_block = nullptr;
    ws->scratch_bytes = 0u;
    CSAMD_HIP(hipMalloc(&ws->scratch_block, align_up(need, 4096u)), "hipMalloc sparse group scratch");
    ws->scratch_bytes = align_up(need, 4096u);
    return 1;
}

int prepare_reduce(const void *src, csamd_sparse_group_reduce_workspace *ws, const csamd_group_mask_config_view *groups,
                   const csamd_row_feature_mask_view *masks, const csamd_sparse_group_filter_params *filter,
                   unsigned int rows, unsigned int cols, unsigned long long values, unsigned int value_type,
                   reduce_prelude *p, bool pack_keep = true) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr && filter != nullptr, "compute_sparse_group_reduce: null argument");
    CSAMD_REQUIRE(value_type_decodable(value_type), "compute_sparse_group_reduce: unsupported value type %u", value_type);
    p->group_count = groups != nullptr ? groups->group_count : 0u;
    CSAMD_REQUIRE(p->group_count <= CSAMD_MAX_MASK_GROUPS, "compute_sparse_group_reduce: group_count %u > %u", p->group_count,
                  CSAMD_MAX_MASK_GROUPS);
    if (!csamd_group_reduce_reserve(ws, rows, cols, sat_u32(values), p->group_count)) return 0;
    if (masks != nullptr) p->masks = *masks;
    else std::memset(&p->masks, 0, sizeof(p->masks));
    p->fgm = groups != nullptr && groups->feature_group_masks != nullptr ? groups->feature_group_masks : ws->feature_group_masks;
    p->keep_bits = nullptr;
    p->keep_words = 0u;
    const unsigned int words = (cols + 31u) / 32u;
    if (pack_keep && p->masks.feature_keep != nullptr && cols != 0u && words <= MG_KEEP_WORDS_MAX && rows != 0u) {
        if (!ensure_scratch(ws, (size_t) words * 4u)) return 0;
        hipStream_t s = (hipStream_t) ws->stream;
        mg_pack_keep_kernel<<<grid_for(words, 256u, 1024u), 256, 0, s>>>(p->masks.feature_keep, cols, (unsigned int *) ws->scratch_block,
                                                                        words);
        CSAMD_HIP(hipGetLastError(), "mg_pack_keep_kernel");
        p->keep_bits = (const unsigned int *) ws->scratch_block;
        p->keep_words = words;
    }
    return 1;
}

gr_source stream_source(const csamd_compressed_view &v, unsigned int value_type) {
    gr_source g;
    std::memset(&g, 0, sizeof(g));
    g.rows = v.rows;
    g.cols = v.cols;
    g.ptr = v.majorPtr;
    g.idx = v.minorIdx;
    g.val = v.val;
    g.value_type = value_type;
    return g;
}

gr_source stream_source(const csamd_sliced_ell_view &v, unsigned int value_type) {
    gr_source g;
    std::memset(&g, 0, sizeof(g));
    g.rows = v.rows;
    g.cols = v.cols;
    g.slice_count = v.slice_count;
    g.slice_rows = v.slice_rows;
    g.slice_offsets = v.slice_row_offsets;
    g.slice_widths = v.slice_widths;
    g.slot_offsets = v.slice_slot_offsets;
    g.idx = v.col_idx;
    g.val = v.val;
    g.value_type = value_type;
    return g;
}

gr_outputs stream_outputs(const reduce_outputs &o) {
    return gr_outputs{o.row_totals, o.max_values, o.detected_features, o.row_keep, o.group_counts, o.group_percentages};
}

// The streaming path over compressed rows / Sliced-ELL (group_reduce.hip).
int run_stream(const gr_source &gs, csamd_sparse_group_reduce_workspace *ws, const reduce_prelude &p,
               const csamd_sparse_group_filter_params &filter) {
    if (!ensure_scratch(ws, gr_scratch_bytes(gs.cols))) return 0;
    return gr_reduce_rows(gs, p.group_count != 0u ? p.fgm : nullptr, p.group_count, p.masks.row_keep, p.masks.feature_keep, filter,
                          stream_outputs(outputs_of(ws, p.group_count)), ws->scratch_block, ws->scratch_bytes, (hipStream_t) ws->stream);
}

template<template<int> class Launch, typename View>
int dispatch_value_type(unsigned int vt, const View &v, unsigned int grid, hipStream_t stream, const reduce_prelude &p,
                        const csamd_sparse_group_filter_params &f, const reduce_outputs &o) {
    switch (vt) {
        case CSAMD_VALUE_F16: Launch<CSAMD_VALUE_F16>::run(v, grid, stream, p, f, o); break;
        case CSAMD_VALUE_BF16: Launch<CSAMD_VALUE_BF16>::run(v, grid, stream, p, f, o); break;
        case CSAMD_VALUE_F32: Launch<CSAMD_VALUE_F32>::run(v, grid, stream, p, f, o); break;
        case CSAMD_VALUE_FP8_E4M3: Launch<CSAMD_VALUE_FP8_E4M3>::run(v, grid, stream, p, f, o); break;
        default: set_error("unsupported value type %u", vt); return 0;
    }
    return hip_ok(hipGetLastError(), "group reduce kernel launch");
}

template<int VT> struct launch_compressed {
    static void run(const csamd_compressed_view &v, unsigned int grid, hipStream_t s, const reduce_prelude &p,
                    const csamd_sparse_group_filter_params &f, const reduce_outputs &o) {
        group_reduce_compressed_kernel<VT><<<grid, MG_THREADS, p.keep_words * 4u, s>>>(v, p.fgm, p.masks, p.group_count, f, o, p.keep_bits, p.keep_words);
    }
};
template<int VT> struct launch_sliced {
    static void run(const csamd_sliced_ell_view &v, unsigned int grid, hipStream_t s, const reduce_prelude &p,
                    const csamd_sparse_group_filter_params &f, const reduce_outputs &o) {
        group_reduce_sliced_kernel<VT><<<grid, MG_THREADS, p.keep_words * 4u, s>>>(v, p.fgm, p.masks, p.group_count, f, o, p.keep_bits, p.keep_words);
    }
};
template<int VT> struct launch_blocked {
    static void run(const csamd_blocked_ell_view &v, unsigned int grid, hipStream_t s, const reduce_prelude &p,
                    const csamd_sparse_group_filter_params &f, const reduce_outputs &o) {
        group_reduce_blocked_kernel<VT><<<grid, MG_THREADS, p.keep_words * 4u, s>>>(v, p.fgm, p.masks, p.group_count, f, o, p.keep_bits, p.keep_words);
    }
};

// ---------------------------------------------------------------------------
// Masked Blocked-ELL -> CSR (manual_reoptimize_masked_sparse, mask_groups.cu:960-1034).
// ---------------------------------------------------------------------------
struct reopt_private {
    void *compressed_block;
    size_t compressed_bytes;
    void *scratch_block;
    size_t scratch_bytes;
    void *gather_block;  // per-row regions of surviving indices | values (one-read path)
    size_t gather_bytes;
};

__device__ __forceinline__ bool reopt_keep(const csamd_blocked_ell_view &src, const csamd_row_feature_mask_view &masks,
                                           unsigned int rb, unsigned int width, unsigned int e, unsigned int row,
                                           unsigned int output_cols, unsigned int *out_col, uint16_t *bits) {
    const unsigned int bs = src.block_size;
    const unsigned int slot = bs != 0u ? e / bs : 0u;
    const unsigned int lib = bs != 0u ? e % bs : 0u;
    const unsigned int bc = width != 0u ? src.blockColIdx[(unsigned long long) rb * width + slot] : CSAMD_INVALID_COL;
    const unsigned int col = bc != CSAMD_INVALID_COL ? bc * bs + lib : src.cols;
    const uint16_t h = ((const uint16_t *) src.val)[(unsigned long long) row * src.ell_cols + e];
    const float v = f16_bits_to_float(h);
    if (col < src.cols && v != 0.0f && feature_participates(masks, col)) {
        const unsigned int oc = masks.feature_remap != nullptr ? masks.feature_remap[col] : col;
        if (oc < output_cols) {
            *out_col = oc;
            *bits = h;
            return true;
        }
    }
    return false;
}

// Output column of every source column (feature kept, remapped, below
// output_cols) or ~0u, so that a chunk of entries needs one gather each.
__global__ void reopt_col_map_kernel(csamd_row_feature_mask_view masks, unsigned int cols, unsigned int output_cols,
                                     unsigned int *__restrict__ col_map) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
        unsigned int o = ~0u;
        if (feature_participates(masks, c)) {
            o = masks.feature_remap != nullptr ? masks.feature_remap[c] : c;
            if (o >= output_cols) o = ~0u;
        }
        col_map[c] = o;
    }
}

// 16-byte chunks of 8 f16 values, when every chunk lies in one block: bit e
// of the mask is set when value e is nonzero at a column that maps to an
// output column (oc[e]).
__device__ __forceinline__ bool reopt_chunks_ok(const csamd_blocked_ell_view &src) {
    return src.block_size != 0u && src.block_size % 8u == 0u && src.ell_cols % 8u == 0u && ((uintptr_t) src.val & 15u) == 0u;
}

__device__ __forceinline__ unsigned int reopt_chunk(const csamd_blocked_ell_view &src, const unsigned int *__restrict__ col_map,
                                                    unsigned int width, unsigned int row, unsigned int c, uint4 *q,
                                                    unsigned int (&oc)[8]) {
    const unsigned int bs = src.block_size;
    const unsigned int t0 = c * 8u, slot = t0 / bs;
    const unsigned int bc = src.blockColIdx[(unsigned long long) (row / bs) * width + slot];
    *q = ((const uint4 *) ((const uint16_t *) src.val + (unsigned long long) row * src.ell_cols))[c];
    const unsigned int col0 = bc != CSAMD_INVALID_COL ? bc * bs + (t0 - slot * bs) : src.cols;
#pragma unroll
    for (unsigned int e = 0; e < 8u; ++e) oc[e] = col0 + e < src.cols ? col_map[col0 + e] : ~0u;  // gathers issued together
    const unsigned int w[4] = {q->x, q->y, q->z, q->w};
    unsigned int m = 0u;
#pragma unroll
    for (unsigned int e = 0; e < 8u; ++e)
        if (oc[e] != ~0u && ((w[e >> 1] >> (16u * (e & 1u))) & 0x7fffu) != 0u) m |= 1u << e;
    return m;
}

// block_keep[b] bit j = column b * bs + j is below cols and maps to an
// output column (bs <= 32).
__global__ void reopt_block_keep_kernel(unsigned int cols, unsigned int bs, const unsigned int *__restrict__ col_map,
                                        unsigned int *__restrict__ block_keep) {
    const unsigned int blocks = (unsigned int) (((unsigned long long) cols + bs - 1u) / bs);
    for (unsigned int b = blockIdx.x * blockDim.x + threadIdx.x; b < blocks; b += gridDim.x * blockDim.x) {
        unsigned int w = 0u;
        for (unsigned int j = 0; j < bs; ++j) {
            const unsigned long long c = (unsigned long long) b * bs + j;
            if (c < cols && col_map[c] != ~0u) w |= 1u << j;
        }
        block_keep[b] = w;
    }
}

// reopt_chunk with the chunk's kept columns taken from one word of its block
// column's keep bits: col_map is read only for surviving values.
__device__ __forceinline__ unsigned int reopt_chunk_keep(const csamd_blocked_ell_view &src, const unsigned int *__restrict__ col_map,
                                                         const unsigned int *__restrict__ block_keep, unsigned int keep_blocks,
                                                         unsigned int width, unsigned int row, unsigned int c, uint4 *q,
                                                         unsigned int (&oc)[8]) {
    const unsigned int bs = src.block_size;
    const unsigned int t0 = c * 8u, slot = t0 / bs, j0 = t0 - slot * bs;
    const unsigned int bc = src.blockColIdx[(unsigned long long) (row / bs) * width + slot];
    *q = ((const uint4 *) ((const uint16_t *) src.val + (unsigned long long) row * src.ell_cols))[c];
    const unsigned int w[4] = {q->x, q->y, q->z, q->w};
    unsigned int nz = 0u;
#pragma unroll
    for (unsigned int e = 0; e < 8u; ++e)
        if (((w[e >> 1] >> (16u * (e & 1u))) & 0x7fffu) != 0u) nz |= 1u << e;
    const unsigned int m = bc < keep_blocks ? nz & (block_keep[bc] >> j0) : 0u;  // CSAMD_INVALID_COL fails the test
#pragma unroll
    for (unsigned int e = 0; e < 8u; ++e) oc[e] = (m >> e) & 1u ? col_map[bc * bs + j0 + e] : ~0u;
    return m;
}

__global__ void __launch_bounds__(MG_THREADS) reopt_count_kernel(csamd_blocked_ell_view src, csamd_row_feature_mask_view masks,
                                                                 unsigned int output_rows, unsigned int output_cols,
                                                                 unsigned int *major_ptr, const unsigned int *__restrict__ col_map,
                                                                 const unsigned int *__restrict__ block_keep, unsigned int keep_blocks) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int bs = src.block_size;
    const unsigned int width = bs != 0u ? src.ell_cols / bs : 0u;
    const bool chunks = reopt_chunks_ok(src);
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (!row_participates(masks, row)) continue;
        const unsigned int out_row = masks.row_remap != nullptr ? masks.row_remap[row] : row;
        if (out_row >= output_rows) continue;
        const unsigned int rb = bs != 0u ? row / bs : 0u;
        unsigned int count = 0u;
        if (chunks) {
            const unsigned int nc = src.ell_cols / 8u;
            for (unsigned int c = lane_id(); c < nc; c += 64u) {
                uint4 q;
                unsigned int oc[8];
                count += (unsigned int) __popc(block_keep != nullptr
                                                   ? reopt_chunk_keep(src, col_map, block_keep, keep_blocks, width, row, c, &q, oc)
                                                   : reopt_chunk(src, col_map, width, row, c, &q, oc));
            }
            count = wave_sum_u32(count);
            if (lane_id() == 0u) major_ptr[out_row + 1u] = count;
            continue;
        }
        for (unsigned int e = lane_id(); e < src.ell_cols; e += 64u) {
            unsigned int oc;
            uint16_t bits;
            count += reopt_keep(src, masks, rb, width, e, row, output_cols, &oc, &bits) ? 1u : 0u;
        }
        count = wave_sum_u32(count);
        if (lane_id() == 0u) major_ptr[out_row + 1u] = count;
    }
}

// In-place exclusive scan of major_ptr[1..rows] into major_ptr[0..rows]
// (single block, loops with a carry).
__global__ void __launch_bounds__(1024) reopt_scan_kernel(unsigned int rows, unsigned int *major_ptr) {
    __shared__ unsigned int smem[16];
    __shared__ unsigned int carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (unsigned int start = 0; start < rows; start += blockDim.x) {
        const unsigned int i = start + threadIdx.x;
        const unsigned int v = i < rows ? major_ptr[i + 1u] : 0u;
        unsigned int total = 0u;
        const unsigned int ex = block_exclusive_scan_u32(v, smem, &total);
        const unsigned int c = carry;
        __syncthreads();
        if (i < rows) major_ptr[i] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) major_ptr[rows] = carry;
}

// One-read form of count + scatter for the 16-byte chunk path: a wave takes
// a row and writes its surviving entries, in position order, to the row's
// own region of a scratch array (row r at r * cap, at most cap entries; a
// row with more sets *overflow and the caller runs count + scatter
// instead), with the count in major_ptr[out_row + 1] and the source row in
// src_row[out_row].
__global__ void __launch_bounds__(MG_THREADS) reopt_gather_kernel(csamd_blocked_ell_view src, csamd_row_feature_mask_view masks,
                                                                  unsigned int output_rows, unsigned int cap,
                                                                  unsigned int *__restrict__ major_ptr, unsigned int *__restrict__ src_row,
                                                                  unsigned int *__restrict__ overflow, unsigned int *__restrict__ t_idx,
                                                                  uint16_t *__restrict__ t_val, const unsigned int *__restrict__ col_map,
                                                                  const unsigned int *__restrict__ block_keep, unsigned int keep_blocks) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int bs = src.block_size;
    const unsigned int width = src.ell_cols / bs;
    const unsigned int nc = src.ell_cols / 8u;
    CSAMD_GLOBAL unsigned int *gi = gmem(t_idx);
    CSAMD_GLOBAL uint16_t *gv = gmem(t_val);
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (!row_participates(masks, row)) continue;
        const unsigned int out_row = masks.row_remap != nullptr ? masks.row_remap[row] : row;
        if (out_row >= output_rows) continue;
        const unsigned long long base = (unsigned long long) row * cap;
        unsigned int w = 0u;
        for (unsigned int c0 = 0u; c0 < nc; c0 += 64u) {
            const unsigned int c = c0 + lane_id();
            uint4 q = make_uint4(0u, 0u, 0u, 0u);
            unsigned int oc[8];
            const unsigned int m = c >= nc ? 0u
                                   : block_keep != nullptr ? reopt_chunk_keep(src, col_map, block_keep, keep_blocks, width, row, c, &q, oc)
                                                           : reopt_chunk(src, col_map, width, row, c, &q, oc);
            const unsigned int cnt = (unsigned int) __popc(m);
            const unsigned int incl = wave_inclusive_scan_u32(cnt);
            unsigned int k = w + incl - cnt;
            w += __shfl(incl, 63);
            const unsigned int qw[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (unsigned int e = 0; e < 8u; ++e) {
                if ((m >> e) & 1u) {
                    if (k < cap) {
                        gi[base + k] = oc[e];
                        gv[base + k] = (uint16_t) (qw[e >> 1] >> (16u * (e & 1u)));
                    }
                    ++k;
                }
            }
        }
        if (lane_id() == 0u) {
            major_ptr[out_row + 1u] = w;
            src_row[out_row] = row;
            if (w > cap) *overflow = 1u;
        }
    }
}

// Streaming form of the gather (keep bits available, block size 8, 16 or
// 32): a wave streams its row 4 chunks per lane ahead through two register
// buffers (value chunk and block column together), the keep bits come from
// an LDS copy, and survivors are written with their source columns (the
// copy pass maps them through col_map), so nothing inside the loop waits on
// a dependent global load.
__global__ void __launch_bounds__(MG_THREADS) reopt_gather_stream_kernel(csamd_blocked_ell_view src, csamd_row_feature_mask_view masks,
                                                                         unsigned int output_rows, unsigned int cap, unsigned int cpb_log2,
                                                                         unsigned int *__restrict__ major_ptr,
                                                                         unsigned int *__restrict__ src_row, unsigned int *__restrict__ overflow,
                                                                         unsigned int *__restrict__ t_idx, uint16_t *__restrict__ t_val,
                                                                         const unsigned int *__restrict__ block_keep, unsigned int keep_blocks) {
    extern __shared__ unsigned int lkeep_w[];  // [keep_blocks]
    const __attribute__((address_space(3))) unsigned int *lkeep = (const __attribute__((address_space(3))) unsigned int *) lkeep_w;
    for (unsigned int k = threadIdx.x; k < keep_blocks; k += blockDim.x) lkeep_w[k] = gmem(block_keep)[k];
    __syncthreads();
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int bs = src.block_size;
    const unsigned int width = src.ell_cols / bs;
    const unsigned int nc = src.ell_cols / 8u;
    const unsigned int cmask = (1u << cpb_log2) - 1u;
    CSAMD_GLOBAL unsigned int *gi = gmem(t_idx);
    CSAMD_GLOBAL uint16_t *gv = gmem(t_val);
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (!row_participates(masks, row)) continue;
        const unsigned int out_row = masks.row_remap != nullptr ? masks.row_remap[row] : row;
        if (out_row >= output_rows) continue;
        const unsigned long long base = (unsigned long long) row * cap;
        unsigned int w = 0u;
        bell_stream_row<uint16_t, 4u>(src, width, cpb_log2, row, [&](unsigned int c, const u32x4 &q, unsigned int bc) {
            unsigned int nz = 0u;
#pragma unroll
            for (unsigned int e = 0; e < 8u; ++e)
                if ((chunk_val<uint16_t>(q, e) & 0x7fffu) != 0u) nz |= 1u << e;
            const unsigned int j0 = (c & cmask) * 8u;
            const unsigned int kb = bc < keep_blocks ? lkeep[bc] : 0u;  // CSAMD_INVALID_COL included
            const unsigned int m = c < nc ? nz & (kb >> j0) : 0u;
            const unsigned int cnt = (unsigned int) __popc(m);
            const unsigned int incl = wave_inclusive_scan_u32(cnt);
            unsigned int k = w + incl - cnt;
            w += (unsigned int) __builtin_amdgcn_readlane((int) incl, 63);
            const unsigned int col0 = bc * bs + j0;
#pragma unroll
            for (unsigned int e = 0; e < 8u; ++e) {
                if ((m >> e) & 1u) {
                    if (k < cap) {
                        gi[base + k] = col0 + e;
                        gv[base + k] = chunk_val<uint16_t>(q, e);
                    }
                    ++k;
                }
            }
        });
        if (lane_id() == 0u) {
            major_ptr[out_row + 1u] = w;
            src_row[out_row] = row;
            if (w > cap) *overflow = 1u;
        }
    }
}

// Copy pass after the gather: output row o's entries from its source row's
// region, one wave per row; col_map (non-null after the streaming gather)
// takes the stored source columns to output columns.
__global__ void __launch_bounds__(MG_THREADS) reopt_copy_kernel(unsigned int output_rows, unsigned int cap,
                                                                const unsigned int *__restrict__ major_ptr,
                                                                const unsigned int *__restrict__ src_row,
                                                                const unsigned int *__restrict__ t_idx, const uint16_t *__restrict__ t_val,
                                                                const unsigned int *__restrict__ col_map,
                                                                unsigned int *__restrict__ minor_idx, uint16_t *__restrict__ val) {
    for (unsigned int o = blockIdx.x * MG_WAVES + (threadIdx.x >> 6); o < output_rows; o += gridDim.x * MG_WAVES) {
        const unsigned int b = major_ptr[o], n = major_ptr[o + 1u] - b;
        if (n == 0u) continue;
        const unsigned long long base = (unsigned long long) src_row[o] * cap;
        for (unsigned int k = lane_id(); k < n; k += 64u) {
            const unsigned int c = gmem(t_idx)[base + k];
            gmem(minor_idx)[b + k] = col_map != nullptr ? gmem(col_map)[c] : c;  // streaming gather: source columns
            gmem(val)[b + k] = gmem(t_val)[base + k];
        }
    }
}

__global__ void __launch_bounds__(MG_THREADS) reopt_scatter_kernel(csamd_blocked_ell_view src, csamd_row_feature_mask_view masks,
                                                                   unsigned int output_rows, unsigned int output_cols,
                                                                   const unsigned int *__restrict__ major_ptr,
                                                                   unsigned int *__restrict__ row_write,
                                                                   unsigned int *__restrict__ minor_idx, uint16_t *__restrict__ val,
                                                                   const unsigned int *__restrict__ col_map,
                                                                   const unsigned int *__restrict__ block_keep, unsigned int keep_blocks) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int bs = src.block_size;
    const unsigned int width = bs != 0u ? src.ell_cols / bs : 0u;
    const bool chunks = reopt_chunks_ok(src);
    for (unsigned int row = blockIdx.x * MG_WAVES + wave; row < src.rows; row += gridDim.x * MG_WAVES) {
        if (!row_participates(masks, row)) continue;
        const unsigned int out_row = masks.row_remap != nullptr ? masks.row_remap[row] : row;
        if (out_row >= output_rows) continue;
        const unsigned int rb = bs != 0u ? row / bs : 0u;
        if (chunks) {
            // per 64 chunks: positions by a wave scan, one cursor claim per group
            const unsigned int nc = src.ell_cols / 8u;
            for (unsigned int c0 = 0u; c0 < nc; c0 += 64u) {
                const unsigned int c = c0 + lane_id();
                uint4 q = make_uint4(0u, 0u, 0u, 0u);
                unsigned int oc[8];
                const unsigned int m = c >= nc ? 0u
                                       : block_keep != nullptr
                                           ? reopt_chunk_keep(src, col_map, block_keep, keep_blocks, width, row, c, &q, oc)
                                           : reopt_chunk(src, col_map, width, row, c, &q, oc);
                const unsigned int cnt = (unsigned int) __popc(m);
                const unsigned int incl = wave_inclusive_scan_u32(cnt);
                const unsigned int total = __shfl(incl, 63);
                if (total == 0u) continue;
                unsigned int base = 0u;
                if (lane_id() == 0u) base = atomicAdd(row_write + out_row, total);
                base = __shfl(base, 0, 64);
                unsigned int w = major_ptr[out_row] + base + incl - cnt;
                const unsigned int qw[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (unsigned int e = 0; e < 8u; ++e) {
                    if ((m >> e) & 1u) {
                        minor_idx[w] = oc[e];
                        val[w] = (uint16_t) (qw[e >> 1] >> (16u * (e & 1u)));
                        ++w;
                    }
                }
            }
            continue;
        }
        const unsigned int ell_round = (src.ell_cols + 63u) & ~63u;
        for (unsigned int e0 = 0u; e0 < ell_round; e0 += 64u) {
            const unsigned int e = e0 + lane_id();
            unsigned int oc = 0u;
            uint16_t bits = 0u;
            const bool keep = e < src.ell_cols && reopt_keep(src, masks, rb, width, e, row, output_cols, &oc, &bits);
            const unsigned long long bal = __ballot(keep);
            if (bal == 0ull) continue;
            unsigned int base = 0u;
            if (lane_id() == 0u) base = atomicAdd(row_write + out_row, (unsigned int) __popcll(bal));
            base = __shfl(base, 0, 64);
            if (keep) {
                const unsigned int w = major_ptr[out_row] + base + (unsigned int) __popcll(bal & lanemask_lt());
                minor_idx[w] = oc;
                val[w] = bits;
            }
        }
    }
}

__global__ void fill_u32_kernel(unsigned int *p, unsigned long long n, unsigned int v) {
    for (unsigned long long i = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long) gridDim.x * blockDim.x)
        p[i] = v;
}

int fill_u32(unsigned int *p, unsigned long long n, unsigned int v, hipStream_t s) {
    if (n == 0ull) return 1;
    fill_u32_kernel<<<grid_for(n, 256u, 4096u), 256, 0, s>>>(p, n, v);
    return hip_ok(hipGetLastError(), "fill_u32_kernel");
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_group_reduce_init(csamd_sparse_group_reduce_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_group_reduce_clear(csamd_sparse_group_reduce_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    if (ws->owns_stream != 0 && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    if (ws->scratch_block != nullptr) (void) hipFree(ws->scratch_block);
    if (ws->feature_block != nullptr) (void) hipFree(ws->feature_block);
    if (ws->row_block != nullptr) (void) hipFree(ws->row_block);
    csamd_group_reduce_init(ws);
}

int csamd_group_reduce_setup(csamd_sparse_group_reduce_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "setup: null workspace");
    csamd_group_reduce_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice sparse group setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags sparse group");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    return 1;
}

int csamd_group_reduce_reserve(csamd_sparse_group_reduce_workspace *ws, unsigned int rows, unsigned int cols,
                               unsigned int values, unsigned int group_count) {
    CSAMD_REQUIRE(ws != nullptr, "reserve: null workspace");
    CSAMD_REQUIRE(group_count <= CSAMD_MAX_MASK_GROUPS, "reserve: group_count %u > %u", group_count, CSAMD_MAX_MASK_GROUPS);
    CSAMD_HIP(hipSetDevice(ws->device >= 0 ? ws->device : 0), "hipSetDevice sparse group reserve");
    if (rows > ws->rows_capacity || group_count > ws->group_capacity) {
        const unsigned int ar = rows > ws->rows_capacity ? rows : ws->rows_capacity;
        const unsigned int ag = group_count > ws->group_capacity ? group_count : ws->group_capacity;
        const size_t n = (size_t) ar, ng = (size_t) ar * ag;
        size_t off_tot = 0, off_max, off_det, off_keep, off_gc, off_gp, bytes;
        off_max = align_up(off_tot + n * sizeof(float), 16u);
        off_det = align_up(off_max + n * sizeof(float), 16u);
        off_keep = align_up(off_det + n * sizeof(unsigned int), 16u);
        off_gc = align_up(off_keep + n, 16u);
        off_gp = align_up(off_gc + ng * sizeof(float), 16u);
        bytes = off_gp + ng * sizeof(float);
        if (ws->row_block != nullptr) (void) hipFree(ws->row_block);
        ws->row_block = nullptr;
        ws->rows_capacity = 0u;
        ws->group_capacity = 0u;
        if (bytes != 0u) CSAMD_HIP(hipMalloc(&ws->row_block, bytes), "hipMalloc sparse group row block");
        char *b = (char *) ws->row_block;
        ws->row_totals = (float *) (b + off_tot);
        ws->max_values = (float *) (b + off_max);
        ws->detected_features = (unsigned int *) (b + off_det);
        ws->row_keep = (unsigned char *) (b + off_keep);
        ws->group_counts = (float *) (b + off_gc);
        ws->group_percentages = (float *) (b + off_gp);
        ws->rows_capacity = ar;
        ws->group_capacity = ag;
    }
    if (cols > ws->cols_capacity) {
        if (ws->feature_block != nullptr) (void) hipFree(ws->feature_block);
        ws->feature_block = nullptr;
        ws->feature_group_masks = nullptr;
        ws->cols_capacity = 0u;
        CSAMD_HIP(hipMalloc(&ws->feature_block, (size_t) cols * sizeof(uint32_t)), "hipMalloc sparse group feature block");
        ws->feature_group_masks = (uint32_t *) ws->feature_block;
        ws->cols_capacity = cols;
    }
    if (values > ws->values_capacity) ws->values_capacity = values;
    return 1;
}

int csamd_upload_feature_group_masks(csamd_sparse_group_reduce_workspace *ws, unsigned int cols, const uint32_t *host_masks) {
    CSAMD_REQUIRE(ws != nullptr, "upload_feature_group_masks: null workspace");
    if (!csamd_group_reduce_reserve(ws, ws->rows_capacity, cols, ws->values_capacity, ws->group_capacity)) return 0;
    if (cols == 0u) return 1;
    hipStream_t s = (hipStream_t) ws->stream;
    if (host_masks == nullptr) return fill_u32(ws->feature_group_masks, cols, 0u, s);
    CSAMD_HIP(hipMemcpyAsync(ws->feature_group_masks, host_masks, (size_t) cols * sizeof(uint32_t), hipMemcpyHostToDevice, s),
              "hipMemcpyAsync feature group masks");
    return 1;
}

int csamd_compute_sparse_group_reduce_compressed_typed(const csamd_compressed_view *src, unsigned int value_type,
                                                       csamd_sparse_group_reduce_workspace *ws,
                                                       const csamd_group_mask_config_view *groups,
                                                       const csamd_row_feature_mask_view *masks,
                                                       const csamd_sparse_group_filter_params *filter,
                                                       csamd_sparse_group_reduce_result *out) {
    reduce_prelude p;
    CSAMD_REQUIRE(src != nullptr, "compute_sparse_group_reduce_compressed: null src");
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW, "compute_sparse_group_reduce_compressed: axis must be row");
    const gr_source gs = stream_source(*src, value_type);
    const bool stream = gr_eligible(gs);
    if (!prepare_reduce(src, ws, groups, masks, filter, src->rows, src->cols, src->nnz, value_type, &p, !stream)) return 0;
    if (stream) {
        if (!run_stream(gs, ws, p, *filter)) return 0;
    } else if (src->rows != 0u) {
        if (!dispatch_value_type<launch_compressed>(value_type, *src, reduce_grid(src->rows), (hipStream_t) ws->stream, p, *filter,
                                                    outputs_of(ws, p.group_count)))
            return 0;
    }
    bind_result(ws, src->rows, p.group_count, out);
    return 1;
}

int csamd_compute_sparse_group_reduce_compressed(const csamd_compressed_view *src, csamd_sparse_group_reduce_workspace *ws,
                                                 const csamd_group_mask_config_view *groups,
                                                 const csamd_row_feature_mask_view *masks,
                                                 const csamd_sparse_group_filter_params *filter,
                                                 csamd_sparse_group_reduce_result *out) {
    return csamd_compute_sparse_group_reduce_compressed_typed(src, CSAMD_VALUE_F16, ws, groups, masks, filter, out);
}

int csamd_compute_sparse_group_reduce_sliced_ell_typed(const csamd_sliced_ell_view *src, unsigned int value_type,
                                                       csamd_sparse_group_reduce_workspace *ws,
                                                       const csamd_group_mask_config_view *groups,
                                                       const csamd_row_feature_mask_view *masks,
                                                       const csamd_sparse_group_filter_params *filter,
                                                       csamd_sparse_group_reduce_result *out) {
    reduce_prelude p;
    CSAMD_REQUIRE(src != nullptr, "compute_sparse_group_reduce(sliced_ell): null src");
    CSAMD_REQUIRE(src->slice_count == 0u || (src->slice_row_offsets != nullptr && src->slice_widths != nullptr &&
                                             src->slice_slot_offsets != nullptr),
                  "compute_sparse_group_reduce(sliced_ell): slice arrays missing");
    const gr_source gs = stream_source(*src, value_type);
    const bool stream = gr_eligible(gs);
    if (!prepare_reduce(src, ws, groups, masks, filter, src->rows, src->cols, src->nnz, value_type, &p, !stream)) return 0;
    if (stream) {
        if (!run_stream(gs, ws, p, *filter)) return 0;
    } else if (src->rows != 0u) {
        if (!dispatch_value_type<launch_sliced>(value_type, *src, reduce_grid(src->rows), (hipStream_t) ws->stream, p, *filter,
                                                outputs_of(ws, p.group_count)))
            return 0;
    }
    bind_result(ws, src->rows, p.group_count, out);
    return 1;
}

int csamd_compute_sparse_group_reduce_sliced_ell(const csamd_sliced_ell_view *src, csamd_sparse_group_reduce_workspace *ws,
                                                 const csamd_group_mask_config_view *groups,
                                                 const csamd_row_feature_mask_view *masks,
                                                 const csamd_sparse_group_filter_params *filter,
                                                 csamd_sparse_group_reduce_result *out) {
    return csamd_compute_sparse_group_reduce_sliced_ell_typed(src, CSAMD_VALUE_F16, ws, groups, masks, filter, out);
}

int csamd_compute_sparse_group_reduce_blocked_ell_typed(const csamd_blocked_ell_view *src, unsigned int value_type,
                                                        csamd_sparse_group_reduce_workspace *ws,
                                                        const csamd_group_mask_config_view *groups,
                                                        const csamd_row_feature_mask_view *masks,
                                                        const csamd_sparse_group_filter_params *filter,
                                                        csamd_sparse_group_reduce_result *out) {
    reduce_prelude p;
    CSAMD_REQUIRE(src != nullptr, "compute_sparse_group_reduce(blocked_ell): null src");
    const gr_bell_source gb{src->rows, src->cols, src->block_size, src->ell_cols, src->blockColIdx, src->val, value_type};
    const bool stream = gr_bell_eligible(gb);
    if (!prepare_reduce(src, ws, groups, masks, filter, src->rows, src->cols,
                        (unsigned long long) src->rows * src->ell_cols, value_type, &p, !stream))
        return 0;
    if (stream) {
        if (!ensure_scratch(ws, gr_scratch_bytes(gb.cols))) return 0;
        if (!gr_reduce_bell(gb, p.group_count != 0u ? p.fgm : nullptr, p.group_count, p.masks.row_keep, p.masks.feature_keep, *filter,
                            stream_outputs(outputs_of(ws, p.group_count)), ws->scratch_block, ws->scratch_bytes,
                            (hipStream_t) ws->stream))
            return 0;
    } else if (src->rows != 0u) {
        if (!dispatch_value_type<launch_blocked>(value_type, *src, reduce_grid(src->rows), (hipStream_t) ws->stream, p, *filter,
                                                 outputs_of(ws, p.group_count)))
            return 0;
    }
    bind_result(ws, src->rows, p.group_count, out);
    return 1;
}

int csamd_compute_sparse_group_reduce_blocked_ell(const csamd_blocked_ell_view *src, csamd_sparse_group_reduce_workspace *ws,
                                                  const csamd_group_mask_config_view *groups,
                                                  const csamd_row_feature_mask_view *masks,
                                                  const csamd_sparse_group_filter_params *filter,
                                                  csamd_sparse_group_reduce_result *out) {
    return csamd_compute_sparse_group_reduce_blocked_ell_typed(src, CSAMD_VALUE_F16, ws, groups, masks, filter, out);
}

// ---------------------------------------------------------------------------
// Fleet (one GPU's shard set).
// ---------------------------------------------------------------------------
void csamd_fleet_init(csamd_sparse_group_reduce_fleet_workspace *fleet) {
    if (fleet == nullptr) return;
    std::memset(fleet, 0, sizeof(*fleet));
}

void csamd_fleet_clear(csamd_sparse_group_reduce_fleet_workspace *fleet) {
    if (fleet == nullptr) return;
    if (fleet->devices != nullptr)
        for (unsigned int i = 0; i < fleet->slot_count; ++i) csamd_group_reduce_clear(fleet->devices + i);
    std::free(fleet->devices);
    std::free(fleet->results);
    std::free(fleet->slots);
    std::free(fleet->device_ids);
    csamd_fleet_init(fleet);
}

int csamd_setup_fleet(csamd_sparse_group_reduce_fleet_workspace *fleet, const csamd_sparse_group_reduce_fleet_config *config) {
    CSAMD_REQUIRE(fleet != nullptr, "setup_fleet: null fleet");
    csamd_fleet_clear(fleet);
    CSAMD_REQUIRE(config == nullptr || config->ranked_nccl == nullptr,
                  "setup_fleet: ranked communicators are not available on a one-GPU build");
    CSAMD_REQUIRE(config == nullptr || config->device_count == 0u || config->device_ids != nullptr,
                  "setup_fleet: device_ids missing");
    const int visible = csamd_device_count();
    CSAMD_REQUIRE(visible > 0, "setup_fleet: no HIP device");
    const unsigned int count = config != nullptr && config->device_count != 0u ? config->device_count : (unsigned int) visible;
    fleet->device_ids = (int *) std::calloc(count, sizeof(int));
    fleet->slots = (unsigned int *) std::calloc(count, sizeof(unsigned int));
    fleet->devices = (csamd_sparse_group_reduce_workspace *) std::calloc(count, sizeof(csamd_sparse_group_reduce_workspace));
    fleet->results = (csamd_sparse_group_reduce_result *) std::calloc(count, sizeof(csamd_sparse_group_reduce_result));
    if (fleet->device_ids == nullptr || fleet->slots == nullptr || fleet->devices == nullptr || fleet->results == nullptr) {
        csamd_fleet_clear(fleet);
        set_error("setup_fleet: host allocation failed");
        return 0;
    }
    fleet->slot_count = count;
    for (unsigned int i = 0; i < count; ++i) csamd_group_reduce_init(fleet->devices + i);
    for (unsigned int i = 0; i < count; ++i) {
        const int dev = config != nullptr && config->device_count != 0u ? config->device_ids[i] : (int) i;
        if (dev < 0 || dev >= visible) {
            csamd_fleet_clear(fleet);
            set_error("setup_fleet: device %d not visible", dev);
            return 0;
        }
        fleet->device_ids[i] = dev;
        fleet->slots[i] = (unsigned int) dev;
        if (!csamd_group_reduce_setup(fleet->devices + i, dev, nullptr)) {
            csamd_fleet_clear(fleet);
            return 0;
        }
    }
    return 1;
}

int csamd_compute_sparse_group_reduce_fleet_blocked_ell(const csamd_blocked_ell_view *src_by_slot,
                                                        csamd_sparse_group_reduce_fleet_workspace *fleet,
                                                        const csamd_group_mask_config_view *groups,
                                                        const csamd_row_feature_mask_view *masks,
                                                        const csamd_sparse_group_filter_params *filter,
                                                        csamd_sparse_group_reduce_fleet_result *out) {
    CSAMD_REQUIRE(src_by_slot != nullptr && fleet != nullptr && filter != nullptr && fleet->slot_count != 0u &&
                      fleet->devices != nullptr,
                  "compute_sparse_group_reduce_fleet: bad arguments");
    const unsigned int cols = src_by_slot[0].cols;
    for (unsigned int i = 0; i < fleet->slot_count; ++i) {
        CSAMD_REQUIRE(src_by_slot[i].cols == cols, "compute_sparse_group_reduce_fleet: column count differs across slots");
        if (!csamd_compute_sparse_group_reduce_blocked_ell(src_by_slot + i, fleet->devices + i, groups, masks, filter,
                                                          fleet->results + i))
            return 0;
    }
    if (out != nullptr) {
        out->slot_count = fleet->slot_count;
        out->leader_index = 0u;
        out->slot_results = fleet->results;
    }
    return 1;
}

int csamd_compute_sparse_group_reduce_fleet_sliced_ell(const csamd_sliced_ell_view *src_by_slot,
                                                       csamd_sparse_group_reduce_fleet_workspace *fleet,
                                                       const csamd_group_mask_config_view *groups,
                                                       const csamd_row_feature_mask_view *masks,
                                                       const csamd_sparse_group_filter_params *filter,
                                                       csamd_sparse_group_reduce_fleet_result *out) {
    CSAMD_REQUIRE(src_by_slot != nullptr && fleet != nullptr && filter != nullptr && fleet->slot_count != 0u &&
                      fleet->devices != nullptr,
                  "compute_sparse_group_reduce_fleet: bad arguments");
    const unsigned int cols = src_by_slot[0].cols;
    for (unsigned int i = 0; i < fleet->slot_count; ++i) {
        CSAMD_REQUIRE(src_by_slot[i].cols == cols, "compute_sparse_group_reduce_fleet: column count differs across slots");
        if (!csamd_compute_sparse_group_reduce_sliced_ell(src_by_slot + i, fleet->devices + i, groups, masks, filter,
                                                         fleet->results + i))
            return 0;
    }
    if (out != nullptr) {
        out->slot_count = fleet->slot_count;
        out->leader_index = 0u;
        out->slot_results = fleet->results;
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Masked reoptimize.
// ---------------------------------------------------------------------------
int csamd_predict_masked_sparse_reoptimization(const csamd_auto_reoptimize_prediction_input *input,
                                               csamd_auto_reoptimize_prediction *out) {
    CSAMD_REQUIRE(input != nullptr && out != nullptr, "predict_masked_sparse_reoptimization: null argument");
    out->should_reoptimize = 0u;
    out->estimated_net_gain_ms = 0.0f;
    return 1;
}

void csamd_reoptimize_init(csamd_masked_sparse_reoptimize_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_reoptimize_clear(csamd_masked_sparse_reoptimize_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    if (ws->owns_stream != 0 && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    reopt_private *st = (reopt_private *) ws->private_workspace;
    if (st != nullptr) {
        if (st->compressed_block != nullptr) (void) hipFree(st->compressed_block);
        if (st->scratch_block != nullptr) (void) hipFree(st->scratch_block);
        if (st->gather_block != nullptr) (void) hipFree(st->gather_block);
        std::free(st);
    }
    csamd_reoptimize_init(ws);
}

int csamd_reoptimize_setup(csamd_masked_sparse_reoptimize_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "reoptimize setup: null workspace");
    csamd_reoptimize_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice masked reoptimize setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags masked reoptimize");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    ws->private_workspace = std::calloc(1u, sizeof(reopt_private));
    CSAMD_REQUIRE(ws->private_workspace != nullptr, "reoptimize setup: host allocation failed");
    return 1;
}

int csamd_manual_reoptimize_masked_sparse(const csamd_blocked_ell_view *src, const csamd_row_feature_mask_view *masks,
                                          const csamd_masked_sparse_reoptimize_config *config,
                                          csamd_masked_sparse_reoptimize_workspace *ws,
                                          csamd_masked_sparse_reoptimize_result *out) {
    CSAMD_REQUIRE(src != nullptr && config != nullptr && ws != nullptr && out != nullptr,
                  "manual_reoptimize_masked_sparse: null argument");
    CSAMD_HIP(hipSetDevice(ws->device >= 0 ? ws->device : 0), "hipSetDevice masked reoptimize");
    if (ws->private_workspace == nullptr) {
        ws->private_workspace = std::calloc(1u, sizeof(reopt_private));
        CSAMD_REQUIRE(ws->private_workspace != nullptr, "reoptimize: host allocation failed");
    }
    reopt_private *st = (reopt_private *) ws->private_workspace;
    hipStream_t s = (hipStream_t) ws->stream;
    csamd_row_feature_mask_view m;
    if (masks != nullptr) m = *masks;
    else std::memset(&m, 0, sizeof(m));
    const unsigned int output_rows = config->output_rows != 0u ? config->output_rows : src->rows;
    const unsigned int output_cols = config->output_cols != 0u ? config->output_cols : src->cols;
    // The reference sizes this as a u32 product (mask_groups.cu:971); here it
    // is computed in 64 bits and rejected if the CSR view cannot index it.
    const unsigned long long max_nnz = (unsigned long long) src->rows * src->ell_cols;
    CSAMD_REQUIRE(max_nnz <= 0xffffffffull, "manual_reoptimize: rows*ell_cols = %llu exceeds u32", max_nnz);
    const size_t major_bytes = (size_t) (output_rows + 1ull) * sizeof(unsigned int);
    const size_t minor_off = align_up(major_bytes, 16u);
    const size_t val_off = align_up(minor_off + max_nnz * sizeof(unsigned int), 16u);
    const size_t bytes = val_off + max_nnz * sizeof(uint16_t);
    if (bytes > st->compressed_bytes) {
        if (st->compressed_block != nullptr) (void) hipFree(st->compressed_block);
        st->compressed_block = nullptr;
        st->compressed_bytes = 0u;
        CSAMD_HIP(hipMalloc(&st->compressed_block, bytes), "hipMalloc masked compressed output");
        st->compressed_bytes = bytes;
    }
    const size_t map_off = align_up((size_t) (output_rows == 0u ? 1u : output_rows) * sizeof(unsigned int), 256u);
    const size_t keep_off = map_off + align_up((size_t) (src->cols == 0u ? 1u : src->cols) * sizeof(unsigned int), 256u);
    const size_t src_off = keep_off + align_up((size_t) (src->cols == 0u ? 1u : src->cols) * sizeof(unsigned int), 256u);
    const size_t flag_off = src_off + align_up((size_t) (output_rows == 0u ? 1u : output_rows) * sizeof(unsigned int), 256u);
    const size_t scratch_needed = flag_off + 256u;
    if (scratch_needed > st->scratch_bytes) {
        if (st->scratch_block != nullptr) (void) hipFree(st->scratch_block);
        st->scratch_block = nullptr;
        st->scratch_bytes = 0u;
        CSAMD_HIP(hipMalloc(&st->scratch_block, scratch_needed), "hipMalloc masked reoptimize scratch");
        st->scratch_bytes = scratch_needed;
    }
    char *base = (char *) st->compressed_block;
    csamd_compressed_view c;
    c.rows = output_rows;
    c.cols = output_cols;
    c.nnz = (unsigned int) max_nnz;
    c.axis = CSAMD_COMPRESSED_BY_ROW;
    c.majorPtr = (unsigned int *) base;
    c.minorIdx = (unsigned int *) (base + minor_off);
    c.val = (csamd_half *) (base + val_off);
    unsigned int *row_write = (unsigned int *) st->scratch_block;
    unsigned int *col_map = (unsigned int *) ((char *) st->scratch_block + map_off);
    // keep bits for the 16-byte chunk path (block_size a multiple of 8, at most 32)
    const bool keep_bits = src->cols != 0u && src->block_size != 0u && src->block_size % 8u == 0u && src->block_size <= 32u;
    unsigned int *block_keep = keep_bits ? (unsigned int *) ((char *) st->scratch_block + keep_off) : nullptr;
    const unsigned int keep_blocks = keep_bits ? (unsigned int) (((unsigned long long) src->cols + src->block_size - 1u) / src->block_size) : 0u;

    unsigned int *src_row = (unsigned int *) ((char *) st->scratch_block + src_off);
    unsigned int *overflow = (unsigned int *) ((char *) st->scratch_block + flag_off);

    if (!fill_u32(c.majorPtr, (unsigned long long) output_rows + 1ull, 0u, s)) return 0;
    if (!fill_u32(row_write, output_rows, 0u, s)) return 0;
    if (!fill_u32(overflow, 1u, 0u, s)) return 0;
    const unsigned int grid = reduce_grid(src->rows);
    if (src->cols != 0u) {
        reopt_col_map_kernel<<<grid_for(src->cols, 256u, 1024u), 256, 0, s>>>(m, src->cols, output_cols, col_map);
        CSAMD_HIP(hipGetLastError(), "reopt_col_map_kernel");
    }
    if (keep_bits) {
        reopt_block_keep_kernel<<<grid_for(keep_blocks, 256u, 1024u), 256, 0, s>>>(src->cols, src->block_size, col_map, block_keep);
        CSAMD_HIP(hipGetLastError(), "reopt_block_keep_kernel");
    }
    // One read of the kept rows (16-byte chunk path): each row's survivors to
    // a region of cap = min(ell_cols, output_cols) entries, enough while a
    // row's columns are distinct and the feature remap is one-to-one, then
    // packed once the row pointers are known; a row with more survivors
    // sends the call through count + scatter.
    const unsigned long long cap = src->ell_cols < output_cols ? src->ell_cols : output_cols;
    const unsigned long long gather_bytes = align_up((size_t) (src->rows * cap) * 4u, 256u) + (size_t) (src->rows * cap) * 2u;
    const bool chunk_path = src->block_size != 0u && src->block_size % 8u == 0u && src->ell_cols % 8u == 0u &&
                            ((uintptr_t) src->val & 15u) == 0u;
    const bool gather = chunk_path && src->rows != 0u && cap != 0u && gather_bytes <= (1ull << 30);
    unsigned int host_nnz = 0u;
    bool done = false;
    if (gather) {
        if (gather_bytes > st->gather_bytes) {
            if (st->gather_block != nullptr) (void) hipFree(st->gather_block);
            st->gather_block = nullptr;
            st->gather_bytes = 0u;
            CSAMD_HIP(hipMalloc(&st->gather_block, gather_bytes), "hipMalloc masked reoptimize gather");
            st->gather_bytes = gather_bytes;
        }
        unsigned int *t_idx = (unsigned int *) st->gather_block;
        uint16_t *t_val = (uint16_t *) ((char *) st->gather_block + align_up((size_t) (src->rows * cap) * 4u, 256u));
        // streaming form: keep bits, a power-of-two number of 16-byte chunks
        // per block (block size 8, 16 or 32) and the keep bits in LDS
        const unsigned int cpb = src->block_size / 8u;
        const bool stream = block_keep != nullptr && (cpb & (cpb - 1u)) == 0u && (size_t) keep_blocks * 4u <= 32768u;
        if (stream) {
            reopt_gather_stream_kernel<<<grid, MG_THREADS, (size_t) keep_blocks * 4u, s>>>(
                *src, m, output_rows, (unsigned int) cap, (unsigned int) __builtin_ctz(cpb), c.majorPtr, src_row, overflow, t_idx, t_val,
                block_keep, keep_blocks);
            CSAMD_HIP(hipGetLastError(), "reopt_gather_stream_kernel");
        } else {
            reopt_gather_kernel<<<grid, MG_THREADS, 0, s>>>(*src, m, output_rows, (unsigned int) cap, c.majorPtr, src_row, overflow,
                                                            t_idx, t_val, col_map, block_keep, keep_blocks);
            CSAMD_HIP(hipGetLastError(), "reopt_gather_kernel");
        }
        reopt_scan_kernel<<<1, 1024, 0, s>>>(output_rows, c.majorPtr);
        CSAMD_HIP(hipGetLastError(), "reopt_scan_kernel");
        unsigned int flag = 0u;
        CSAMD_HIP(hipMemcpyAsync(&flag, overflow, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync reoptimize flag");
        CSAMD_HIP(hipMemcpyAsync(&host_nnz, c.majorPtr + output_rows, sizeof(unsigned int), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync masked nnz");
        CSAMD_HIP(hipStreamSynchronize(s), "hipStreamSynchronize masked reoptimize");
        if (flag == 0u) {
            reopt_copy_kernel<<<grid_for(output_rows, MG_WAVES, (unsigned int) device_cu_count() * 16u), MG_THREADS, 0, s>>>(
                output_rows, (unsigned int) cap, c.majorPtr, src_row, t_idx, t_val, stream ? col_map : nullptr, c.minorIdx,
                (uint16_t *) c.val);
            CSAMD_HIP(hipGetLastError(), "reopt_copy_kernel");
            CSAMD_HIP(hipStreamSynchronize(s), "hipStreamSynchronize masked reoptimize");
            done = true;
        } else if (!fill_u32(c.majorPtr, (unsigned long long) output_rows + 1ull, 0u, s)) {
            return 0;
        }
    }
    if (!done) {
        if (src->rows != 0u) {
            reopt_count_kernel<<<grid, MG_THREADS, 0, s>>>(*src, m, output_rows, output_cols, c.majorPtr, col_map, block_keep,
                                                           keep_blocks);
            CSAMD_HIP(hipGetLastError(), "reopt_count_kernel");
        }
        reopt_scan_kernel<<<1, 1024, 0, s>>>(output_rows, c.majorPtr);
        CSAMD_HIP(hipGetLastError(), "reopt_scan_kernel");
        if (src->rows != 0u) {
            reopt_scatter_kernel<<<grid, MG_THREADS, 0, s>>>(*src, m, output_rows, output_cols, c.majorPtr, row_write, c.minorIdx,
                                                             (uint16_t *) c.val, col_map, block_keep, keep_blocks);
            CSAMD_HIP(hipGetLastError(), "reopt_scatter_kernel");
        }
        CSAMD_HIP(hipMemcpyAsync(&host_nnz, c.majorPtr + output_rows, sizeof(unsigned int), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync masked nnz");
        CSAMD_HIP(hipStreamSynchronize(s), "hipStreamSynchronize masked reoptimize");
    }
    c.nnz = host_nnz;
    out->compressed = c;
    std::memset(&out->blocked_ell, 0, sizeof(out->blocked_ell));
    out->metadata.kept_rows = output_rows;
    out->metadata.kept_features = output_cols;
    out->metadata.live_nnz = host_nnz;
    const unsigned long long rbc = src->block_size != 0u ? (src->rows + src->block_size - 1ull) / src->block_size : 0ull;
    const unsigned long long w = src->block_size != 0u ? src->ell_cols / src->block_size : 0ull;
    out->metadata.bytes_before = (unsigned long long) src->rows * src->ell_cols * 2ull + rbc * w * 4ull;
    out->metadata.bytes_after = (unsigned long long) (output_rows + 1ull) * 4ull + (unsigned long long) host_nnz * 6ull;
    out->metadata.output_layout = CSAMD_LAYOUT_COMPRESSED;
    return 1;
}

} // extern "C"
