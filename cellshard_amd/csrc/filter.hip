// Masked Blocked-ELL -> CSR compaction with row / column remaps, and the
// same followed by a major-nnz bucket plan and rebuild (the step after the
// grouped reduce: QC keep -> compact -> rebucket).
//
// Reference: src/convert/filtered_blocked_ell_to_compressed.cuh:125-415
// (filtered_blocked_ell_{stats,result,workspace}, filter_blocked_ell_to_compressed,
// build_bucketed_filtered_blocked_ell_major_view) and
// src/bucket/operators/major_nnz.cuh:11-213 (build_bucketed_major_view).
//
// An entry of source row r at ell position e survives when the row is kept
// (keep_rows null or nonzero) and maps to out_row = row_remap[r] (or r) <
// output_rows, its column block is valid, col < cols, the stored value is
// nonzero, the column is kept, and out_col = col_remap[col] (or col) <
// output_cols.  Output rows hold their entries in source ell order.  The
// row pointers are a plain exclusive scan of the per-row counts: row i of
// the output is [ptr[i], ptr[i + 1]).  (The reference writes each count one
// place to the right and then runs an exclusive scan, which shifts every
// row by one more place; the row contents here are what its emit kernel
// intends.)  Two kept rows mapped to one output row are an error.
//
// gfx950 plan: one workgroup per source row, threads across the ell
// positions (16-byte value groups, four in flight per thread); a count
// pass, one exclusive scan, then an emit pass that places each thread's
// entries with a block prefix sum.  Padding (zero values) is rejected
// before any gather; for block sizes up to 32 that hold whole 16-byte
// groups, a group's kept columns come from one word of per-block-column
// keep bits, so the column map is read only for surviving values.  Rows that are not whole 16-byte groups take a
// one-wave-per-row path.
#include "common.hh"
#include "stream.hh"

#include <cstdlib>
#include <cstring>

namespace csamd {
namespace {

constexpr unsigned int FL_THREADS = 256u;
constexpr unsigned int FL_WAVES = FL_THREADS / 64u;
constexpr unsigned int FL_LDS_SLOTS = 8192u;  // block columns of a row block staged in LDS (32 KB)

template<typename S> __device__ __forceinline__ bool fl_nonzero(S raw);
template<> __device__ __forceinline__ bool fl_nonzero<uint16_t>(uint16_t raw) { return (raw & 0x7fffu) != 0u; }
template<> __device__ __forceinline__ bool fl_nonzero<float>(float raw) { return raw != 0.0f; }
template<> __device__ __forceinline__ bool fl_nonzero<uint8_t>(uint8_t raw) { return (raw & 0x7fu) != 0u; }

struct fl_masks {
    unsigned int output_rows, output_cols;
    unsigned int row_base;  // shard row of the part's first row (the output row without a remap)
    const unsigned char *keep_rows, *keep_cols;
    const unsigned int *row_remap, *col_remap;
    const unsigned int *col_map;  // keep and remap of every column in one lookup (~0u: dropped), or null
    // Kept-column bits of each block column (bit j: column bc * block_size + j
    // survives), for block sizes up to 32 that are whole 16-byte groups; null
    // otherwise.  keep_blocks = ceil(cols / block_size).
    const unsigned int *block_keep;
    unsigned int keep_blocks;
};

// col_map[c] = output column of source column c, or ~0u when it is dropped
// (not kept, or remapped past output_cols): one gather per surviving value
// instead of two dependent ones.
__global__ void fl_col_map_kernel(unsigned int cols, fl_masks m, unsigned int *__restrict__ col_map) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
        unsigned int o = ~0u;
        if (m.keep_cols == nullptr || m.keep_cols[c] != 0u) {
            o = m.col_remap != nullptr ? m.col_remap[c] : c;
            if (o >= m.output_cols) o = ~0u;
        }
        col_map[c] = o;
    }
}

// block_keep[b] bit j = column b * bs + j is kept (below cols, col_map not ~0u).
__global__ void fl_block_keep_kernel(unsigned int cols, unsigned int bs, const unsigned int *__restrict__ col_map,
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

// Output row of source row r, or output_rows when the row is dropped.
__device__ __forceinline__ unsigned int fl_out_row(const fl_masks &m, unsigned int r) {
    if (m.keep_rows != nullptr && m.keep_rows[r] == 0u) return m.output_rows;
    const unsigned int o = m.row_remap != nullptr ? m.row_remap[r] : m.row_base + r;
    return o < m.output_rows ? o : m.output_rows;
}

// Output column of ell position e of row r (block row rb), or ~0u if the
// entry is dropped; *v gets its stored value.
template<typename S>
__device__ __forceinline__ unsigned int fl_out_col(const csamd_blocked_ell_view &src, const S *val, const fl_masks &m, unsigned int width,
                                                   unsigned int rb, unsigned int r, unsigned int e, S *v) {
    const unsigned int bs = src.block_size;
    const unsigned int slot = e / bs;
    const unsigned int bc = src.blockColIdx[(unsigned long long) rb * width + slot];
    if (bc == CSAMD_INVALID_COL) return ~0u;
    const unsigned int col = bc * bs + (e - slot * bs);
    if (col >= src.cols) return ~0u;
    *v = val[(unsigned long long) r * src.ell_cols + e];
    if (!fl_nonzero<S>(*v)) return ~0u;
    if (m.keep_cols != nullptr && m.keep_cols[col] == 0u) return ~0u;
    const unsigned int oc = m.col_remap != nullptr ? m.col_remap[col] : col;
    return oc < m.output_cols ? oc : ~0u;
}

// A row's values are read in 16-byte groups (PER values per lane, U
// groups per lane in flight) when the value rows are whole, 16-byte aligned
// groups (ell_cols * sizeof(S) % 16 == 0 and an aligned array); otherwise
// one value per lane at a time.  Groups are in position order: group g of a
// pass holds positions base + (g * 64 + lane) * PER + k.
template<typename S> union fl_group {
    uint4 q;
    S v[16u / sizeof(S)];
};

// The row block's block columns, staged in LDS or read in place.  The two
// are read through pointers of their own address spaces: a generic pointer
// that may point at either would make every read a flat access, which the
// compiler counts against both memory counters (it then waits for all
// outstanding value loads at each read).
struct fl_slots_ref {
    const __attribute__((address_space(3))) unsigned int *lds;
    const CSAMD_GLOBAL unsigned int *glob;
    bool in_lds;
    __device__ __forceinline__ unsigned int operator[](unsigned int k) const { return in_lds ? lds[k] : glob[k]; }
};

// Surviving entries of group g (bit k: position k of the lane's group);
// slots: the row block's block columns.
template<typename S>
__device__ __forceinline__ unsigned int fl_group_mask(const csamd_blocked_ell_view &src, const fl_masks &m, const fl_slots_ref &slots,
                                                      unsigned int e, const fl_group<S> &g, unsigned int *oc) {
    constexpr unsigned int PER = 16u / sizeof(S);
    unsigned int mask = 0u;
    if (m.block_keep != nullptr) {
        // The group lies in one block column (block_size % PER == 0): one
        // lookup of its kept-column bits, and col_map only for survivors.
        unsigned int nz = 0u;
#pragma unroll
        for (unsigned int k = 0; k < PER; ++k) {
            oc[k] = ~0u;
            if (e + k < src.ell_cols && fl_nonzero<S>(g.v[k])) nz |= 1u << k;
        }
        if (nz == 0u) return 0u;
        const unsigned int slot = e / src.block_size;
        const unsigned int bc = slots[slot];
        if (bc >= m.keep_blocks) return 0u;  // CSAMD_INVALID_COL included
        const unsigned int j0 = e - slot * src.block_size;
        mask = nz & (m.block_keep[bc] >> j0);
#pragma unroll
        for (unsigned int k = 0; k < PER; ++k)
            if ((mask >> k) & 1u) oc[k] = m.col_map[bc * src.block_size + j0 + k];
        return mask;
    }
#pragma unroll
    for (unsigned int k = 0; k < PER; ++k) {
        oc[k] = ~0u;
        if (e + k >= src.ell_cols || !fl_nonzero<S>(g.v[k])) continue;  // most padding stops here, before any gather
        const unsigned int pos = e + k;
        const unsigned int slot = pos / src.block_size;
        const unsigned int bc = slots[slot];
        if (bc == CSAMD_INVALID_COL) continue;
        const unsigned int col = bc * src.block_size + (pos - slot * src.block_size);
        if (col >= src.cols) continue;
        const unsigned int o = m.col_map[col];
        if (o == ~0u) continue;
        oc[k] = o;
        mask |= 1u << k;
    }
    return mask;
}

// The row block's block columns: staged in LDS when they fit (the caller
// syncs), else read in place.
__device__ __forceinline__ fl_slots_ref fl_slots(const csamd_blocked_ell_view &src, unsigned int width, unsigned int rb,
                                                 unsigned int *lds_slots, unsigned int lds_cap) {
    const CSAMD_GLOBAL unsigned int *g = gmem((const unsigned int *) src.blockColIdx) + (unsigned long long) rb * width;
    const bool in_lds = width <= lds_cap;
    if (in_lds)
        for (unsigned int k = threadIdx.x; k < width; k += blockDim.x) lds_slots[k] = g[k];
    return fl_slots_ref{(const __attribute__((address_space(3))) unsigned int *) lds_slots, g, in_lds};
}

template<typename S, unsigned int U>
__device__ __forceinline__ void fl_load_groups(const S *row, unsigned int ell_cols, unsigned int e0, fl_group<S> (&g)[U]) {
    constexpr unsigned int PER = 16u / sizeof(S);
    const unsigned int lane = lane_id();
#pragma unroll
    for (unsigned int u = 0; u < U; ++u) {
        const unsigned int e = e0 + (u * 64u + lane) * PER;
        g[u].q = e < ell_cols ? *(const uint4 *) (row + e) : make_uint4(0u, 0u, 0u, 0u);
    }
}

// Count pass: counts[out_row] = surviving entries of the row; claim[] marks
// output rows already taken (a second claim sets *dup).  Grouped: one
// workgroup per row (its four waves take interleaved group rows, so a CU
// streams few rows at a time); otherwise one wave per row.
template<typename S, unsigned int U>
__global__ void __launch_bounds__(FL_THREADS) fl_count_kernel(csamd_blocked_ell_view src, const S *__restrict__ val, fl_masks m,
                                                              bool grouped, unsigned int slot_cap, unsigned int *__restrict__ counts,
                                                              unsigned int *__restrict__ claim, unsigned int *__restrict__ dup) {
    constexpr unsigned int PER = 16u / sizeof(S);
    const unsigned int width = src.ell_cols / src.block_size;
    const unsigned int lane = lane_id(), wave = threadIdx.x >> 6;
    __shared__ unsigned int red[FL_WAVES];
    extern __shared__ unsigned int lds_slots[];
    if (grouped) {
        unsigned int cur_rb = ~0u;
        fl_slots_ref slots{nullptr, nullptr, false};
        for (unsigned int r = blockIdx.x; r < src.rows; r += gridDim.x) {
            const unsigned int o = fl_out_row(m, r);
            if (o >= m.output_rows) continue;
            const unsigned int rb = r / src.block_size;
            if (rb != cur_rb) {
                __syncthreads();
                slots = fl_slots(src, width, rb, lds_slots, slot_cap);
                cur_rb = rb;
                __syncthreads();
            }
            const S *row = val + (unsigned long long) r * src.ell_cols;
            unsigned int n = 0u;
            for (unsigned int e0 = wave * 64u * U * PER; e0 < src.ell_cols; e0 += FL_WAVES * 64u * U * PER) {
                fl_group<S> g[U];
                fl_load_groups<S, U>(row, src.ell_cols, e0, g);
#pragma unroll
                for (unsigned int u = 0; u < U; ++u) {
                    unsigned int oc[PER];
                    n += (unsigned int) __popc(fl_group_mask<S>(src, m, slots, e0 + (u * 64u + lane) * PER, g[u], oc));
                }
            }
            n = wave_sum_u32(n);
            if (lane == 0u) red[wave] = n;
            __syncthreads();
            if (threadIdx.x == 0u) {
                unsigned int t = 0u;
                for (unsigned int k = 0; k < FL_WAVES; ++k) t += red[k];
                if (atomicExch(&claim[o], 1u) != 0u) *dup = o + 1u;
                counts[o] = t;
            }
            __syncthreads();
        }
        return;
    }
    for (unsigned int r = blockIdx.x * FL_WAVES + wave; r < src.rows; r += gridDim.x * FL_WAVES) {
        const unsigned int o = fl_out_row(m, r);
        if (o >= m.output_rows) continue;
        const unsigned int rb = r / src.block_size;
        unsigned int n = 0u;
        for (unsigned int e = lane; e < src.ell_cols; e += 64u) {
            S v;
            n += fl_out_col<S>(src, val, m, width, rb, r, e, &v) != ~0u ? 1u : 0u;
        }
        n = wave_sum_u32(n);
        if (lane == 0u) {
            if (atomicExch(&claim[o], 1u) != 0u) *dup = o + 1u;
            counts[o] = n;
        }
    }
}

// Emit pass, in position order.  Grouped: one workgroup per row, a block
// prefix sum of the threads' surviving counts per pass of FL_THREADS x U
// groups places each thread's entries; otherwise one wave per row with a
// ballot per 64-position chunk.
template<typename S, unsigned int U>
__global__ void __launch_bounds__(FL_THREADS) fl_emit_kernel(csamd_blocked_ell_view src, const S *__restrict__ val, fl_masks m,
                                                             bool grouped, unsigned int slot_cap, const unsigned int *__restrict__ ptr,
                                                             unsigned int *__restrict__ out_idx, S *__restrict__ out_val) {
    constexpr unsigned int PER = 16u / sizeof(S);
    const unsigned int width = src.ell_cols / src.block_size;
    const unsigned int lane = lane_id();
    __shared__ unsigned int smem[16];
    extern __shared__ unsigned int lds_slots[];
    if (grouped) {
        unsigned int cur_rb = ~0u;
        fl_slots_ref slots{nullptr, nullptr, false};
        for (unsigned int r = blockIdx.x; r < src.rows; r += gridDim.x) {
            const unsigned int o = fl_out_row(m, r);
            if (o >= m.output_rows) continue;
            const unsigned int rb = r / src.block_size;
            if (rb != cur_rb) {
                __syncthreads();
                slots = fl_slots(src, width, rb, lds_slots, slot_cap);
                cur_rb = rb;
                __syncthreads();
            }
            const S *row = val + (unsigned long long) r * src.ell_cols;
            unsigned int w = ptr[o];
            // thread t takes groups t, t + FL_THREADS, ... of the pass (position order = thread order per group row)
            for (unsigned int e0 = 0; e0 < src.ell_cols; e0 += FL_THREADS * U * PER) {
                fl_group<S> g[U];
#pragma unroll
                for (unsigned int u = 0; u < U; ++u) {
                    const unsigned int e = e0 + (u * FL_THREADS + threadIdx.x) * PER;
                    g[u].q = e < src.ell_cols ? *(const uint4 *) (row + e) : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (unsigned int u = 0; u < U; ++u) {
                    unsigned int oc[PER];
                    const unsigned int mask = fl_group_mask<S>(src, m, slots, e0 + (u * FL_THREADS + threadIdx.x) * PER, g[u], oc);
                    const unsigned int c = (unsigned int) __popc(mask);
                    unsigned int total = 0u;
                    unsigned int k = w + block_exclusive_scan_u32(c, smem, &total);
#pragma unroll
                    for (unsigned int j = 0; j < PER; ++j) {
                        if ((mask >> j) & 1u) {
                            out_idx[k] = oc[j];
                            out_val[k] = g[u].v[j];
                            ++k;
                        }
                    }
                    w += total;
                }
            }
        }
        return;
    }
    for (unsigned int r = blockIdx.x * FL_WAVES + (threadIdx.x >> 6); r < src.rows; r += gridDim.x * FL_WAVES) {
        const unsigned int o = fl_out_row(m, r);
        if (o >= m.output_rows) continue;
        const unsigned int rb = r / src.block_size;
        unsigned int w = ptr[o];
        for (unsigned int e0 = 0; e0 < src.ell_cols; e0 += 64u) {
            const unsigned int e = e0 + lane;
            S v = S(0);
            const unsigned int oc = e < src.ell_cols ? fl_out_col<S>(src, val, m, width, rb, r, e, &v) : ~0u;
            const unsigned long long bal = __ballot(oc != ~0u);
            if (oc != ~0u) {
                const unsigned int k = w + (unsigned int) __popcll(bal & lanemask_lt());
                out_idx[k] = oc;
                out_val[k] = v;
            }
            w += (unsigned int) __popcll(bal);
        }
    }
}

// Gather pass, the one read of the kept value rows when every row's
// survivors fit a fixed region: one workgroup per row, as the grouped emit
// pass, but a row's surviving entries go in position order to its own
// region of a scratch array (source row r at r * cap, at most cap entries;
// a row with more sets *overflow and the caller falls back to count + emit).
// The row's count, its source row and the output-row claim are recorded as
// in the count pass.
template<typename S, unsigned int U>
__global__ void __launch_bounds__(FL_THREADS) fl_gather_kernel(csamd_blocked_ell_view src, const S *__restrict__ val, fl_masks m,
                                                               unsigned int slot_cap, unsigned int row0, unsigned int cap,
                                                               unsigned int *__restrict__ counts, unsigned int *__restrict__ claim,
                                                               unsigned int *__restrict__ dup, unsigned int *__restrict__ src_row,
                                                               unsigned int *__restrict__ overflow, unsigned int *__restrict__ t_idx,
                                                               S *__restrict__ t_val) {
    constexpr unsigned int PER = 16u / sizeof(S);
    const unsigned int width = src.ell_cols / src.block_size;
    __shared__ unsigned int smem[16];
    extern __shared__ unsigned int lds_slots[];
    unsigned int cur_rb = ~0u;
    fl_slots_ref slots{nullptr, nullptr, false};
    CSAMD_GLOBAL unsigned int *gi = gmem(t_idx);
    CSAMD_GLOBAL S *gv = gmem(t_val);
    for (unsigned int r = blockIdx.x; r < src.rows; r += gridDim.x) {
        const unsigned int o = fl_out_row(m, r);
        if (o >= m.output_rows) continue;
        const unsigned int rb = r / src.block_size;
        if (rb != cur_rb) {
            __syncthreads();
            slots = fl_slots(src, width, rb, lds_slots, slot_cap);
            cur_rb = rb;
            __syncthreads();
        }
        const S *row = val + (unsigned long long) r * src.ell_cols;
        const unsigned long long base = (unsigned long long) (row0 + r) * cap;
        unsigned int w = 0u;
        for (unsigned int e0 = 0; e0 < src.ell_cols; e0 += FL_THREADS * U * PER) {
            fl_group<S> g[U];
#pragma unroll
            for (unsigned int u = 0; u < U; ++u) {
                const unsigned int e = e0 + (u * FL_THREADS + threadIdx.x) * PER;
                g[u].q = e < src.ell_cols ? *(const uint4 *) (row + e) : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (unsigned int u = 0; u < U; ++u) {
                unsigned int oc[PER];
                const unsigned int mask = fl_group_mask<S>(src, m, slots, e0 + (u * FL_THREADS + threadIdx.x) * PER, g[u], oc);
                unsigned int total = 0u;
                unsigned int k = w + block_exclusive_scan_u32((unsigned int) __popc(mask), smem, &total);
#pragma unroll
                for (unsigned int j = 0; j < PER; ++j) {
                    if ((mask >> j) & 1u) {
                        if (k < cap) {
                            gi[base + k] = oc[j];
                            gv[base + k] = g[u].v[j];
                        }
                        ++k;
                    }
                }
                w += total;
            }
        }
        if (threadIdx.x == 0u) {
            if (atomicExch(&claim[o], 1u) != 0u) *dup = o + 1u;
            counts[o] = w;
            src_row[o] = row0 + r;
            if (w > cap) *overflow = 1u;
        }
    }
}

// Streaming gather pass (block-keep bits available, a power-of-two number
// of 16-byte groups per block, the keep bits and the row block's block
// columns staged in LDS): like fl_gather_kernel, but each thread takes 16
// consecutive values per step (16 / PER consecutive groups), the next
// step's loads are in flight while the current one is consumed, one block
// scan per step orders the survivors, and nothing in the loop reads global
// memory besides the value stream (LDS reads do not wait for the loads in
// flight).  Survivors are written with their source columns; the copy pass
// maps them through col_map.
template<typename S>
__global__ void __launch_bounds__(FL_THREADS) fl_gather_stream_kernel(csamd_blocked_ell_view src, const S *__restrict__ val, fl_masks m,
                                                                      unsigned int cpb_log2, unsigned int row0, unsigned int cap,
                                                                      unsigned int *__restrict__ counts, unsigned int *__restrict__ claim,
                                                                      unsigned int *__restrict__ dup, unsigned int *__restrict__ src_row,
                                                                      unsigned int *__restrict__ overflow, unsigned int *__restrict__ t_idx,
                                                                      S *__restrict__ t_val) {
    constexpr unsigned int PER = 16u / sizeof(S);
    constexpr unsigned int UC = 16u / PER;           // groups per thread and step (16 values)
    constexpr unsigned int STEP = FL_THREADS * UC;   // groups per step
    const unsigned int bs = src.block_size;
    const unsigned int width = src.ell_cols / bs;
    const unsigned int ng = src.ell_cols / PER;
    const unsigned int cpb_mask = (1u << cpb_log2) - 1u;
    __shared__ unsigned int smem[16];
    extern __shared__ unsigned int lds_dyn[];  // [keep_blocks] keep bits, then [width] block columns
    typedef const __attribute__((address_space(3))) unsigned int *lds_ptr;
    unsigned int *lkeep_w = lds_dyn;
    unsigned int *lslots_w = lds_dyn + ((m.keep_blocks + 3u) & ~3u);
    const lds_ptr lkeep = (lds_ptr) lkeep_w;
    const lds_ptr lslots = (lds_ptr) lslots_w;
    for (unsigned int k = threadIdx.x; k < m.keep_blocks; k += FL_THREADS) lkeep_w[k] = gmem(m.block_keep)[k];
    CSAMD_GLOBAL unsigned int *gi = gmem(t_idx);
    CSAMD_GLOBAL S *gv = gmem(t_val);
    unsigned int cur_rb = ~0u;
    for (unsigned int r = blockIdx.x; r < src.rows; r += gridDim.x) {
        const unsigned int o = fl_out_row(m, r);
        if (o >= m.output_rows) continue;
        const unsigned int rb = r / bs;
        if (rb != cur_rb) {
            __syncthreads();
            const CSAMD_GLOBAL unsigned int *g = gmem((const unsigned int *) src.blockColIdx) + (unsigned long long) rb * width;
            for (unsigned int k = threadIdx.x; k < width; k += FL_THREADS) lslots_w[k] = g[k];
            cur_rb = rb;
            __syncthreads();
        }
        const CSAMD_GLOBAL u32x4 *rq = (const CSAMD_GLOBAL u32x4 *) gmem(val + (unsigned long long) r * src.ell_cols);
        const unsigned long long base = (unsigned long long) (row0 + r) * cap;
        unsigned int w = 0u;
        auto load = [&](unsigned int g0, u32x4 (&buf)[UC]) {
#pragma unroll
            for (unsigned int u = 0; u < UC; ++u) buf[u] = rq[min(g0 + threadIdx.x * UC + u, ng - 1u)];
        };
        auto consume = [&](unsigned int g0, const u32x4 (&buf)[UC]) {
            unsigned int mask[UC], col0[UC];
            unsigned int cnt = 0u;
#pragma unroll
            for (unsigned int u = 0; u < UC; ++u) {
                const unsigned int gidx = g0 + threadIdx.x * UC + u;
                unsigned int nz = 0u;
#pragma unroll
                for (unsigned int k = 0; k < PER; ++k)
                    if (fl_nonzero<S>(chunk_val<S>(buf[u], k))) nz |= 1u << k;
                const unsigned int bc = lslots[min(gidx, ng - 1u) >> cpb_log2];
                const unsigned int j0 = (gidx & cpb_mask) * PER;
                const unsigned int kb = bc < m.keep_blocks ? lkeep[bc] : 0u;  // CSAMD_INVALID_COL included
                mask[u] = gidx < ng ? nz & (kb >> j0) : 0u;
                col0[u] = bc * bs + j0;
                cnt += (unsigned int) __popc(mask[u]);
            }
            unsigned int total = 0u;
            unsigned int k = w + block_exclusive_scan_u32(cnt, smem, &total);
#pragma unroll
            for (unsigned int u = 0; u < UC; ++u) {
#pragma unroll
                for (unsigned int j = 0; j < PER; ++j) {
                    if ((mask[u] >> j) & 1u) {
                        if (k < cap) {
                            gi[base + k] = col0[u] + j;
                            gv[base + k] = chunk_val<S>(buf[u], j);
                        }
                        ++k;
                    }
                }
            }
            w += total;
        };
        if (ng != 0u) {
            u32x4 A[UC], B[UC];
            load(0u, A);
            for (unsigned int g0 = 0u; g0 < ng; g0 += 2u * STEP) {
                load(g0 + STEP, B);
                consume(g0, A);
                if (g0 + STEP >= ng) break;
                load(g0 + 2u * STEP, A);
                consume(g0 + STEP, B);
            }
        }
        if (threadIdx.x == 0u) {
            if (atomicExch(&claim[o], 1u) != 0u) *dup = o + 1u;
            counts[o] = w;
            src_row[o] = row0 + r;
            if (w > cap) *overflow = 1u;
        }
    }
}

// Copy pass after the gather: output row o's entries from its source row's
// region, one wave per row; col_map (non-null after the streaming gather)
// takes the stored source columns to output columns.
template<typename S>
__global__ void fl_copy_kernel(unsigned int output_rows, unsigned int cap, const unsigned int *__restrict__ ptr,
                               const unsigned int *__restrict__ src_row, const unsigned int *__restrict__ t_idx,
                               const S *__restrict__ t_val, const unsigned int *__restrict__ col_map, unsigned int *__restrict__ out_idx,
                               S *__restrict__ out_val) {
    const unsigned int lane = lane_id();
    const unsigned int nw = blockDim.x >> 6;
    for (unsigned int o = blockIdx.x * nw + (threadIdx.x >> 6); o < output_rows; o += gridDim.x * nw) {
        const unsigned int b = ptr[o], n = ptr[o + 1u] - b;
        if (n == 0u) continue;
        const unsigned long long base = (unsigned long long) src_row[o] * cap;
        for (unsigned int k = lane; k < n; k += 64u) {
            const unsigned int c = gmem(t_idx)[base + k];
            gmem(out_idx)[b + k] = col_map != nullptr ? gmem(col_map)[c] : c;  // streaming gather: source columns
            gmem(out_val)[b + k] = gmem(t_val)[base + k];
        }
    }
}

// Workspace-owned device memory (grow-only).
struct fl_private {
    void *ptr_block;     // counts [output_rows] | claim [output_rows] | dup, kept rows / slots | ptr [output_rows + 1] | col map | scan scratch
    size_t ptr_bytes;
    void *val_block;     // minor idx [max_nnz] | values [max_nnz]
    size_t val_bytes;
    void *bucket_block;  // plan arrays, rebuilt ptr, offsets, sort / scan scratch
    size_t bucket_bytes;
    void *rebuilt_block; // rebuilt minor idx | values
    size_t rebuilt_bytes;
    void *bell_block;    // repacked Blocked-ELL: blockColIdx | values
    size_t bell_bytes;
    void *tmp_block;     // gather pass: per-row regions of surviving indices | values
    size_t tmp_bytes;
    csamd_convert_workspace cws;  // scratch of the Blocked-ELL build
    int cws_ready;
};

void free_private(void *p) {
    if (p == nullptr) return;
    fl_private *st = (fl_private *) p;
    void *blocks[6] = {st->ptr_block, st->val_block, st->bucket_block, st->rebuilt_block, st->bell_block, st->tmp_block};
    for (void *b : blocks)
        if (b != nullptr) (void) hipFree(b);
    if (st->cws_ready) {
        st->cws.stream = nullptr;  // the stream belongs to the caller's workspace
        st->cws.owns_stream = 0;
        csamd_convert_clear(&st->cws);
    }
    std::free(st);
}

// kept[0] += rows kept (keep null: all), kept[1] += ell_cols per kept row.
__global__ void fl_kept_kernel(unsigned int rows, const unsigned char *__restrict__ keep_rows, unsigned int ell_cols,
                               unsigned long long *__restrict__ kept) {
    unsigned int n = 0u;
    for (unsigned int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x)
        n += keep_rows == nullptr || keep_rows[r] != 0u ? 1u : 0u;
    n = wave_sum_u32(n);
    if (lane_id() == 0u && n != 0u) {
        atomicAdd(&kept[0], (unsigned long long) n);
        atomicAdd(&kept[1], (unsigned long long) n * ell_cols);
    }
}

int grow(void **p, size_t *have, size_t need) {
    if (need <= *have && *p != nullptr) return 1;
    if (*p != nullptr) (void) hipFree(*p);
    *p = nullptr;
    *have = 0u;
    CSAMD_HIP(hipMalloc(p, need == 0u ? 256u : need), "hipMalloc filtered blocked ell workspace");
    *have = need == 0u ? 256u : need;
    return 1;
}

template<typename S> struct fl_launch {
    static int run(const csamd_blocked_ell_view *src, const fl_masks &m, unsigned int *counts, unsigned int *claim, unsigned int *dup,
                   const unsigned int *ptr, unsigned int *out_idx, void *out_val, bool emit, hipStream_t s) {
        const bool grouped = ((unsigned long long) src->ell_cols * sizeof(S)) % 16ull == 0ull && ((uintptr_t) src->val & 15u) == 0u;
        const unsigned int grid = grid_for(src->rows, grouped ? 1u : FL_WAVES, (unsigned int) device_cu_count() * 16u);
        const unsigned int width = src->ell_cols / src->block_size;
        const unsigned int cap = grouped && width <= FL_LDS_SLOTS ? width : 0u;  // slot table staged in LDS
        const size_t lds = (size_t) cap * 4u;
        if (!emit) {
            fl_count_kernel<S, 4u><<<grid, FL_THREADS, lds, s>>>(*src, (const S *) src->val, m, grouped, cap, counts, claim, dup);
            return hip_ok(hipGetLastError(), "fl_count_kernel");
        }
        fl_emit_kernel<S, 4u><<<grid, FL_THREADS, lds, s>>>(*src, (const S *) src->val, m, grouped, cap, ptr, out_idx, (S *) out_val);
        return hip_ok(hipGetLastError(), "fl_emit_kernel");
    }
};

// Gather / copy launches (rows whose value rows are whole 16-byte groups).
template<typename S> struct fl_gather_launch {
    static int run(const csamd_blocked_ell_view *src, const fl_masks &m, unsigned int row0, unsigned int cap, unsigned int *counts,
                   unsigned int *claim, unsigned int *dup, unsigned int *src_row, unsigned int *overflow, unsigned int *t_idx,
                   void *t_val, bool stream, hipStream_t s) {
        const unsigned int grid = grid_for(src->rows, 1u, (unsigned int) device_cu_count() * 16u);
        const unsigned int width = src->ell_cols / src->block_size;
        if (stream) {
            const unsigned int cpb = src->block_size / (16u / (unsigned int) sizeof(S));
            const unsigned int cpb_log2 = (unsigned int) __builtin_ctz(cpb);
            const size_t lds = ((size_t) ((m.keep_blocks + 3u) & ~3u) + width) * 4u;
            fl_gather_stream_kernel<S><<<grid, FL_THREADS, lds, s>>>(*src, (const S *) src->val, m, cpb_log2, row0, cap, counts, claim,
                                                                    dup, src_row, overflow, t_idx, (S *) t_val);
            return hip_ok(hipGetLastError(), "fl_gather_stream_kernel");
        }
        const unsigned int slot_cap = width <= FL_LDS_SLOTS ? width : 0u;
        fl_gather_kernel<S, 4u><<<grid, FL_THREADS, (size_t) slot_cap * 4u, s>>>(*src, (const S *) src->val, m, slot_cap, row0, cap,
                                                                                counts, claim, dup, src_row, overflow, t_idx, (S *) t_val);
        return hip_ok(hipGetLastError(), "fl_gather_kernel");
    }
};
template<typename S> struct fl_copy_launch {
    static int run(unsigned int output_rows, unsigned int cap, const unsigned int *ptr, const unsigned int *src_row,
                   const unsigned int *t_idx, const void *t_val, const unsigned int *col_map, unsigned int *out_idx, void *out_val,
                   hipStream_t s) {
        fl_copy_kernel<S><<<grid_for(output_rows, FL_WAVES, (unsigned int) device_cu_count() * 16u), FL_THREADS, 0, s>>>(
            output_rows, cap, ptr, src_row, t_idx, (const S *) t_val, col_map, out_idx, (S *) out_val);
        return hip_ok(hipGetLastError(), "fl_copy_kernel");
    }
};

template<template<typename> class K, typename... Args>
int fl_dispatch(unsigned int vt, Args... args) {
    switch (value_type_bytes(vt)) {
        case 1: return K<uint8_t>::run(args...);
        case 2: return K<uint16_t>::run(args...);
        case 4: return K<float>::run(args...);
        default: set_error("filter_blocked_ell_to_compressed: unsupported value type %u", vt); return 0;
    }
}

// Filtered CSR over consecutive partitions (partition p holds source rows
// [row0_p, row0_p + parts[p].rows); keep_rows / row_remap index those
// source rows).  Sets out->filtered and, in *kept, the kept row count and
// the kept rows' ell slots.
int filter_parts(const csamd_blocked_ell_view *parts, unsigned int part_count, unsigned int value_type, unsigned int output_rows,
                 unsigned int output_cols, const unsigned char *d_keep_rows, const unsigned char *d_keep_cols,
                 const unsigned int *d_row_remap, const unsigned int *d_col_remap, int device, hipStream_t s, void **priv,
                 csamd_compressed_view *filtered, unsigned long long kept[2], const char *who) {
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "%s: unsupported value type %u", who, value_type);
    CSAMD_REQUIRE(part_count == 0u || parts != nullptr, "%s: null partitions", who);
    unsigned long long max_nnz = 0ull, cols = 0ull;
    for (unsigned int p = 0; p < part_count; ++p) {
        const csamd_blocked_ell_view &v = parts[p];
        CSAMD_REQUIRE(v.rows == 0u || v.block_size != 0u, "%s: block_size is 0", who);
        CSAMD_REQUIRE(v.block_size == 0u || v.ell_cols % v.block_size == 0u, "%s: ell_cols not a multiple of block_size", who);
        CSAMD_REQUIRE(v.rows == 0u || v.ell_cols == 0u || (v.blockColIdx != nullptr && v.val != nullptr), "%s: source arrays missing",
                      who);
        max_nnz += v.nnz;  // a kept entry is a nonzero stored value
        cols = v.cols > cols ? v.cols : cols;
    }
    CSAMD_REQUIRE(max_nnz <= 0xffffffffull, "%s: %llu stored values exceed u32", who, max_nnz);
    CSAMD_HIP(hipSetDevice(device >= 0 ? device : 0), "hipSetDevice filtered blocked ell");
    if (*priv == nullptr) {
        *priv = std::calloc(1u, sizeof(fl_private));
        CSAMD_REQUIRE(*priv != nullptr, "%s: host allocation failed", who);
    }
    fl_private *st = (fl_private *) *priv;
    const size_t off_claim = align_up((size_t) output_rows * 4u, 256u);
    const size_t off_dup = off_claim + align_up((size_t) output_rows * 4u, 256u);
    const size_t off_ptr = off_dup + 256u;
    const size_t off_map = off_ptr + align_up(((size_t) output_rows + 1u) * 4u, 256u);
    const size_t off_keep = off_map + align_up((size_t) cols * 4u, 256u);
    const size_t off_src = off_keep + align_up((size_t) cols * 4u, 256u);
    const size_t off_scan = off_src + align_up((size_t) output_rows * 4u, 256u);
    if (!grow(&st->ptr_block, &st->ptr_bytes, off_scan + exclusive_scan_scratch_bytes(output_rows))) return 0;
    const size_t off_val = align_up((size_t) max_nnz * 4u, 256u);
    if (!grow(&st->val_block, &st->val_bytes, off_val + (size_t) max_nnz * (size_t) vb)) return 0;
    char *pb = (char *) st->ptr_block;
    unsigned int *counts = (unsigned int *) pb;
    unsigned int *claim = (unsigned int *) (pb + off_claim);
    unsigned int *dup = (unsigned int *) (pb + off_dup);
    unsigned int *overflow = (unsigned int *) (pb + off_dup + 4u);  // read back with dup
    unsigned long long *d_kept = (unsigned long long *) (pb + off_dup + 64u);
    unsigned int *src_row = (unsigned int *) (pb + off_src);
    unsigned int *ptr = (unsigned int *) (pb + off_ptr);
    unsigned int *out_idx = (unsigned int *) st->val_block;
    void *out_val = (char *) st->val_block + off_val;
    std::memset(filtered, 0, sizeof(*filtered));
    filtered->rows = output_rows;
    filtered->cols = output_cols;
    filtered->axis = CSAMD_COMPRESSED_BY_ROW;
    filtered->majorPtr = ptr;
    filtered->minorIdx = out_idx;
    filtered->val = (csamd_half *) out_val;
    if (!zero_bytes(pb, off_ptr, s)) return 0;  // counts, claims, dup flag, kept totals
    // kept rows and their ell slots (counted whatever the output range)
    unsigned int row0 = 0u;
    for (unsigned int p = 0; p < part_count; ++p) {
        if (parts[p].rows != 0u) {
            fl_kept_kernel<<<grid_for(parts[p].rows, 256u, 1024u), 256, 0, s>>>(
                parts[p].rows, d_keep_rows != nullptr ? d_keep_rows + row0 : nullptr, parts[p].ell_cols, d_kept);
            CSAMD_HIP(hipGetLastError(), "fl_kept_kernel");
        }
        row0 += parts[p].rows;
    }
    unsigned int live = 0u;
    if (output_rows != 0u && output_cols != 0u) {
        unsigned int *col_map = (unsigned int *) (pb + off_map);
        if (cols != 0u) {
            const fl_masks mc{output_rows, output_cols, 0u, nullptr, d_keep_cols, nullptr, d_col_remap, col_map};
            fl_col_map_kernel<<<grid_for(cols, 256u, 1024u), 256, 0, s>>>((unsigned int) cols, mc, col_map);
            CSAMD_HIP(hipGetLastError(), "fl_col_map_kernel");
        }
        unsigned int *block_keep = (unsigned int *) (pb + off_keep);
        unsigned int keep_bs = 0u;  // block size block_keep currently holds (stream order keeps reuse safe)
        // One read of the kept rows instead of two (count, then emit): when
        // every part's value rows are whole aligned 16-byte groups, a row's
        // survivors go to a region of cap = min(ell_cols, output_cols)
        // entries (enough while a row's columns are distinct and the column
        // remap is one-to-one; a row with more falls back to count + emit),
        // and a copy pass packs the regions once the row pointers are known.
        unsigned long long rows_all = 0ull, max_ell = 0ull;
        bool aligned = true;
        for (unsigned int p = 0; p < part_count; ++p) {
            const csamd_blocked_ell_view &v = parts[p];
            rows_all += v.rows;
            max_ell = v.ell_cols > max_ell ? v.ell_cols : max_ell;
            if (v.rows != 0u && v.ell_cols != 0u)
                aligned = aligned && ((unsigned long long) v.ell_cols * (unsigned long long) vb) % 16ull == 0ull &&
                          ((uintptr_t) v.val & 15u) == 0u;
        }
        const unsigned long long cap = max_ell < output_cols ? max_ell : output_cols;
        const unsigned long long tmp_need = rows_all * cap * (4ull + (unsigned long long) vb);
        const unsigned long long tmp_budget = 2ull * max_nnz * (4ull + (unsigned long long) vb) > (256ull << 20)
                                                  ? 2ull * max_nnz * (4ull + (unsigned long long) vb) : (256ull << 20);
        const bool gather = aligned && cap != 0u && tmp_need <= tmp_budget && rows_all <= 0xffffffffull;
        unsigned int *t_idx = nullptr;
        void *t_val = nullptr;
        if (gather) {
            const size_t off_tv = align_up((size_t) (rows_all * cap) * 4u, 256u);
            if (!grow(&st->tmp_block, &st->tmp_bytes, off_tv + (size_t) (rows_all * cap) * (size_t) vb)) return 0;
            t_idx = (unsigned int *) st->tmp_block;
            t_val = (char *) st->tmp_block + off_tv;
        }
        // The streaming gather applies when every non-empty part has block
        // keep bits, a power-of-two number of 16-byte groups per block and
        // LDS room for its keep bits and a row block's block columns; it
        // stores source columns, which the copy pass maps.
        bool stream = gather;
        for (unsigned int p = 0; p < part_count && stream; ++p) {
            const csamd_blocked_ell_view &v = parts[p];
            if (v.rows == 0u || v.ell_cols == 0u) continue;
            const bool keep_bits = cols != 0u && v.cols == cols && v.block_size != 0u && v.block_size <= 32u &&
                                   v.block_size % (16u / (unsigned int) vb) == 0u;
            const unsigned int cpb = keep_bits ? v.block_size / (16u / (unsigned int) vb) : 0u;
            const unsigned long long kb = (cols + v.block_size - 1u) / (v.block_size != 0u ? v.block_size : 1u);
            stream = keep_bits && (cpb & (cpb - 1u)) == 0u && (kb + 4ull + v.ell_cols / v.block_size) * 4ull <= 48ull * 1024ull;
        }
        // mode 0: count, 1: emit, 2: gather
        auto pass = [&](int mode) -> int {
            unsigned int r0 = 0u;
            for (unsigned int p = 0; p < part_count; ++p) {
                const csamd_blocked_ell_view &v = parts[p];
                const bool keep_bits = cols != 0u && v.cols == cols && v.block_size != 0u && v.block_size <= 32u &&
                                       v.block_size % (16u / (unsigned int) vb) == 0u;
                if (keep_bits && v.rows != 0u && v.ell_cols != 0u && v.block_size != keep_bs) {
                    fl_block_keep_kernel<<<grid_for((cols + v.block_size - 1u) / v.block_size, 256u, 1024u), 256, 0, s>>>(
                        (unsigned int) cols, v.block_size, col_map, block_keep);
                    CSAMD_HIP(hipGetLastError(), "fl_block_keep_kernel");
                    keep_bs = v.block_size;
                }
                const fl_masks m{output_rows, output_cols, r0, d_keep_rows != nullptr ? d_keep_rows + r0 : nullptr, d_keep_cols,
                                 d_row_remap != nullptr ? d_row_remap + r0 : nullptr, d_col_remap, col_map,
                                 keep_bits ? block_keep : nullptr,
                                 keep_bits ? (unsigned int) ((cols + v.block_size - 1u) / v.block_size) : 0u};
                if (v.rows != 0u && v.ell_cols != 0u) {
                    const int ok = mode == 2 ? fl_dispatch<fl_gather_launch>(value_type, &v, m, r0, (unsigned int) cap, counts, claim, dup,
                                                                             src_row, overflow, t_idx, t_val, stream, s)
                                             : fl_dispatch<fl_launch>(value_type, &v, m, counts, claim, dup, (const unsigned int *) ptr,
                                                                      out_idx, out_val, mode == 1, s);
                    if (!ok) return 0;
                }
                r0 += v.rows;
            }
            return 1;
        };
        // counts -> row pointers; the live count, duplicate-row and overflow flags to the host
        auto finish_counts = [&](unsigned int h[3]) -> int {
            if (!exclusive_scan_u32(counts, output_rows, ptr, nullptr, pb + off_scan, exclusive_scan_scratch_bytes(output_rows), s))
                return 0;
            CSAMD_HIP(hipMemcpyAsync(&h[0], ptr + output_rows, 4u, hipMemcpyDeviceToHost, s), "D2H filtered nnz");
            CSAMD_HIP(hipMemcpyAsync(&h[1], dup, 8u, hipMemcpyDeviceToHost, s), "D2H filtered flags");
            CSAMD_HIP(hipStreamSynchronize(s), "filtered blocked ell sync");
            CSAMD_REQUIRE(h[1] == 0u, "%s: row_remap maps two kept rows to output row %u", who, h[1] - 1u);
            return 1;
        };
        unsigned int h[3] = {0u, 0u, 0u};
        bool done = false;
        if (gather) {
            if (!pass(2) || !finish_counts(h)) return 0;
            if (h[2] == 0u) {  // every row fit its region
                live = h[0];
                CSAMD_REQUIRE(live <= max_nnz, "%s: %u surviving entries exceed the source nnz %llu", who, live, max_nnz);
                if (!fl_dispatch<fl_copy_launch>(value_type, output_rows, (unsigned int) cap, (const unsigned int *) ptr,
                                                 (const unsigned int *) src_row, (const unsigned int *) t_idx, (const void *) t_val,
                                                 stream ? (const unsigned int *) col_map : nullptr,
                                                 out_idx, out_val, s))
                    return 0;
                done = true;
            } else if (!zero_bytes(pb, off_dup + 64u, s)) {  // counts, claims and flags again (not the kept totals) for count + emit
                return 0;
            }
        }
        if (!done) {
            if (!pass(0) || !finish_counts(h)) return 0;
            live = h[0];
            CSAMD_REQUIRE(live <= max_nnz, "%s: %u surviving entries exceed the source nnz %llu", who, live, max_nnz);
            if (!pass(1)) return 0;
        }
    } else if (output_rows != 0u) {
        if (!zero_bytes(ptr, ((size_t) output_rows + 1u) * 4u, s)) return 0;
    } else {
        if (!zero_bytes(ptr, 4u, s)) return 0;
    }
    filtered->nnz = live;
    CSAMD_HIP(hipMemcpyAsync(kept, d_kept, 16u, hipMemcpyDeviceToHost, s), "D2H kept rows");
    CSAMD_HIP(hipStreamSynchronize(s), "filtered blocked ell sync");
    return 1;
}

// bucket::build_bucketed_major_view over the filtered CSR: plan (stable
// ascending nnz order, equal-count offsets) and rebuild into workspace memory.
int bucket_view(fl_private *st, const csamd_compressed_view &filtered, unsigned int value_type, unsigned int requested,
                hipStream_t s, csamd_compressed_major_bucket_result *out) {
    const unsigned int n = filtered.rows;
    const unsigned int bc = csamd_clamp_bucket_count(n, requested);
    size_t sort_b = 0u, scan_b = 0u;
    if (!csamd_major_nnz_bucket_sort_scratch_bytes(n, &sort_b) || !csamd_major_nnz_bucket_scan_scratch_bytes(n, &scan_b)) return 0;
    const size_t a = align_up((size_t) n * 4u, 256u);
    const size_t off_sorted = a, off_in = 2u * a, off_out = 3u * a, off_inv = 4u * a, off_ptr = 5u * a;
    const size_t off_offsets = off_ptr + align_up(((size_t) n + 1u) * 4u, 256u);
    const size_t off_sort = off_offsets + align_up(((size_t) bc + 1u) * 4u, 256u);
    const size_t off_scan = off_sort + align_up(sort_b, 256u);
    if (!grow(&st->bucket_block, &st->bucket_bytes, off_scan + scan_b)) return 0;
    const int vb = value_type_bytes(value_type);
    const size_t nnz = filtered.nnz;
    const size_t off_rval = align_up(nnz * 4u, 256u);
    if (!grow(&st->rebuilt_block, &st->rebuilt_bytes, off_rval + nnz * (size_t) vb)) return 0;
    char *b = (char *) st->bucket_block;
    unsigned int *major_nnz = (unsigned int *) b;
    unsigned int *sorted = (unsigned int *) (b + off_sorted);
    unsigned int *order_in = (unsigned int *) (b + off_in);
    unsigned int *order_out = (unsigned int *) (b + off_out);
    unsigned int *inverse = (unsigned int *) (b + off_inv);
    unsigned int *rptr = (unsigned int *) (b + off_ptr);
    unsigned int *offsets = (unsigned int *) (b + off_offsets);
    unsigned int *ridx = (unsigned int *) st->rebuilt_block;
    void *rval = (char *) st->rebuilt_block + off_rval;
    if (!csamd_build_major_nnz_bucket_plan_raw(filtered.majorPtr, n, major_nnz, sorted, order_in, order_out, offsets, bc,
                                               b + off_sort, sort_b, (csamd_stream_t) s))
        return 0;
    if (!csamd_rebuild_compressed_major_order_raw(filtered.majorPtr, filtered.minorIdx, filtered.val, value_type, n, order_out, rptr,
                                                  ridx, rval, inverse, b + off_scan, scan_b, (csamd_stream_t) s))
        return 0;
    if (n == 0u && !zero_bytes(rptr, 4u, s)) return 0;
    out->rebuilt = filtered;
    out->rebuilt.majorPtr = rptr;
    out->rebuilt.minorIdx = ridx;
    out->rebuilt.val = (csamd_half *) rval;
    out->plan.major_dim = n;
    out->plan.bucket_count = bc;
    out->plan.major_order = order_out;
    out->plan.major_nnz_sorted = sorted;
    out->plan.bucket_offsets = offsets;
    out->plan.inverse_major_order = inverse;
    return 1;
}

void fill_stats(csamd_filtered_blocked_ell_stats *st, unsigned int kept_rows, unsigned int output_rows, unsigned int output_cols,
                unsigned int live, unsigned long long slots, int vb) {
    st->kept_rows = kept_rows;
    st->output_rows = output_rows;
    st->output_cols = output_cols;
    st->live_nnz = live;
    st->kept_row_slots = slots;
    st->dead_slots = slots > live ? slots - live : 0ull;
    st->live_fill_ratio = slots == 0ull ? 1.0 : (double) live / (double) slots;
    st->dead_value_bytes = (size_t) st->dead_slots * (size_t) vb;
}

int filter_impl(const csamd_blocked_ell_view *src, unsigned int value_type, unsigned int output_rows, unsigned int output_cols,
                const unsigned char *d_keep_rows, const unsigned char *d_keep_cols, const unsigned int *d_row_remap,
                const unsigned int *d_col_remap, csamd_filtered_blocked_ell_workspace *ws, csamd_filtered_blocked_ell_result *out) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr, "filter_blocked_ell_to_compressed: null argument");
    csamd_filtered_blocked_ell_result local;
    std::memset(&local, 0, sizeof(local));
    unsigned long long kept[2] = {0ull, 0ull};
    if (!filter_parts(src, 1u, value_type, output_rows, output_cols, d_keep_rows, d_keep_cols, d_row_remap, d_col_remap, ws->device,
                      (hipStream_t) ws->stream, &ws->private_workspace, &local.filtered, kept, "filter_blocked_ell_to_compressed"))
        return 0;
    // the reference reports kept_rows = output_rows and prices kept_rows x ell_cols slots (:359-370)
    const bool any = output_rows != 0u && output_cols != 0u;
    fill_stats(&local.stats, output_rows, output_rows, output_cols, local.filtered.nnz,
               any ? (unsigned long long) output_rows * src->ell_cols : 0ull, value_type_bytes(value_type));
    if (out != nullptr) *out = local;
    return 1;
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_filtered_blocked_ell_init(csamd_filtered_blocked_ell_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_filtered_blocked_ell_clear(csamd_filtered_blocked_ell_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    free_private(ws->private_workspace);
    if (ws->owns_stream && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    csamd_filtered_blocked_ell_init(ws);
}

int csamd_filtered_blocked_ell_setup(csamd_filtered_blocked_ell_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "filtered blocked ell setup: null workspace");
    csamd_filtered_blocked_ell_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice filtered blocked ell setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags filtered blocked ell");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    return 1;
}

int csamd_filter_blocked_ell_to_compressed(const csamd_blocked_ell_view *src, unsigned int value_type, unsigned int output_rows,
                                           unsigned int output_cols, const unsigned char *d_keep_rows,
                                           const unsigned char *d_keep_cols, const unsigned int *d_row_remap,
                                           const unsigned int *d_col_remap, csamd_filtered_blocked_ell_workspace *ws,
                                           csamd_filtered_blocked_ell_result *out) {
    return filter_impl(src, value_type, output_rows, output_cols, d_keep_rows, d_keep_cols, d_row_remap, d_col_remap, ws, out);
}

int csamd_build_bucketed_filtered_blocked_ell_major_view(const csamd_blocked_ell_view *src, unsigned int value_type,
                                                         unsigned int output_rows, unsigned int output_cols,
                                                         const unsigned char *d_keep_rows, const unsigned char *d_keep_cols,
                                                         const unsigned int *d_row_remap, const unsigned int *d_col_remap,
                                                         unsigned int requested_bucket_count,
                                                         csamd_filtered_blocked_ell_workspace *ws,
                                                         csamd_filtered_blocked_ell_result *out) {
    csamd_filtered_blocked_ell_result local;
    if (!filter_impl(src, value_type, output_rows, output_cols, d_keep_rows, d_keep_cols, d_row_remap, d_col_remap, ws, &local))
        return 0;
    if (!bucket_view((fl_private *) ws->private_workspace, local.filtered, value_type, requested_bucket_count, (hipStream_t) ws->stream,
                     &local.bucketed))
        return 0;
    local.has_bucketed = 1;
    if (out != nullptr) *out = local;
    return 1;
}

void csamd_repack_init(csamd_sharded_blocked_ell_repack_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_repack_clear(csamd_sharded_blocked_ell_repack_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    free_private(ws->private_workspace);
    if (ws->owns_stream && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    csamd_repack_init(ws);
}

int csamd_repack_setup(csamd_sharded_blocked_ell_repack_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "repack setup: null workspace");
    csamd_repack_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice repack setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags repack");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    return 1;
}

int csamd_build_repacked_blocked_ell(const csamd_blocked_ell_view *parts, unsigned int part_count, unsigned int value_type,
                                     const csamd_sharded_blocked_ell_repack_config *cfg, const unsigned char *d_keep_rows,
                                     const unsigned char *d_keep_cols, const unsigned int *d_row_remap,
                                     const unsigned int *d_col_remap, unsigned long shard_id,
                                     csamd_sharded_blocked_ell_repack_workspace *ws, csamd_sharded_blocked_ell_repack_result *out) {
    const char *who = "build_repacked_shard_blocked_ell";
    CSAMD_REQUIRE(cfg != nullptr && ws != nullptr && out != nullptr, "%s: null argument", who);
    CSAMD_REQUIRE(cfg->block_size != 0u, "%s: block_size is 0", who);
    hipStream_t s = (hipStream_t) ws->stream;
    csamd_sharded_blocked_ell_repack_result local;
    std::memset(&local, 0, sizeof(local));
    unsigned long long kept[2] = {0ull, 0ull};
    if (!filter_parts(parts, part_count, value_type, cfg->output_rows, cfg->output_cols, d_keep_rows, d_keep_cols, d_row_remap,
                      d_col_remap, ws->device, s, &ws->private_workspace, &local.filtered, kept, who))
        return 0;
    fl_private *st = (fl_private *) ws->private_workspace;
    const unsigned int live = local.filtered.nnz;
    const int vb = value_type_bytes(value_type);
    local.shard_id = shard_id;
    // kept rows counted over the shard, slots priced at each kept row's own ell width
    // (repack/kernels/sharded_blocked_ell.cuh:43-51)
    csamd_filtered_blocked_ell_stats fs;
    fill_stats(&fs, (unsigned int) kept[0], cfg->output_rows, cfg->output_cols, live, kept[1], vb);
    local.stats.kept_rows = fs.kept_rows;
    local.stats.output_rows = fs.output_rows;
    local.stats.output_cols = fs.output_cols;
    local.stats.live_nnz = fs.live_nnz;
    local.stats.kept_row_slots = fs.kept_row_slots;
    local.stats.dead_slots = fs.dead_slots;
    local.stats.live_fill_ratio = fs.live_fill_ratio;
    local.stats.dead_value_bytes = fs.dead_value_bytes;
    local.stats.ell_width = 0u;
    local.rebuilt.rows = cfg->output_rows;
    local.rebuilt.cols = cfg->output_cols;
    local.rebuilt.nnz = live;
    local.rebuilt.block_size = cfg->block_size;
    if (cfg->output_rows == 0u || cfg->output_cols == 0u || live == 0u) {
        // nothing to bucket or pack (:541-560)
        local.bucketed.rebuilt = local.filtered;
        local.bucketed.plan.major_dim = cfg->output_rows;
        local.bucketed.plan.bucket_count = 1u;
        local.bucket_plan = local.bucketed.plan;
        *out = local;
        return 1;
    }
    if (!bucket_view(st, local.filtered, value_type, cfg->requested_bucket_count, s, &local.bucketed)) return 0;
    local.bucket_plan = local.bucketed.plan;
    // the bucketed rows repacked as Blocked-ELL of the requested block size
    if (!st->cws_ready) {
        csamd_convert_init(&st->cws);
        if (!csamd_convert_setup(&st->cws, ws->device >= 0 ? ws->device : 0, ws->stream)) return 0;
        st->cws_ready = 1;
    }
    st->cws.stream = ws->stream;
    const unsigned int bs = cfg->block_size;
    unsigned int width = 0u;
    unsigned long long total = 0ull;
    double fill = 0.0;
    if (!csamd_blocked_ell_shape(&local.bucketed.rebuilt, bs, &st->cws, &width, &total, &fill)) return 0;
    const unsigned long long row_blocks = (cfg->output_rows + (unsigned long long) bs - 1ull) / bs;
    const unsigned long long ell_cols = (unsigned long long) width * bs;
    CSAMD_REQUIRE(ell_cols <= 0xffffffffull, "%s: ell width %u x block %u exceeds u32", who, width, bs);
    const size_t off_v = align_up((size_t) (row_blocks * width) * 4u, 256u);
    if (!grow(&st->bell_block, &st->bell_bytes, off_v + (size_t) cfg->output_rows * (size_t) ell_cols * (size_t) vb)) return 0;
    csamd_blocked_ell_view bell;
    std::memset(&bell, 0, sizeof(bell));
    bell.block_size = bs;
    bell.ell_cols = (unsigned int) ell_cols;
    bell.blockColIdx = (unsigned int *) st->bell_block;
    bell.val = (csamd_half *) ((char *) st->bell_block + off_v);
    if (!csamd_blocked_ell_from_compressed(&local.bucketed.rebuilt, value_type, &bell, &st->cws)) return 0;
    local.rebuilt = bell;
    local.rebuilt.nnz = live;  // as the reference reports it (:714)
    local.stats.ell_width = width;
    *out = local;
    return 1;
}

} // extern "C"
