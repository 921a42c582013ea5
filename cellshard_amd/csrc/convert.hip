// CSR -> Blocked-ELL (shape, tuner, emit), CSR/CSC -> Sliced-ELL (layout,
// tuner, emit), Blocked-ELL -> Sliced-ELL.
//
// Layout rules are the reference's (src/convert/blocked_ell_from_compressed.cuh,
// src/convert/sliced_ell_from_compressed.cuh, src/convert/sliced_ell_from_blocked_ell.cuh):
//   Blocked-ELL: per row block, the unique column blocks touched by any stored
//   entry, ascending, fill the first slots; W = max over row blocks; unused
//   slots hold 0xffffffff; value (r, c) sits at r*ell_cols + slot*B + c%B;
//   nnz = count of nonzero stored values (a duplicate (r, c) keeps the last).
//   Sliced-ELL: uniform slices of S = clamp(S, 1, rows) rows; width = max row
//   nnz in the slice; row slots follow CSR order; padding col 0xffffffff, 0.
//
// gfx950 mapping: one workgroup per row block with an LDS bitset over column
// blocks (popcount ranks give the ascending slot of every block column);
// one wave64 per row for Sliced-ELL emits (coalesced copy + padding).
#include "common.hh"
#include "sell_write.hh"
#include "stream.hh"

#include <cstdlib>

namespace csamd {
namespace {

constexpr unsigned int CV_THREADS = 256u;
constexpr unsigned int BELL_MAX_BITSET_WORDS = 8192u;  // 262,144 column blocks (32 KB LDS)
constexpr unsigned int BELL_MAX_BLOCK_SIZE = 8192u;    // rows of a block (their row pointers sit in LDS)

// Marks in `bits` the column blocks touched by rows [r0, r1) (entries
// [rp[0], rp[r1 - r0]), row pointers in LDS), four loads in flight per
// thread.  CHECK: also sets *unsorted when some row is not strictly
// ascending (each thread walks its entries' rows forward).
template<bool CHECK>
__device__ __forceinline__ void bell_mark(const csamd_compressed_view &src, unsigned int nrows, unsigned int bs, const unsigned int *rp,
                                          unsigned int *bits, unsigned int *unsorted) {
    const unsigned int b = rp[0], e = rp[nrows];
    unsigned int row = 0u, rend = rp[1];
    for (unsigned int i0 = b; i0 < e; i0 += 4u * CV_THREADS) {
        unsigned int c[4], n[4];
#pragma unroll
        for (unsigned int u = 0; u < 4u; ++u) {
            const unsigned int i = i0 + u * CV_THREADS + threadIdx.x;
            c[u] = i < e ? src.minorIdx[i] : CSAMD_INVALID_COL;
            if (CHECK) n[u] = i + 1u < e ? src.minorIdx[i + 1u] : CSAMD_INVALID_COL;
        }
#pragma unroll
        for (unsigned int u = 0; u < 4u; ++u) {
            const unsigned int i = i0 + u * CV_THREADS + threadIdx.x;
            if (i >= e) continue;
            if (c[u] < src.cols) {
                const unsigned int bc = c[u] / bs;
                atomicOr(&bits[bc >> 5], 1u << (bc & 31u));
            }
            if (CHECK) {
                while (i >= rend) rend = rp[++row + 1u];
                if (i + 1u < rend && c[u] >= n[u]) *unsorted = 1u;
            }
        }
    }
}

__global__ void __launch_bounds__(CV_THREADS) bell_shape_kernel(csamd_compressed_view src, unsigned int bs, unsigned int row_blocks,
                                                                unsigned int nwords, unsigned int *__restrict__ counts,
                                                                unsigned int *__restrict__ max_width,
                                                                unsigned long long *__restrict__ total) {
    extern __shared__ unsigned int bits[];
    unsigned int *rp = bits + nwords;  // [bs + 1] row pointers of the block
    __shared__ unsigned int red[CV_THREADS / 64u];
    for (unsigned int rb = blockIdx.x; rb < row_blocks; rb += gridDim.x) {
        for (unsigned int w = threadIdx.x; w < nwords; w += CV_THREADS) bits[w] = 0u;
        const unsigned int r0 = rb * bs;
        const unsigned int r1 = r0 + bs < src.rows ? r0 + bs : src.rows;
        for (unsigned int r = threadIdx.x; r <= r1 - r0; r += CV_THREADS) rp[r] = src.majorPtr[r0 + r];
        __syncthreads();
        bell_mark<false>(src, r1 - r0, bs, rp, bits, nullptr);
        __syncthreads();
        unsigned int u = 0u;
        for (unsigned int w = threadIdx.x; w < nwords; w += CV_THREADS) u += (unsigned int) __popc(bits[w]);
        u = wave_sum_u32(u);
        if (lane_id() == 0u) red[threadIdx.x >> 6] = u;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int t = 0u;
            for (unsigned int k = 0; k < CV_THREADS / 64u; ++k) t += red[k];
            counts[rb] = t;
            atomicMax(max_width, t);
            atomicAdd(total, (unsigned long long) t);
        }
        __syncthreads();
    }
}

template<typename S>
__device__ __forceinline__ bool stored_nonzero(S raw);
template<> __device__ __forceinline__ bool stored_nonzero<uint16_t>(uint16_t raw) { return (raw & 0x7fffu) != 0u; }
template<> __device__ __forceinline__ bool stored_nonzero<float>(float raw) { return raw != 0.0f; }
template<> __device__ __forceinline__ bool stored_nonzero<uint8_t>(uint8_t raw) { return (raw & 0x7fu) != 0u; }

// Blocked-ELL emit, in three launches.
//
// bell_blocks_kernel, one workgroup per row block: the set of column blocks
// the block's rows touch (LDS bitset), the block's slot of every column
// block (popcount prefix per bitset word), blockColIdx (ascending, then
// invalid padding), and whether some row is not strictly ascending.  The
// bitset and its prefix go to a table for the next launches.
__global__ void __launch_bounds__(CV_THREADS) bell_blocks_kernel(csamd_compressed_view src, unsigned int bs, unsigned int row_blocks,
                                                                 unsigned int width, unsigned int nwords,
                                                                 unsigned int *__restrict__ block_col_idx,
                                                                 unsigned int *__restrict__ tab, unsigned int *__restrict__ unsorted,
                                                                 unsigned int *__restrict__ need_w) {
    extern __shared__ unsigned int lds[];
    unsigned int *bits = lds;
    unsigned int *prefix = lds + nwords;
    unsigned int *rp = prefix + nwords;  // [bs + 1] row pointers of the block
    __shared__ unsigned int red[CV_THREADS / 64u];
    __shared__ unsigned int unsorted_flag;
    for (unsigned int rb = blockIdx.x; rb < row_blocks; rb += gridDim.x) {
        for (unsigned int w = threadIdx.x; w < nwords; w += CV_THREADS) bits[w] = 0u;
        if (threadIdx.x == 0) unsorted_flag = 0u;
        const unsigned int r0 = rb * bs;
        const unsigned int r1 = r0 + bs < src.rows ? r0 + bs : src.rows;
        for (unsigned int r = threadIdx.x; r <= r1 - r0; r += CV_THREADS) rp[r] = src.majorPtr[r0 + r];
        __syncthreads();
        // rows with non-increasing columns (possible duplicates) take the serial path
        bell_mark<true>(src, r1 - r0, bs, rp, bits, &unsorted_flag);
        __syncthreads();
        // exclusive prefix of popcounts over the bitset words (serial chunks of 256)
        {
            unsigned int carry = 0u;
            for (unsigned int w0 = 0; w0 < nwords; w0 += CV_THREADS) {
                const unsigned int w = w0 + threadIdx.x;
                const unsigned int pc = w < nwords ? (unsigned int) __popc(bits[w]) : 0u;
                unsigned int tot = 0u;
                const unsigned int ex = block_exclusive_scan_u32(pc, red, &tot);
                if (w < nwords) prefix[w] = carry + ex;
                carry += tot;
            }
        }
        __syncthreads();
        // block column ids in ascending order, then invalid padding; a block
        // needing more than `width` slots is reported (the call fails) and
        // nothing is written past its slots
        const unsigned long long slot_base = (unsigned long long) rb * width;
        const unsigned int unique = nwords != 0u ? prefix[nwords - 1u] + (unsigned int) __popc(bits[nwords - 1u]) : 0u;
        if (unique > width && threadIdx.x == 0) atomicMax(need_w, unique);
        for (unsigned int w = threadIdx.x; w < nwords; w += CV_THREADS) {
            unsigned int m = bits[w];
            unsigned int s = prefix[w];
            while (m != 0u && s < width) {
                const unsigned int bit = (unsigned int) __builtin_ctz(m);
                block_col_idx[slot_base + s] = w * 32u + bit;
                ++s;
                m &= m - 1u;
            }
        }
        for (unsigned int s = unique + threadIdx.x; s < width; s += CV_THREADS) block_col_idx[slot_base + s] = CSAMD_INVALID_COL;
        unsigned int *t = tab + (size_t) rb * 2u * nwords;
        for (unsigned int w = threadIdx.x; w < 2u * nwords; w += CV_THREADS) t[w] = lds[w];
        if (threadIdx.x == 0) unsorted[rb] = unsorted_flag;
        __syncthreads();
    }
}

// Value position of column c (< cols) in its row: slot of its column block
// times bs plus the lane.
__device__ __forceinline__ unsigned int bell_pos(const unsigned int *bits, const unsigned int *prefix, unsigned int bs, unsigned int c) {
    const unsigned int bc = c / bs;
    const unsigned int slot = prefix[bc >> 5] + (unsigned int) __popc(bits[bc >> 5] & ((1u << (bc & 31u)) - 1u));
    return slot * bs + (c - bc * bs);
}

// bell_rows_kernel, one workgroup per row: the value row is written whole
// (padding included), in segments of SEG_WORDS words staged in LDS
// (zeros, then the row's entries that fall in the segment), so every output
// line is written once by one coalesced store.  A row's entries are read
// once into registers (up to E per thread; longer rows read theirs again
// for each segment).  Rows of an unsorted row block are only zeroed here
// (bell_replay_kernel fills them).  Rows that are not whole, aligned 4-byte
// words (narrow values with an odd ell width) are written value by value.
template<typename S, unsigned int E, unsigned int SEG_WORDS>
__global__ void __launch_bounds__(CV_THREADS) bell_rows_kernel(csamd_compressed_view src, const S *__restrict__ sval, unsigned int bs,
                                                               unsigned int nwords, unsigned int ell_cols,
                                                               const unsigned int *__restrict__ tab,
                                                               const unsigned int *__restrict__ unsorted, S *__restrict__ dval,
                                                               unsigned long long *__restrict__ nnz_total) {
    extern __shared__ unsigned int lds[];
    unsigned int *bits = lds;
    unsigned int *prefix = lds + nwords;
    unsigned int *seg = lds + ((2u * nwords + 3u) & ~3u);  // [SEG_WORDS], 16-byte aligned
    S *segv = (S *) seg;
    constexpr unsigned int PER_WORD = 4u / sizeof(S);
    constexpr unsigned int SEG = SEG_WORDS * PER_WORD;  // values per segment
    constexpr unsigned int INV = 0xffffffffu;
    const bool words = ell_cols % PER_WORD == 0u && ((uintptr_t) dval & 3u) == 0u;
    const bool quads = ((unsigned long long) ell_cols * sizeof(S)) % 16ull == 0ull && ((uintptr_t) dval & 15u) == 0u;
    unsigned int nz = 0u;
    for (unsigned int r = blockIdx.x; r < src.rows; r += gridDim.x) {
        const unsigned int rb = r / bs;
        S *orow = dval + (unsigned long long) r * ell_cols;
        if (unsorted[rb] != 0u) {
            for (unsigned int k = threadIdx.x; k < ell_cols; k += CV_THREADS) orow[k] = S(0);
            continue;
        }
        const unsigned int *t = tab + (size_t) rb * 2u * nwords;
        for (unsigned int w = threadIdx.x; w < 2u * nwords; w += CV_THREADS) lds[w] = t[w];
        const unsigned int b = src.majorPtr[r], e = src.majorPtr[r + 1u];
        __syncthreads();
        const bool cached = e - b <= E * CV_THREADS;
        unsigned int pos[E];
        S v[E];
        if (cached) {
#pragma unroll
            for (unsigned int k = 0; k < E; ++k) {
                const unsigned int i = b + threadIdx.x + k * CV_THREADS;
                pos[k] = INV;
                v[k] = S(0);
                if (i < e) {
                    const unsigned int c = src.minorIdx[i];
                    if (c < src.cols) {
                        pos[k] = bell_pos(bits, prefix, bs, c);
                        v[k] = sval[i];
                        nz += stored_nonzero<S>(v[k]) ? 1u : 0u;
                    }
                }
            }
        }
        for (unsigned int p0 = 0; p0 < ell_cols; p0 += SEG) {
            const unsigned int ne = ell_cols - p0 < SEG ? ell_cols - p0 : SEG;
            const unsigned int nw = (ne + PER_WORD - 1u) / PER_WORD;
            for (unsigned int w = 4u * threadIdx.x; w < nw; w += 4u * CV_THREADS) *(uint4 *) (seg + w) = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
            if (cached) {
#pragma unroll
                for (unsigned int k = 0; k < E; ++k)
                    if (pos[k] - p0 < ne) segv[pos[k] - p0] = v[k];
            } else {
                for (unsigned int i = b + threadIdx.x; i < e; i += CV_THREADS) {
                    const unsigned int c = src.minorIdx[i];
                    if (c >= src.cols) continue;
                    const unsigned int q = bell_pos(bits, prefix, bs, c) - p0;
                    if (q >= ne) continue;
                    const S x = sval[i];
                    segv[q] = x;
                    nz += stored_nonzero<S>(x) ? 1u : 0u;
                }
            }
            __syncthreads();
            if (quads) {  // nw is a multiple of 4 here; streaming stores
                typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
                u32x4 *ow = (u32x4 *) (orow + p0);
                for (unsigned int w = threadIdx.x; 4u * w < nw; w += CV_THREADS)
                    __builtin_nontemporal_store(*(const u32x4 *) (seg + 4u * w), ow + w);
            } else if (words) {
                unsigned int *ow = (unsigned int *) (orow + p0);
                for (unsigned int w = threadIdx.x; w < nw; w += CV_THREADS) ow[w] = seg[w];
            } else {
                for (unsigned int k = threadIdx.x; k < ne; k += CV_THREADS) orow[p0 + k] = segv[k];
            }
            __syncthreads();
        }
    }
    nz = wave_sum_u32(nz);
    if (lane_id() == 0u && nz != 0u) atomicAdd(nnz_total, (unsigned long long) nz);
}

// bell_replay_kernel, one thread per unsorted row block (after
// bell_rows_kernel zeroed its rows): CSR-order replay, the last write of a
// repeated column wins, nnz tracks zero / nonzero transitions.
template<typename S>
__global__ void bell_replay_kernel(csamd_compressed_view src, const S *__restrict__ sval, unsigned int bs, unsigned int row_blocks,
                                   unsigned int nwords, unsigned int ell_cols, const unsigned int *__restrict__ tab,
                                   const unsigned int *__restrict__ unsorted, S *__restrict__ dval,
                                   unsigned long long *__restrict__ nnz_total) {
    for (unsigned int rb = blockIdx.x * blockDim.x + threadIdx.x; rb < row_blocks; rb += gridDim.x * blockDim.x) {
        if (unsorted[rb] == 0u) continue;
        const unsigned int *bits = tab + (size_t) rb * 2u * nwords;
        const unsigned int *prefix = bits + nwords;
        const unsigned int r0 = rb * bs;
        const unsigned int r1 = r0 + bs < src.rows ? r0 + bs : src.rows;
        long long count = 0;
        for (unsigned int r = r0; r < r1; ++r) {
            for (unsigned int i = src.majorPtr[r]; i < src.majorPtr[r + 1u]; ++i) {
                const unsigned int c = src.minorIdx[i];
                if (c >= src.cols) continue;
                const unsigned int q = bell_pos(bits, prefix, bs, c);
                if (q >= ell_cols) continue;  // width too small: reported by bell_blocks_kernel
                S *dst = dval + (unsigned long long) r * ell_cols + q;
                const bool prev = stored_nonzero<S>(*dst);
                const bool next = stored_nonzero<S>(sval[i]);
                if (!prev && next) ++count;
                else if (prev && !next) --count;
                *dst = sval[i];
            }
        }
        if (count != 0) atomicAdd(nnz_total, (unsigned long long) count);
    }
}

// ---------------------------------------------------------------------------
// Sliced-ELL.
// ---------------------------------------------------------------------------
__global__ void csr_row_nnz_kernel(const unsigned int *__restrict__ ptr, unsigned int rows, unsigned int *__restrict__ row_nnz) {
    for (unsigned int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) row_nnz[r] = ptr[r + 1u] - ptr[r];
}

__global__ void csc_row_nnz_kernel(const unsigned int *__restrict__ minor, unsigned int nnz, unsigned int rows,
                                   unsigned int *__restrict__ row_nnz, unsigned int *__restrict__ bad) {
    for (unsigned long long i = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; i < nnz;
         i += (unsigned long long) gridDim.x * blockDim.x) {
        const unsigned int r = minor[i];
        if (r < rows) atomicAdd(row_nnz + r, 1u);
        else *bad = 1u;
    }
}

// 16-byte chunks of a Blocked-ELL value row, when every chunk lies in one
// block (bs and ell_cols multiples of the chunk, an aligned value array):
// bit e of the mask is set when value e of chunk c is nonzero at a valid
// column < cols; *col0 is the chunk's first column.
template<typename S> union bell_chunk_vals {
    uint4 q;
    S v[16u / sizeof(S)];
};

template<typename S>
__device__ __forceinline__ bool bell_chunks_ok(const csamd_blocked_ell_view &src) {
    constexpr unsigned int PER = 16u / sizeof(S);
    return src.block_size != 0u && src.block_size % PER == 0u && src.ell_cols % PER == 0u && ((uintptr_t) src.val & 15u) == 0u;
}

template<typename S>
__device__ __forceinline__ unsigned int bell_chunk_mask(const csamd_blocked_ell_view &src, unsigned int width, unsigned int row,
                                                        unsigned int c, bell_chunk_vals<S> *g, unsigned int *col0) {
    constexpr unsigned int PER = 16u / sizeof(S);
    const unsigned int bs = src.block_size;
    const unsigned int t0 = c * PER, slot = t0 / bs;
    const unsigned int bc = src.blockColIdx[(unsigned long long) (row / bs) * width + slot];
    g->q = ((const uint4 *) ((const S *) src.val + (unsigned long long) row * src.ell_cols))[c];
    *col0 = bc * bs + (t0 - slot * bs);
    unsigned int m = 0u;
    if (bc != CSAMD_INVALID_COL) {
#pragma unroll
        for (unsigned int e = 0; e < PER; ++e)
            if (*col0 + e < src.cols && stored_nonzero<S>(g->v[e])) m |= 1u << e;
    }
    return m;
}

// Bit e set when value e of the chunk is nonzero at a valid column < cols
// (block column bc, first column col0).
template<typename S>
__device__ __forceinline__ unsigned int chunk_mask(const u32x4 &q, unsigned int bc, unsigned int col0, unsigned int cols) {
    constexpr unsigned int PER = 16u / sizeof(S);
    unsigned int m = 0u;
#pragma unroll
    for (unsigned int e = 0; e < PER; ++e)
        if (col0 + e < cols && stored_nonzero<S>(chunk_val<S>(q, e))) m |= 1u << e;
    return bc != CSAMD_INVALID_COL ? m : 0u;
}

// log2 of the 16-byte chunks per block when the streaming walk applies
// (chunked layout, a power-of-two number of chunks per block), else -1.
template<typename S>
__device__ __forceinline__ int bell_stream_log2(const csamd_blocked_ell_view &src) {
    if (!bell_chunks_ok<S>(src)) return -1;
    const unsigned int cpb = src.block_size / (16u / (unsigned int) sizeof(S));
    return (cpb & (cpb - 1u)) == 0u ? __ffs((int) cpb) - 1 : -1;
}

constexpr unsigned int BELL_STREAM_U = 4u;  // chunks per lane per streaming step

// Blocked-ELL row nnz: nonzero values at valid columns (sliced_ell_from_blocked_ell.cuh:36-62).
template<typename S>
__global__ void bell_row_nnz_kernel(csamd_blocked_ell_view src, unsigned int *__restrict__ row_nnz) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int nw = blockDim.x >> 6;
    const unsigned int bs = src.block_size;
    const unsigned int width = bs != 0u ? src.ell_cols / bs : 0u;
    const S *val = (const S *) src.val;
    const int slog = bell_stream_log2<S>(src);
    for (unsigned int row = blockIdx.x * nw + wave; row < src.rows; row += gridDim.x * nw) {
        const unsigned int rb = row / bs;
        unsigned int count = 0u;
        if (slog >= 0) {
            constexpr unsigned int PER = 16u / sizeof(S);
            const unsigned int nc = src.ell_cols / PER;
            const unsigned int cmask = (1u << slog) - 1u;
            bell_stream_row<S, BELL_STREAM_U>(src, width, (unsigned int) slog, row, [&](unsigned int c, const u32x4 &q, unsigned int bc) {
                const unsigned int col0 = bc * bs + (c & cmask) * PER;
                if (c < nc) count += (unsigned int) __popc(chunk_mask<S>(q, bc, col0, src.cols));
            });
            count = wave_sum_u32(count);
            if (lane_id() == 0u) row_nnz[row] = count;
            continue;
        }
        for (unsigned int e = lane_id(); e < width * bs; e += 64u) {
            const unsigned int slot = e / bs;
            const unsigned int bc = src.blockColIdx[(unsigned long long) rb * width + slot];
            if (bc == CSAMD_INVALID_COL) continue;
            const unsigned int col = bc * bs + (e - slot * bs);
            if (col >= src.cols) continue;
            if (stored_nonzero<S>(val[(unsigned long long) row * src.ell_cols + e])) ++count;
        }
        count = wave_sum_u32(count);
        if (lane_id() == 0u) row_nnz[row] = count;
    }
}

// Per slice: width = max row nnz, slot count = rows_in_slice * width.
__global__ void sell_slices_kernel(const unsigned int *__restrict__ row_nnz, unsigned int rows, unsigned int S,
                                   unsigned int slice_count, unsigned int *__restrict__ offsets, unsigned int *__restrict__ widths,
                                   unsigned int *__restrict__ slot_counts) {
    for (unsigned int s = blockIdx.x * blockDim.x + threadIdx.x; s < slice_count; s += gridDim.x * blockDim.x) {
        const unsigned int r0 = s * S;
        const unsigned int r1 = r0 + S < rows ? r0 + S : rows;
        unsigned int m = 0u;
        for (unsigned int r = r0; r < r1; ++r) m = row_nnz[r] > m ? row_nnz[r] : m;
        widths[s] = m;
        offsets[s] = r0;
        const unsigned long long sc = (unsigned long long) (r1 - r0) * m;
        slot_counts[s] = sc > 0xffffffffull ? 0xffffffffu : (unsigned int) sc;
        if (s == slice_count - 1u) offsets[slice_count] = rows;
    }
}

// Padded slot total for a candidate slice height (evaluate_sliced_candidate_).
__global__ void sell_padded_kernel(const unsigned int *__restrict__ row_nnz, unsigned int rows, unsigned int S,
                                   unsigned int slice_count, unsigned long long *__restrict__ total) {
    unsigned long long acc = 0ull;
    for (unsigned int s = blockIdx.x * blockDim.x + threadIdx.x; s < slice_count; s += gridDim.x * blockDim.x) {
        const unsigned int r0 = s * S;
        const unsigned int r1 = r0 + S < rows ? r0 + S : rows;
        unsigned int m = 0u;
        for (unsigned int r = r0; r < r1; ++r) m = row_nnz[r] > m ? row_nnz[r] : m;
        acc += (unsigned long long) (r1 - r0) * m;
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0u && acc != 0ull) atomicAdd(total, acc);
}

__device__ __forceinline__ unsigned int sell_slice_of(const csamd_sliced_ell_view &d, unsigned int row) {
    unsigned int slice;
    if (d.slice_rows != 0u) slice = row / d.slice_rows;
    else {
        unsigned int lo = 0u, hi = d.slice_count - 1u;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (d.slice_row_offsets[mid] <= row) lo = mid;
            else hi = mid - 1u;
        }
        slice = lo;
    }
    return slice;
}

// One wave per row: copy the CSR segment, pad the rest of the row's slots.
template<typename S>
__global__ void sell_emit_csr_kernel(csamd_compressed_view src, const S *__restrict__ sval, csamd_sliced_ell_view dst,
                                     S *__restrict__ dval) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int nw = blockDim.x >> 6;
    const bool quads = sell_quads_ok<S>(dst.col_idx, dval);
    for (unsigned int row = blockIdx.x * nw + wave; row < src.rows; row += gridDim.x * nw) {
        const unsigned int slice = sell_slice_of(dst, row);
        if (slice >= dst.slice_count) continue;
        const unsigned int r0 = dst.slice_row_offsets[slice];
        const unsigned int width = dst.slice_widths[slice];
        const unsigned long long base = (unsigned long long) dst.slice_slot_offsets[slice] + (unsigned long long) (row - r0) * width;
        const unsigned int b = src.majorPtr[row], e = src.majorPtr[row + 1u];
        const unsigned int n = e - b < width ? e - b : width;
        sell_write_row<S>(gmem((const unsigned int *) src.minorIdx), gmem(sval), b, n, width, base, gmem(dst.col_idx), gmem(dval), quads,
                          ~0ull);
    }
}

// One wave per row: compact nonzero valid entries in block-slot order.
template<typename S>
__global__ void sell_emit_bell_kernel(csamd_blocked_ell_view src, csamd_sliced_ell_view dst, S *__restrict__ dval) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int nw = blockDim.x >> 6;
    const unsigned int bs = src.block_size;
    const unsigned int width_blocks = bs != 0u ? src.ell_cols / bs : 0u;
    const S *sval = (const S *) src.val;
    const int slog = bell_stream_log2<S>(src);
    for (unsigned int row = blockIdx.x * nw + wave; row < src.rows; row += gridDim.x * nw) {
        const unsigned int slice = sell_slice_of(dst, row);
        if (slice >= dst.slice_count) continue;
        const unsigned int r0 = dst.slice_row_offsets[slice];
        const unsigned int width = dst.slice_widths[slice];
        const unsigned long long base = (unsigned long long) dst.slice_slot_offsets[slice] + (unsigned long long) (row - r0) * width;
        const unsigned int rb = row / bs;
        const unsigned int total = width_blocks * bs;
        unsigned int written = 0u;
        if (slog >= 0) {
            constexpr unsigned int PER = 16u / sizeof(S);
            const unsigned int nc = src.ell_cols / PER;
            const unsigned int cmask = (1u << slog) - 1u;
            CSAMD_GLOBAL unsigned int *ocol = gmem(dst.col_idx) + base;
            CSAMD_GLOBAL S *oval = gmem(dval) + base;
            bell_stream_row<S, BELL_STREAM_U>(src, width_blocks, (unsigned int) slog, row, [&](unsigned int c, const u32x4 &q, unsigned int bc) {
                const unsigned int col0 = bc * bs + (c & cmask) * PER;
                const unsigned int m = c < nc ? chunk_mask<S>(q, bc, col0, src.cols) : 0u;
                const unsigned int cnt = (unsigned int) __popc(m);
                const unsigned int incl = wave_inclusive_scan_u32(cnt);
                unsigned int pos = written + incl - cnt;
#pragma unroll
                for (unsigned int e = 0; e < PER; ++e) {
                    if ((m >> e) & 1u) {
                        if (pos < width) {
                            ocol[pos] = col0 + e;
                            oval[pos] = chunk_val<S>(q, e);
                        }
                        ++pos;
                    }
                }
                written += (unsigned int) __builtin_amdgcn_readlane((int) incl, 63);
            });
        } else {
        for (unsigned int e0 = 0u; e0 < total; e0 += 64u) {
            const unsigned int e = e0 + lane_id();
            bool keep = false;
            unsigned int col = 0u;
            S v = (S) 0;
            if (e < total) {
                const unsigned int slot = e / bs;
                const unsigned int bc = src.blockColIdx[(unsigned long long) rb * width_blocks + slot];
                if (bc != CSAMD_INVALID_COL) {
                    col = bc * bs + (e - slot * bs);
                    v = sval[(unsigned long long) row * src.ell_cols + e];
                    keep = col < src.cols && stored_nonzero<S>(v);
                }
            }
            const unsigned long long bal = __ballot(keep);
            const unsigned int pos = written + (unsigned int) __popcll(bal & lanemask_lt());
            if (keep && pos < width) {
                dst.col_idx[base + pos] = col;
                dval[base + pos] = v;
            }
            written += (unsigned int) __popcll(bal);
        }
        }
        sell_pad_row<S>(gmem(dst.col_idx), gmem(dval), base, written, width, sell_quads_ok<S>(dst.col_idx, dval));
    }
}

unsigned int rows_grid(unsigned int rows, unsigned int waves_per_block) {
    return grid_for((rows + waves_per_block - 1u) / waves_per_block, 1u, (unsigned int) device_cu_count() * 32u);
}

int layout_from_row_nnz(const unsigned int *d_row_nnz, unsigned int rows, unsigned int slice_rows, unsigned int *d_offsets,
                        unsigned int *d_widths, unsigned int *d_slot_offsets, csamd_convert_workspace *ws, void *scratch_after,
                        unsigned int *out_slice_count, unsigned long long *out_total) {
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int S = rows == 0u ? 0u : (slice_rows < 1u ? 1u : (slice_rows > rows ? rows : slice_rows));
    const unsigned int count = rows == 0u ? 0u : (rows + S - 1u) / S;
    if (out_slice_count != nullptr) *out_slice_count = count;
    if (count == 0u) {
        const unsigned int zero = 0u;
        if (d_offsets != nullptr) CSAMD_HIP(hipMemcpyAsync(d_offsets, &zero, 4u, hipMemcpyHostToDevice, s), "H2D offsets");
        CSAMD_HIP(hipStreamSynchronize(s), "sync");
        if (out_total != nullptr) *out_total = 0ull;
        return 1;
    }
    char *p = (char *) scratch_after;
    unsigned int *slot_counts = (unsigned int *) p;
    p += align_up((size_t) count * 4u, 256u);
    unsigned long long *slot_off64 = (unsigned long long *) p;
    p += align_up((size_t) (count + 1u) * 8u, 256u);
    void *scan_tmp = p;
    sell_slices_kernel<<<grid_for(count, 256u, 8192u), 256, 0, s>>>(d_row_nnz, rows, S, count, d_offsets, d_widths, slot_counts);
    CSAMD_HIP(hipGetLastError(), "sell_slices_kernel");
    if (!exclusive_scan_u32_to_u64(slot_counts, count, slot_off64, scan_tmp, exclusive_scan_scratch_bytes(count), s)) return 0;
    unsigned long long total = 0ull;
    CSAMD_HIP(hipMemcpyAsync(&total, slot_off64 + count, 8u, hipMemcpyDeviceToHost, s), "D2H total slots");
    CSAMD_HIP(hipStreamSynchronize(s), "sell layout sync");
    CSAMD_REQUIRE(total <= 0xffffffffull, "sliced_ell layout: %llu slots exceed the u32 slot offsets", total);
    if (!exclusive_scan_u32(slot_counts, count, d_slot_offsets, nullptr, scan_tmp, exclusive_scan_scratch_bytes(count), s, false))
        return 0;
    CSAMD_HIP(hipStreamSynchronize(s), "sell layout sync 2");
    if (out_total != nullptr) *out_total = total;
    return 1;
}

size_t layout_scratch_bytes(unsigned int rows, unsigned int slice_rows) {
    const unsigned int S = rows == 0u ? 1u : (slice_rows < 1u ? 1u : (slice_rows > rows ? rows : slice_rows));
    const unsigned int count = rows == 0u ? 1u : (rows + S - 1u) / S;
    return align_up((size_t) rows * 4u, 256u) + align_up((size_t) count * 4u, 256u) + align_up((size_t) (count + 1u) * 8u, 256u) +
           exclusive_scan_scratch_bytes(count) + 4096u;
}

template<template<typename> class K, typename... Args>
int bits_dispatch(unsigned int vt, Args... args) {
    const int bytes = value_type_bytes(vt);
    switch (bytes) {
        case 1: return K<uint8_t>::run(args...);
        case 2: return K<uint16_t>::run(args...);
        case 4: return K<float>::run(args...);
        default: set_error("unsupported value type %u", vt); return 0;
    }
}

} // namespace
} // namespace csamd

using namespace csamd;

namespace {

template<typename S> struct bell_emit_launch {
    static int run(const csamd_compressed_view *src, unsigned int row_blocks, unsigned int width, unsigned int nwords,
                   csamd_blocked_ell_view *dst, unsigned int *tab, unsigned int *unsorted, unsigned long long *d_nnz,
                   unsigned int *need_w, hipStream_t s) {
        const unsigned int cus = (unsigned int) device_cu_count();
        {
            const size_t lds = ((size_t) 2u * nwords + dst->block_size + 1u) * sizeof(unsigned int);
            auto k = bell_blocks_kernel;
            CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            k<<<grid_for(row_blocks, 1u, cus * 8u), CV_THREADS, lds, s>>>(*src, dst->block_size, row_blocks, width, nwords,
                                                                          dst->blockColIdx, tab, unsorted, need_w);
            CSAMD_HIP(hipGetLastError(), "bell_blocks_kernel");
        }
        {
            // value row segments of 16 or 32 KB (CSAMD_BELL_SEG = 4096 / 8192 words, for sweeps)
            const char *env = std::getenv("CSAMD_BELL_SEG");
            const unsigned int segw = env != nullptr && std::atoi(env) == 8192 ? 8192u : 4096u;
            const size_t lds = ((((size_t) 2u * nwords + 3u) & ~(size_t) 3u) + segw) * sizeof(unsigned int);
            auto k = segw == 8192u ? bell_rows_kernel<S, 16u, 8192u> : bell_rows_kernel<S, 16u, 4096u>;
            CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
            k<<<grid_for(src->rows, 1u, cus * 16u), CV_THREADS, lds, s>>>(*src, (const S *) src->val, dst->block_size, nwords,
                                                                          dst->ell_cols, tab, unsorted, (S *) dst->val, d_nnz);
            CSAMD_HIP(hipGetLastError(), "bell_rows_kernel");
        }
        bell_replay_kernel<S><<<grid_for(row_blocks, 64u, 1024u), 64, 0, s>>>(*src, (const S *) src->val, dst->block_size, row_blocks,
                                                                             nwords, dst->ell_cols, tab, unsorted, (S *) dst->val, d_nnz);
        return hip_ok(hipGetLastError(), "bell_replay_kernel");
    }
};

template<typename S> struct sell_emit_csr_launch {
    static int run(const csamd_compressed_view *src, csamd_sliced_ell_view *dst, hipStream_t s) {
        sell_emit_csr_kernel<S><<<rows_grid(src->rows, 4u), 256, 0, s>>>(*src, (const S *) src->val, *dst, (S *) dst->val);
        return hip_ok(hipGetLastError(), "sell_emit_csr_kernel");
    }
};

template<typename S> struct sell_emit_bell_launch {
    static int run(const csamd_blocked_ell_view *src, csamd_sliced_ell_view *dst, hipStream_t s) {
        sell_emit_bell_kernel<S><<<rows_grid(src->rows, 4u), 256, 0, s>>>(*src, *dst, (S *) dst->val);
        return hip_ok(hipGetLastError(), "sell_emit_bell_kernel");
    }
};

template<typename S> struct bell_row_nnz_launch {
    static int run(const csamd_blocked_ell_view *src, unsigned int *row_nnz, hipStream_t s) {
        bell_row_nnz_kernel<S><<<rows_grid(src->rows, 4u), 256, 0, s>>>(*src, row_nnz);
        return hip_ok(hipGetLastError(), "bell_row_nnz_kernel");
    }
};

int check_csr(const csamd_compressed_view *src, const char *who) {
    CSAMD_REQUIRE(src != nullptr, "%s: null source", who);
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW, "%s: source axis must be row", who);
    CSAMD_REQUIRE(src->rows == 0u || src->majorPtr != nullptr, "%s: null majorPtr", who);
    return 1;
}

int bell_shape_impl(const csamd_compressed_view *src, unsigned int bs, csamd_convert_workspace *ws, unsigned int *out_w,
                    unsigned long long *out_total) {
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int row_blocks = (src->rows + bs - 1u) / bs;
    const unsigned int col_blocks = (src->cols + bs - 1u) / bs;
    const unsigned int nwords = (col_blocks + 31u) / 32u;
    CSAMD_REQUIRE(nwords <= BELL_MAX_BITSET_WORDS, "blocked_ell: %u column blocks exceed the LDS bitset", col_blocks);
    CSAMD_REQUIRE(bs <= BELL_MAX_BLOCK_SIZE, "blocked_ell: block size %u exceeds %u", bs, BELL_MAX_BLOCK_SIZE);
    char *scratch = (char *) convert_scratch(ws, align_up((size_t) row_blocks * 4u, 256u) + 256u);
    if (scratch == nullptr) return 0;
    unsigned int *counts = (unsigned int *) (scratch + 256u);
    unsigned int *d_w = (unsigned int *) scratch;
    unsigned long long *d_total = (unsigned long long *) (scratch + 64u);
    CSAMD_HIP(hipMemsetAsync(scratch, 0, 128u, s), "memset shape");
    if (row_blocks != 0u) {
        auto k = bell_shape_kernel;
        const size_t lds = ((size_t) nwords + bs + 1u) * 4u;
        CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        k<<<grid_for(row_blocks, 1u, (unsigned int) device_cu_count() * 8u), CV_THREADS, lds, s>>>(*src, bs, row_blocks, nwords, counts, d_w,
                                                                                               d_total);
        CSAMD_HIP(hipGetLastError(), "bell_shape_kernel");
    }
    unsigned int w = 0u;
    unsigned long long total = 0ull;
    CSAMD_HIP(hipMemcpyAsync(&w, d_w, 4u, hipMemcpyDeviceToHost, s), "D2H width");
    CSAMD_HIP(hipMemcpyAsync(&total, d_total, 8u, hipMemcpyDeviceToHost, s), "D2H total");
    CSAMD_HIP(hipStreamSynchronize(s), "shape sync");
    *out_w = w;
    *out_total = total;
    return 1;
}

} // namespace

extern "C" {

int csamd_blocked_ell_shape(const csamd_compressed_view *src, unsigned int block_size, csamd_convert_workspace *ws,
                            unsigned int *out_ell_width, unsigned long long *out_total_blocks, double *out_fill_ratio) {
    if (!check_csr(src, "blocked_ell_shape")) return 0;
    CSAMD_REQUIRE(ws != nullptr && out_ell_width != nullptr, "blocked_ell_shape: null argument");
    CSAMD_REQUIRE(block_size != 0u, "blocked_ell_shape: block_size is 0");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    *out_ell_width = 0u;
    if (out_total_blocks != nullptr) *out_total_blocks = 0ull;
    if (out_fill_ratio != nullptr) *out_fill_ratio = 0.0;
    if (src->rows == 0u || src->cols == 0u) return 1;
    unsigned int w = 0u;
    unsigned long long total = 0ull;
    if (!bell_shape_impl(src, block_size, ws, &w, &total)) return 0;
    const unsigned int row_blocks = (src->rows + block_size - 1u) / block_size;
    *out_ell_width = w;
    if (out_total_blocks != nullptr) *out_total_blocks = total;
    if (out_fill_ratio != nullptr)
        *out_fill_ratio = (row_blocks == 0u || w == 0u) ? 1.0 : (double) total / ((double) row_blocks * (double) w);
    return 1;
}

int csamd_choose_blocked_ell_block_size(const csamd_compressed_view *src, const unsigned int *candidates,
                                        unsigned int candidate_count, unsigned int value_bytes, csamd_convert_workspace *ws,
                                        csamd_blocked_ell_tune_result *out) {
    if (!check_csr(src, "choose_blocked_ell_block_size")) return 0;
    CSAMD_REQUIRE(out != nullptr && candidates != nullptr && candidate_count != 0u && ws != nullptr,
                  "choose_blocked_ell_block_size: bad arguments");
    if (value_bytes == 0u) value_bytes = 2u;
    csamd_blocked_ell_tune_result best{0u, 0.0, 0u};
    for (unsigned int i = 0; i < candidate_count; ++i) {
        const unsigned int bs = candidates[i];
        if (bs == 0u) continue;
        unsigned int w = 0u;
        unsigned long long total = 0ull;
        double fill = 0.0;
        if (!csamd_blocked_ell_shape(src, bs, ws, &w, &total, &fill)) return 0;
        if (src->rows == 0u || src->cols == 0u) fill = 0.0;
        const size_t padded = (size_t) src->rows * (size_t) w * (size_t) bs * (size_t) value_bytes;
        if (best.block_size == 0u || padded + 1u < best.padded_bytes ||
            (padded == best.padded_bytes && fill > best.fill_ratio + 1.0e-9) ||
            (padded == best.padded_bytes && fill + 1.0e-9 >= best.fill_ratio && bs > best.block_size)) {
            best.block_size = bs;
            best.fill_ratio = fill;
            best.padded_bytes = padded;
        }
    }
    CSAMD_REQUIRE(best.block_size != 0u, "choose_blocked_ell_block_size: no usable candidate");
    *out = best;
    return 1;
}

int csamd_blocked_ell_from_compressed(const csamd_compressed_view *src, unsigned int value_type, csamd_blocked_ell_view *dst,
                                      csamd_convert_workspace *ws) {
    if (!check_csr(src, "blocked_ell_from_compressed")) return 0;
    CSAMD_REQUIRE(dst != nullptr && ws != nullptr, "blocked_ell_from_compressed: null argument");
    CSAMD_REQUIRE(dst->block_size != 0u, "blocked_ell_from_compressed: block_size is 0");
    CSAMD_REQUIRE(dst->ell_cols % dst->block_size == 0u, "blocked_ell_from_compressed: ell_cols not a multiple of block_size");
    CSAMD_REQUIRE(value_type_bytes(value_type) != 0 && value_type != CSAMD_VALUE_F64, "blocked_ell_from_compressed: bad value type");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int bs = dst->block_size;
    const unsigned int row_blocks = (src->rows + bs - 1u) / bs;
    const unsigned int width = dst->ell_cols / bs;
    const unsigned int col_blocks = (src->cols + bs - 1u) / bs;
    const unsigned int nwords = (col_blocks + 31u) / 32u;
    CSAMD_REQUIRE(nwords <= BELL_MAX_BITSET_WORDS, "blocked_ell: %u column blocks exceed the LDS bitset", col_blocks);
    CSAMD_REQUIRE(bs <= BELL_MAX_BLOCK_SIZE, "blocked_ell: block size %u exceeds %u", bs, BELL_MAX_BLOCK_SIZE);
    dst->rows = src->rows;
    dst->cols = src->cols;
    if (src->rows == 0u) {
        dst->nnz = 0u;
        return 1;
    }
    if (width == 0u) {  // no slots: every row block must be empty
        unsigned int need_w = 0u;
        unsigned long long total = 0ull;
        if (!bell_shape_impl(src, bs, ws, &need_w, &total)) return 0;
        CSAMD_REQUIRE(need_w == 0u, "blocked_ell_from_compressed: ell width 0 < required %u", need_w);
        dst->nnz = 0u;
        return 1;
    }
    CSAMD_REQUIRE(dst->blockColIdx != nullptr && dst->val != nullptr, "blocked_ell_from_compressed: destination arrays missing");
    // scratch: nnz and the widest row block's need (first 256 bytes), per
    // row block the bitset and its prefix, unsorted flags
    const size_t tab_bytes = align_up((size_t) row_blocks * 2u * nwords * 4u, 256u);
    char *scratch = (char *) convert_scratch(ws, 256u + tab_bytes + align_up((size_t) row_blocks * 4u, 256u));
    if (scratch == nullptr) return 0;
    unsigned long long *d_nnz = (unsigned long long *) scratch;
    unsigned int *d_need = (unsigned int *) (scratch + 64u);
    unsigned int *tab = (unsigned int *) (scratch + 256u);
    unsigned int *unsorted = (unsigned int *) (scratch + 256u + tab_bytes);
    CSAMD_HIP(hipMemsetAsync(scratch, 0, 128u, s), "memset nnz");
    if (!bits_dispatch<bell_emit_launch>(value_type, src, row_blocks, width, nwords, dst, tab, unsorted, d_nnz, d_need, s)) return 0;
    unsigned long long hres[9] = {};
    CSAMD_HIP(hipMemcpyAsync(hres, scratch, 72u, hipMemcpyDeviceToHost, s), "D2H bell nnz");
    CSAMD_HIP(hipStreamSynchronize(s), "bell emit sync");
    const unsigned long long nnz = hres[0];
    const unsigned int need = (unsigned int) hres[8];
    CSAMD_REQUIRE(need <= width, "blocked_ell_from_compressed: ell width %u < required %u", width, need);
    CSAMD_REQUIRE(nnz <= 0xffffffffull, "blocked_ell_from_compressed: nnz overflow");
    dst->nnz = (unsigned int) nnz;
    return 1;
}

int csamd_sliced_ell_layout(const csamd_compressed_view *src, unsigned int slice_rows, unsigned int *d_slice_row_offsets,
                            unsigned int *d_slice_widths, unsigned int *d_slice_slot_offsets, csamd_convert_workspace *ws,
                            unsigned int *out_slice_count, unsigned long long *out_total_slots) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr, "sliced_ell_layout: null argument");
    CSAMD_REQUIRE(slice_rows != 0u, "sliced_ell_layout: slice_rows is 0");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int rows = src->rows;
    char *scratch = (char *) convert_scratch(ws, layout_scratch_bytes(rows, slice_rows) + 256u);
    if (scratch == nullptr) return 0;
    unsigned int *bad = (unsigned int *) scratch;
    unsigned int *row_nnz = (unsigned int *) (scratch + 256u);
    void *after = scratch + 256u + align_up((size_t) rows * 4u, 256u);
    if (rows != 0u) {
        if (src->axis == CSAMD_COMPRESSED_BY_ROW) {
            csr_row_nnz_kernel<<<grid_for(rows, 256u, 8192u), 256, 0, s>>>(src->majorPtr, rows, row_nnz);
            CSAMD_HIP(hipGetLastError(), "csr_row_nnz_kernel");
        } else {
            CSAMD_HIP(hipMemsetAsync(scratch, 0, 256u + (size_t) rows * 4u, s), "memset row nnz");
            if (src->nnz != 0u) {
                csc_row_nnz_kernel<<<grid_for(src->nnz, 256u, 8192u), 256, 0, s>>>(src->minorIdx, src->nnz, rows, row_nnz, bad);
                CSAMD_HIP(hipGetLastError(), "csc_row_nnz_kernel");
            }
            unsigned int hbad = 0u;
            CSAMD_HIP(hipMemcpyAsync(&hbad, bad, 4u, hipMemcpyDeviceToHost, s), "D2H bad");
            CSAMD_HIP(hipStreamSynchronize(s), "sync");
            CSAMD_REQUIRE(hbad == 0u, "sliced_ell_layout: CSC row index out of range");
        }
    }
    return layout_from_row_nnz(row_nnz, rows, slice_rows, d_slice_row_offsets, d_slice_widths, d_slice_slot_offsets, ws, after,
                               out_slice_count, out_total_slots);
}

int csamd_sliced_ell_layout_from_blocked_ell(const csamd_blocked_ell_view *src, unsigned int value_type, unsigned int slice_rows,
                                             unsigned int *d_slice_row_offsets, unsigned int *d_slice_widths,
                                             unsigned int *d_slice_slot_offsets, csamd_convert_workspace *ws,
                                             unsigned int *out_slice_count, unsigned long long *out_total_slots) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr, "sliced_ell_layout_from_blocked_ell: null argument");
    CSAMD_REQUIRE(slice_rows != 0u && src->block_size != 0u, "sliced_ell_layout_from_blocked_ell: zero slice_rows/block_size");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int rows = src->rows;
    char *scratch = (char *) convert_scratch(ws, layout_scratch_bytes(rows, slice_rows) + 256u);
    if (scratch == nullptr) return 0;
    unsigned int *row_nnz = (unsigned int *) (scratch + 256u);
    void *after = scratch + 256u + align_up((size_t) rows * 4u, 256u);
    if (rows != 0u) {
        if (!bits_dispatch<bell_row_nnz_launch>(value_type, src, row_nnz, s)) return 0;
    }
    return layout_from_row_nnz(row_nnz, rows, slice_rows, d_slice_row_offsets, d_slice_widths, d_slice_slot_offsets, ws, after,
                               out_slice_count, out_total_slots);
}

int csamd_choose_sliced_ell_slice_rows(const csamd_compressed_view *src, const unsigned int *candidates,
                                       unsigned int candidate_count, csamd_convert_workspace *ws,
                                       csamd_sliced_ell_tune_result *out) {
    if (!check_csr(src, "choose_sliced_ell_slice_rows")) return 0;
    CSAMD_REQUIRE(candidates != nullptr && candidate_count != 0u && out != nullptr && ws != nullptr,
                  "choose_sliced_ell_slice_rows: bad arguments");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int rows = src->rows;
    const unsigned long long nnz = src->nnz;
    char *scratch = (char *) convert_scratch(ws, align_up((size_t) rows * 4u, 256u) + 512u);
    if (scratch == nullptr) return 0;
    unsigned long long *d_total = (unsigned long long *) scratch;
    unsigned int *row_nnz = (unsigned int *) (scratch + 256u);
    if (rows != 0u) {
        csr_row_nnz_kernel<<<grid_for(rows, 256u, 8192u), 256, 0, s>>>(src->majorPtr, rows, row_nnz);
        CSAMD_HIP(hipGetLastError(), "csr_row_nnz_kernel");
    }
    csamd_sliced_ell_tune_result best;
    std::memset(&best, 0, sizeof(best));
    best.score = __builtin_inf();
    for (unsigned int i = 0; i < candidate_count; ++i) {
        if (candidates[i] == 0u) continue;
        csamd_sliced_ell_tune_result r;
        std::memset(&r, 0, sizeof(r));
        if (rows == 0u) {
            r.storage_ratio = 1.0;
            r.score = 1.0;
        } else {
            const unsigned int S = candidates[i] > rows ? rows : candidates[i];
            const unsigned int count = (rows + S - 1u) / S;
            CSAMD_HIP(hipMemsetAsync(d_total, 0, 8u, s), "memset");
            sell_padded_kernel<<<grid_for(count, 256u, 4096u), 256, 0, s>>>(row_nnz, rows, S, count, d_total);
            CSAMD_HIP(hipGetLastError(), "sell_padded_kernel");
            unsigned long long padded = 0ull;
            CSAMD_HIP(hipMemcpyAsync(&padded, d_total, 8u, hipMemcpyDeviceToHost, s), "D2H padded");
            CSAMD_HIP(hipStreamSynchronize(s), "sync");
            r.slice_rows = S;
            r.slice_count = count;
            r.padded_slots = padded;
            r.padded_bytes = padded * 6u + ((size_t) count * 2u + 1u) * 4u;
            r.storage_ratio = nnz == 0ull ? 1.0 : (double) padded / (double) nnz;
            r.slice_penalty = 0.25 * ((double) count / (double) rows);
            r.score = r.storage_ratio + r.slice_penalty;
        }
        if (best.slice_rows == 0u || r.score + 1.0e-12 < best.score ||
            (r.score <= best.score + 1.0e-12 && r.padded_bytes < best.padded_bytes) ||
            (r.score <= best.score + 1.0e-12 && r.padded_bytes == best.padded_bytes && r.slice_rows > best.slice_rows)) {
            best = r;
        }
    }
    CSAMD_REQUIRE(!(best.slice_rows == 0u && rows != 0u), "choose_sliced_ell_slice_rows: no usable candidate");
    *out = best;
    return 1;
}

int csamd_sliced_ell_from_compressed(const csamd_compressed_view *src, unsigned int value_type, csamd_sliced_ell_view *dst,
                                     csamd_convert_workspace *ws) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && ws != nullptr, "sliced_ell_from_compressed: null argument");
    CSAMD_REQUIRE(value_type_bytes(value_type) != 0 && value_type != CSAMD_VALUE_F64, "sliced_ell_from_compressed: bad value type");
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW,
                  "sliced_ell_from_compressed: transpose a CSC source to CSR first (ascending columns per row)");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    dst->rows = src->rows;
    dst->cols = src->cols;
    dst->nnz = src->nnz;
    if (src->rows == 0u || dst->slice_count == 0u) return 1;
    return bits_dispatch<sell_emit_csr_launch>(value_type, src, dst, (hipStream_t) ws->stream);
}

int csamd_sliced_ell_from_blocked_ell(const csamd_blocked_ell_view *src, unsigned int value_type, csamd_sliced_ell_view *dst,
                                      csamd_convert_workspace *ws) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && ws != nullptr, "sliced_ell_from_blocked_ell: null argument");
    CSAMD_REQUIRE(src->block_size != 0u, "sliced_ell_from_blocked_ell: block_size is 0");
    CSAMD_REQUIRE(value_type_bytes(value_type) != 0 && value_type != CSAMD_VALUE_F64, "sliced_ell_from_blocked_ell: bad value type");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    dst->rows = src->rows;
    dst->cols = src->cols;
    dst->nnz = src->nnz;
    if (src->rows == 0u || dst->slice_count == 0u) return 1;
    return bits_dispatch<sell_emit_bell_launch>(value_type, src, dst, (hipStream_t) ws->stream);
}

} // extern "C"
