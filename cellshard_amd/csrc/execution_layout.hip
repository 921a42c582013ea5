// Bucketed execution partitions: a Blocked-ELL (Sliced-ELL) partition split
// into a few segments of row blocks (rows) with similar valid widths, each
// segment stored at its own width, plus the row permutation both ways.
//
// Reference (host, single thread): src/io/csh5/execution_builders_part.hh
//   Blocked-ELL: build_execution_bucket_layout :1035-1098 (stable sort of row
//   blocks by valid slots, the partial last block left in place; equal-count
//   cuts floor(b n / B); equal widths collapse to one segment),
//   allocate/fill :1100-1204, build_bucketed_execution_partition :1206-1246
//   (one segment when the split does not shrink the packed bytes),
//   choose_bucket_count_for_part :1248-1275 (fewest bytes over 1..8 buckets,
//   ties to fewer);
//   Sliced-ELL: build_sliced_execution_bucket_layout :1312-1429 (stable sort
//   of rows by valid slots, then the k-segment split of fewest bytes, equal
//   adjacent widths merged), fill :1498-1533, build :1535-1568, choose
//   :1570-1596.
//   Packed-byte estimators: include/CellShard/io/pack/packfile.cuh:109-116,
//   src/io/pack/packfile.cu:179-188 (2-byte values, as the reference counts).
//
// Here the widths, the stable sorts and the segment fills run on the GPU
// (one wave per row / workgroup per row block); the layout rules, on a few
// integers per row block or row, run on the host.  The segments are device
// views; the library owns their memory until the clear call.
#include "common.hh"

#include <algorithm>
#include <vector>

namespace csamd {
namespace {

constexpr unsigned int EL_MAX_BUCKETS = 8u;
constexpr unsigned int EL_FILL_THREADS = 256u;
constexpr unsigned long long EL_INF = ~0ull / 4ull;

// packfile.cuh:110-116 with the 16-byte disk header and 2-byte values
unsigned long long bell_packed_bytes(unsigned long long rows, unsigned long long ell_cols, unsigned long long bs) {
    const unsigned long long rb = bs != 0ull ? (rows + bs - 1ull) / bs : 0ull;
    const unsigned long long w = bs != 0ull ? ell_cols / bs : 0ull;
    return 16ull + 4ull + 4ull + rb * w * 4ull + rows * ell_cols * 2ull;
}

// packfile.cu:179-188 with the 13-byte header and 2-byte values
unsigned long long sell_packed_bytes(unsigned long long slices, unsigned long long slots) {
    return 13ull + 4ull + (slices + 1ull) * 4ull + slices * 4ull + slots * 4ull + slots * 2ull;
}

// ---------------------------------------------------------------------------
// Device kernels.
// ---------------------------------------------------------------------------
__global__ void bell_valid_width_kernel(const unsigned int *__restrict__ bci, unsigned int row_blocks, unsigned int W,
                                        unsigned int *__restrict__ widths, unsigned int *__restrict__ order) {
    const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
    for (unsigned int rb = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; rb < row_blocks; rb += nw) {
        unsigned int n = 0u;
        for (unsigned int s = lane_id(); s < W; s += 64u) n += bci[(size_t) rb * W + s] != CSAMD_INVALID_COL ? 1u : 0u;
        n = wave_sum_u32(n);
        if (lane_id() == 0u) {
            widths[rb] = n;
            order[rb] = rb;
        }
    }
}

__device__ __forceinline__ unsigned int sell_slice_of(const csamd_sliced_ell_view &v, unsigned int row) {
    if (v.slice_rows != 0u) {
        const unsigned int s = row / v.slice_rows;
        return s < v.slice_count ? s : v.slice_count - 1u;
    }
    unsigned int lo = 0u, hi = v.slice_count - 1u;
    while (lo < hi) {
        const unsigned int mid = (lo + hi + 1u) >> 1;
        if (v.slice_row_offsets[mid] <= row) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

__global__ void sell_row_width_kernel(csamd_sliced_ell_view v, unsigned int *__restrict__ widths, unsigned int *__restrict__ order) {
    const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
    for (unsigned int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < v.rows; row += nw) {
        const unsigned int s = sell_slice_of(v, row);
        const unsigned int w = v.slice_widths[s];
        const size_t base = (size_t) v.slice_slot_offsets[s] + (size_t) (row - v.slice_row_offsets[s]) * w;
        unsigned int n = 0u;
        for (unsigned int k = lane_id(); k < w; k += 64u) n += v.col_idx[base + k] != CSAMD_INVALID_COL ? 1u : 0u;
        n = wave_sum_u32(n);
        if (lane_id() == 0u) {
            widths[row] = n;
            order[row] = row;
        }
    }
}

struct seg_table {
    unsigned int count;
    unsigned int begin[EL_MAX_BUCKETS + 1u];  // in sorted positions (row blocks or rows)
    unsigned int row0[EL_MAX_BUCKETS + 1u];   // first exec row of each segment
    unsigned int width[EL_MAX_BUCKETS];       // slots (Blocked-ELL: column blocks) per row of the segment
    unsigned int *idx[EL_MAX_BUCKETS];        // blockColIdx / col_idx of the segment
    void *val[EL_MAX_BUCKETS];
};

__device__ __forceinline__ unsigned int seg_of(const seg_table &T, unsigned int pos) {
    unsigned int s = 0u;
    while (s + 1u < T.count && T.begin[s + 1u] <= pos) ++s;
    return s;
}

// One workgroup per exec row block (position p of the order): the source
// row block's valid slots, compacted in slot order, and their values.
template<typename S>
__global__ void __launch_bounds__(EL_FILL_THREADS) bell_fill_kernel(csamd_blocked_ell_view src, const unsigned int *__restrict__ order,
                                                                   unsigned int row_blocks, seg_table T,
                                                                   unsigned int *__restrict__ exec_to_canon,
                                                                   unsigned int *__restrict__ canon_to_exec) {
    extern __shared__ unsigned int map[];  // [W] destination slot of each source slot, or ~0
    __shared__ unsigned int scan_smem[16];
    __shared__ unsigned int valid_total;
    const unsigned int p = blockIdx.x;
    if (p >= row_blocks) return;
    const unsigned int B = src.block_size;
    const unsigned int W = src.ell_cols / B;
    const unsigned int src_rb = order[p];
    const unsigned int s = seg_of(T, p);
    const unsigned int j = p - T.begin[s];
    const unsigned int Wd = T.width[s];
    const unsigned int rows_in = src.rows - src_rb * B < B ? src.rows - src_rb * B : B;
    unsigned int carry = 0u;
    for (unsigned int t0 = 0u; t0 < W; t0 += EL_FILL_THREADS) {
        const unsigned int t = t0 + threadIdx.x;
        const unsigned int bc = t < W ? src.blockColIdx[(size_t) src_rb * W + t] : CSAMD_INVALID_COL;
        const unsigned int valid = bc != CSAMD_INVALID_COL ? 1u : 0u;
        unsigned int tot = 0u;
        const unsigned int o = carry + block_exclusive_scan_u32(valid, scan_smem, &tot);
        if (t < W) {
            map[t] = valid ? o : ~0u;
            if (valid) T.idx[s][(size_t) j * Wd + o] = bc;
        }
        carry += tot;
    }
    if (threadIdx.x == 0u) valid_total = carry;
    __syncthreads();
    const unsigned int nvalid = valid_total;
    for (unsigned int k = nvalid + threadIdx.x; k < Wd; k += EL_FILL_THREADS) T.idx[s][(size_t) j * Wd + k] = CSAMD_INVALID_COL;
    const S *sv = (const S *) src.val;
    S *dv = (S *) T.val[s];
    const size_t dst_cols = (size_t) Wd * B;
    const size_t per_row = (size_t) W * B;
    for (size_t e = threadIdx.x; e < (size_t) rows_in * per_row; e += EL_FILL_THREADS) {
        const unsigned int r = (unsigned int) (e / per_row);
        const unsigned int in_row = (unsigned int) (e - (size_t) r * per_row);
        const unsigned int t = in_row / B, l = in_row - t * B;
        const unsigned int d = map[t];
        if (d != ~0u) dv[((size_t) j * B + r) * dst_cols + (size_t) d * B + l] = sv[((size_t) src_rb * B + r) * per_row + in_row];
    }
    const size_t tail = (size_t) (Wd - nvalid) * B;
    for (size_t e = threadIdx.x; e < (size_t) rows_in * tail; e += EL_FILL_THREADS) {
        const unsigned int r = (unsigned int) (e / tail);
        const size_t k = e - (size_t) r * tail;
        dv[((size_t) j * B + r) * dst_cols + (size_t) nvalid * B + k] = S(0);
    }
    for (unsigned int r = threadIdx.x; r < rows_in; r += EL_FILL_THREADS) {
        const unsigned int exec = T.row0[s] + j * B + r;
        exec_to_canon[exec] = src_rb * B + r;
        canon_to_exec[src_rb * B + r] = exec;
    }
}

// One wave per exec row (position p of the order): the source row's valid
// slots compacted into the segment's single slice, padded to its width.
template<typename S>
__global__ void sell_fill_kernel(csamd_sliced_ell_view src, const unsigned int *__restrict__ order, seg_table T,
                                 unsigned int *__restrict__ exec_to_canon, unsigned int *__restrict__ canon_to_exec) {
    const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
    const unsigned int lane = lane_id();
    const S *sv = (const S *) src.val;
    for (unsigned int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; p < src.rows; p += nw) {
        const unsigned int row = order[p];
        const unsigned int s = seg_of(T, p);
        const unsigned int dr = p - T.begin[s];
        const unsigned int Wd = T.width[s];
        const unsigned int sl = sell_slice_of(src, row);
        const unsigned int w = src.slice_widths[sl];
        const size_t base = (size_t) src.slice_slot_offsets[sl] + (size_t) (row - src.slice_row_offsets[sl]) * w;
        unsigned int *di = T.idx[s] + (size_t) dr * Wd;
        S *dv = (S *) T.val[s] + (size_t) dr * Wd;
        unsigned int out = 0u;
        for (unsigned int k0 = 0u; k0 < w; k0 += 64u) {
            const unsigned int k = k0 + lane;
            const unsigned int c = k < w ? src.col_idx[base + k] : CSAMD_INVALID_COL;
            const bool valid = c != CSAMD_INVALID_COL;
            const unsigned long long m = __ballot(valid);
            const unsigned int before = (unsigned int) __popcll(m & ((1ull << lane) - 1ull));
            if (valid) {
                di[out + before] = c;
                dv[out + before] = sv[base + k];
            }
            out += (unsigned int) __popcll(m);
        }
        for (unsigned int k = out + lane; k < Wd; k += 64u) {
            di[k] = CSAMD_INVALID_COL;
            dv[k] = S(0);
        }
        if (lane == 0u) {
            exec_to_canon[T.row0[s] + dr] = row;
            canon_to_exec[row] = T.row0[s] + dr;
        }
    }
}

__global__ void iota_kernel(unsigned int *__restrict__ p, unsigned int n) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

// ---------------------------------------------------------------------------
// Host: widths, sorted order, layouts.
// ---------------------------------------------------------------------------
struct owned {
    std::vector<void *> dev;
    std::vector<void *> host;
    void *d(size_t bytes) {
        void *q = nullptr;
        if (hipMalloc(&q, bytes == 0u ? 16u : bytes) != hipSuccess) return nullptr;
        dev.push_back(q);
        return q;
    }
    void *h(size_t bytes) {
        void *q = std::calloc(1u, bytes == 0u ? 16u : bytes);
        if (q != nullptr) host.push_back(q);
        return q;
    }
    void release() {
        for (void *q : dev) (void) hipFree(q);
        for (void *q : host) std::free(q);
        dev.clear();
        host.clear();
    }
};

unsigned int bits_for(unsigned int n) {
    unsigned int b = 0u;
    while (b < 32u && (n >> b) != 0u) ++b;
    return b;
}

// Widths (valid slots) per item and the order stable-sorted by width over
// the first `sortable` items (the rest keep their place); both on the host.
int sorted_widths(unsigned int n, unsigned int sortable, unsigned int *d_widths, unsigned int *d_order, hipStream_t s,
                  std::vector<unsigned int> &widths, std::vector<unsigned int> &order) {
    widths.assign(n, 0u);
    order.assign(n, 0u);
    if (n == 0u) return 1;
    CSAMD_HIP(hipMemcpyAsync(widths.data(), d_widths, (size_t) n * 4u, hipMemcpyDeviceToHost, s), "D2H widths");
    CSAMD_HIP(hipStreamSynchronize(s), "widths sync");
    unsigned int wmax = 0u;
    for (unsigned int i = 0; i < sortable; ++i) wmax = std::max(wmax, widths[i]);
    if (sortable > 1u) {
        const size_t sb = radix_sort_pairs_u32_scratch_bytes(sortable);
        void *scratch = nullptr, *kout = nullptr, *vout = nullptr;
        CSAMD_HIP(hipMalloc(&scratch, sb), "hipMalloc sort scratch");
        int ok = hip_ok(hipMalloc(&kout, (size_t) sortable * 4u), "hipMalloc sort keys") &&
                 hip_ok(hipMalloc(&vout, (size_t) sortable * 4u), "hipMalloc sort vals");
        if (ok)
            ok = radix_sort_pairs_u32(d_widths, (unsigned int *) kout, d_order, (unsigned int *) vout, sortable, bits_for(wmax), scratch,
                                      sb, s) &&
                 hip_ok(hipMemcpyAsync(d_order, vout, (size_t) sortable * 4u, hipMemcpyDeviceToDevice, s), "D2D order");
        if (ok) ok = hip_ok(hipStreamSynchronize(s), "sort sync");
        (void) hipFree(scratch);
        if (kout != nullptr) (void) hipFree(kout);
        if (vout != nullptr) (void) hipFree(vout);
        if (!ok) return 0;
    }
    CSAMD_HIP(hipMemcpy(order.data(), d_order, (size_t) n * 4u, hipMemcpyDeviceToHost), "D2H order");
    return 1;
}

struct layout {
    std::vector<unsigned int> begin;  // [segments + 1] sorted positions
    std::vector<unsigned int> width;  // [segments]
};

// execution_builders_part.hh:1035-1098
layout bell_layout(const std::vector<unsigned int> &widths, const std::vector<unsigned int> &order, unsigned int requested) {
    layout L;
    const unsigned int n = (unsigned int) order.size();
    if (n == 0u) {
        L.begin.push_back(0u);
        return L;
    }
    const unsigned int buckets = std::max(1u, std::min(requested, n));
    L.begin.push_back(0u);
    for (unsigned int b = 0; b < buckets; ++b) {
        const unsigned int lo = (unsigned int) (((unsigned long long) b * n) / buckets);
        const unsigned int hi = (unsigned int) (((unsigned long long) (b + 1u) * n) / buckets);
        unsigned int w = 0u;
        for (unsigned int q = lo; q < hi; ++q) w = std::max(w, widths[order[q]]);
        L.width.push_back(w);
        L.begin.push_back(hi);
    }
    if (L.width.size() > 1u && std::all_of(L.width.begin(), L.width.end(), [&](unsigned int w) { return w == L.width[0]; })) {
        const unsigned int w = L.width[0];
        L.begin.assign({0u, n});
        L.width.assign(1u, w);
    }
    return L;
}

unsigned long long bell_layout_bytes(const layout &L, const std::vector<unsigned int> &order, unsigned int rows, unsigned int bs) {
    unsigned long long total = 0ull;
    for (size_t s = 0; s < L.width.size(); ++s) {
        unsigned long long seg_rows = 0ull;
        for (unsigned int q = L.begin[s]; q < L.begin[s + 1]; ++q) {
            const unsigned int rb = order[q];
            seg_rows += std::min<unsigned long long>(bs, (unsigned long long) rows - (unsigned long long) rb * bs);
        }
        total += bell_packed_bytes(seg_rows, (unsigned long long) L.width[s] * bs, bs);
    }
    return total;
}

// build_bucketed_execution_partition's byte rule: a split that does not
// shrink the packed bytes falls back to one segment (counted at the
// original bytes).
layout bell_bucketed(const std::vector<unsigned int> &widths, const std::vector<unsigned int> &order, unsigned int requested,
                     unsigned int rows, unsigned int bs, unsigned int ell_cols, unsigned long long *bytes) {
    layout L = bell_layout(widths, order, requested);
    unsigned long long b = bell_layout_bytes(L, order, rows, bs);
    const unsigned long long original = bell_packed_bytes(rows, ell_cols, bs);
    if (L.width.size() > 1u && b >= original) {
        L = bell_layout(widths, order, 1u);
        b = original;
    }
    *bytes = b;
    return L;
}

// The k-segment split of the stably sorted widths with the fewest bytes
// (execution_builders_part.hh:1312-1429).  A segment of sorted rows (i, j]
// costs F + S x w[j-1] x (j - i) (width = its largest, the last).  Exactly
// k = min(requested, rows) non-empty segments; dp_k(j) over j uses the
// lines y = dp_{k-1}(i) - S i x queried at x = w[j-1] (non-decreasing), so
// a monotone lower envelope answers each query in amortised O(1): the query
// pointer moves on only to a strictly lower line, a middle line leaves the
// envelope when the two lines around it meet at or below it.  Adjacent
// segments of equal width merge.
layout sell_layout(const std::vector<unsigned int> &w, unsigned int requested) {
    layout L;
    const unsigned int n = (unsigned int) w.size();
    if (n == 0u) {
        L.begin.push_back(0u);
        return L;
    }
    const long long F = (long long) sell_packed_bytes(1ull, 0ull), S = 6;
    const unsigned int K = std::max(1u, std::min(requested, n));
    std::vector<unsigned long long> prev(n + 1u, EL_INF), cur(n + 1u, EL_INF);
    std::vector<std::vector<unsigned int>> cut(K + 1u, std::vector<unsigned int>(n + 1u, 0u));
    prev[0] = 0ull;
    struct line {
        long long m, c;  // y = m x + c
        unsigned int at;
    };
    std::vector<line> env;
    for (unsigned int k = 1; k <= K; ++k) {
        std::fill(cur.begin(), cur.end(), EL_INF);
        env.clear();
        size_t head = 0u;
        auto value = [](const line &l, long long x) { return l.c + l.m * x; };
        auto add = [&](unsigned int i) {
            const line nl{-S * (long long) i, (long long) prev[i], i};
            while (env.size() >= head + 2u) {
                const line &a = env[env.size() - 2u], &b = env.back();
                // b is off the envelope when a and nl meet at or below it
                const __int128 lhs = (__int128) (nl.c - a.c) * (__int128) (a.m - b.m);
                const __int128 rhs = (__int128) (b.c - a.c) * (__int128) (a.m - nl.m);
                if (lhs <= rhs) env.pop_back();
                else break;
            }
            env.push_back(nl);
        };
        if (prev[k - 1u] != EL_INF) add(k - 1u);
        for (unsigned int j = k; j <= n; ++j) {
            const long long x = (long long) w[j - 1u];
            while (head + 1u < env.size() && value(env[head + 1u], x) < value(env[head], x)) ++head;
            if (head < env.size()) {
                cur[j] = (unsigned long long) (F + S * x * (long long) j + value(env[head], x));
                cut[k][j] = env[head].at;
            }
            if (prev[j] != EL_INF) add(j);
        }
        prev.swap(cur);
    }
    std::vector<unsigned int> off(K + 1u, 0u), wid(K, 0u);
    unsigned int j = n;
    off[K] = n;
    for (unsigned int k = K; k != 0u; --k) {
        const unsigned int i = cut[k][j];
        off[k - 1u] = i;
        wid[k - 1u] = w[j - 1u];
        j = i;
    }
    L.begin.push_back(0u);
    for (unsigned int k = 0; k < K; ++k) {
        if (!L.width.empty() && L.width.back() == wid[k]) L.begin.back() = off[k + 1u];
        else {
            L.width.push_back(wid[k]);
            L.begin.push_back(off[k + 1u]);
        }
    }
    return L;
}

unsigned long long sell_layout_bytes(const layout &L) {
    unsigned long long total = 0ull;
    for (size_t s = 0; s < L.width.size(); ++s) {
        const unsigned long long rows = L.begin[s + 1] - L.begin[s];
        total += sell_packed_bytes(rows != 0ull ? 1ull : 0ull, rows * L.width[s]);
    }
    return total;
}

layout sell_bucketed(const std::vector<unsigned int> &sorted_w, unsigned int requested, unsigned long long original,
                     unsigned long long *bytes) {
    layout L = sell_layout(sorted_w, requested);
    unsigned long long b = sell_layout_bytes(L);
    if (L.width.size() > 1u && b >= original) {
        L = sell_layout(sorted_w, 1u);
        b = original;
    }
    *bytes = b;
    return L;
}

int check_bell(const csamd_blocked_ell_view *src) {
    CSAMD_REQUIRE(src != nullptr, "bucketed blocked_ell: null source");
    CSAMD_REQUIRE(src->block_size != 0u && src->ell_cols % src->block_size == 0u, "bucketed blocked_ell: bad block size");
    CSAMD_REQUIRE(src->rows == 0u || src->ell_cols == 0u || (src->blockColIdx != nullptr && src->val != nullptr),
                  "bucketed blocked_ell: null arrays");
    return 1;
}

int bell_prepare(const csamd_blocked_ell_view *src, csamd_convert_workspace *ws, std::vector<unsigned int> &widths,
                 std::vector<unsigned int> &order, unsigned int **d_order_out, owned &mem) {
    const unsigned int bs = src->block_size;
    const unsigned int rb = (src->rows + bs - 1u) / bs;
    const unsigned int W = src->ell_cols / bs;
    hipStream_t s = (hipStream_t) ws->stream;
    unsigned int *d_w = (unsigned int *) mem.d((size_t) rb * 4u), *d_o = (unsigned int *) mem.d((size_t) rb * 4u);
    CSAMD_REQUIRE(d_w != nullptr && d_o != nullptr, "bucketed blocked_ell: device allocation failed");
    if (rb != 0u) {  // W == 0: every width is 0
        bell_valid_width_kernel<<<grid_for((unsigned long long) rb * 64u, 256u, 4096u), 256, 0, s>>>(src->blockColIdx, rb, W, d_w, d_o);
        CSAMD_HIP(hipGetLastError(), "bell widths");
    }
    const bool partial_last = rb != 0u && src->rows - (rb - 1u) * bs != bs;
    if (!sorted_widths(rb, partial_last ? rb - 1u : rb, d_w, d_o, s, widths, order)) return 0;
    *d_order_out = d_o;
    return 1;
}

int sell_prepare(const csamd_sliced_ell_view *src, csamd_convert_workspace *ws, std::vector<unsigned int> &widths,
                 std::vector<unsigned int> &order, unsigned int **d_order_out, owned &mem) {
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int n = src->rows;
    unsigned int *d_w = (unsigned int *) mem.d((size_t) n * 4u), *d_o = (unsigned int *) mem.d((size_t) n * 4u);
    CSAMD_REQUIRE(d_w != nullptr && d_o != nullptr, "bucketed sliced_ell: device allocation failed");
    if (n != 0u) {
        sell_row_width_kernel<<<grid_for((unsigned long long) n * 64u, 256u, 4096u), 256, 0, s>>>(*src, d_w, d_o);
        CSAMD_HIP(hipGetLastError(), "sell widths");
    }
    if (!sorted_widths(n, n, d_w, d_o, s, widths, order)) return 0;
    *d_order_out = d_o;
    return 1;
}

unsigned long long sell_total_slots_host(const csamd_sliced_ell_view *src) {
    if (src->slice_count == 0u) return 0ull;
    std::vector<unsigned int> off(src->slice_count + 1u), w(src->slice_count);
    if (hipMemcpy(off.data(), src->slice_row_offsets, off.size() * 4u, hipMemcpyDeviceToHost) != hipSuccess) return EL_INF;
    if (hipMemcpy(w.data(), src->slice_widths, w.size() * 4u, hipMemcpyDeviceToHost) != hipSuccess) return EL_INF;
    unsigned long long t = 0ull;
    for (unsigned int s = 0; s < src->slice_count; ++s) t += (unsigned long long) (off[s + 1u] - off[s]) * w[s];
    return t;
}

struct el_state {
    owned mem;
};

} // namespace

unsigned long long exec_bell_packed_bytes(unsigned long long rows, unsigned long long ell_cols, unsigned long long block_size) {
    return bell_packed_bytes(rows, ell_cols, block_size);
}
unsigned long long exec_sell_packed_bytes(unsigned long long slices, unsigned long long slots) { return sell_packed_bytes(slices, slots); }

void *exec_storage_new() { return new el_state(); }
void *exec_storage_dev(void *storage, size_t bytes) { return ((el_state *) storage)->mem.d(bytes); }
void *exec_storage_host(void *storage, size_t bytes) { return ((el_state *) storage)->mem.h(bytes); }

} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_bucketed_blocked_ell_clear(csamd_bucketed_blocked_ell *m) {
    if (m == nullptr) return;
    el_state *st = (el_state *) m->storage;
    if (st != nullptr) {
        st->mem.release();
        delete st;
    }
    std::memset(m, 0, sizeof(*m));
}

void csamd_bucketed_sliced_ell_clear(csamd_bucketed_sliced_ell *m) {
    if (m == nullptr) return;
    el_state *st = (el_state *) m->storage;
    if (st != nullptr) {
        st->mem.release();
        delete st;
    }
    std::memset(m, 0, sizeof(*m));
}

int csamd_choose_bucket_count_blocked_ell(const csamd_blocked_ell_view *src, csamd_convert_workspace *ws, unsigned int *bucket_count,
                                          unsigned long long *bucketed_bytes) {
    if (!check_bell(src)) return 0;
    CSAMD_REQUIRE(ws != nullptr && bucket_count != nullptr && bucketed_bytes != nullptr, "choose_bucket_count: null argument");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    owned mem;
    std::vector<unsigned int> widths, order;
    unsigned int *d_order = nullptr;
    const int ok = bell_prepare(src, ws, widths, order, &d_order, mem);
    mem.release();
    if (!ok) return 0;
    const unsigned int rb = (unsigned int) order.size();
    const unsigned int maxb = std::max(1u, std::min(EL_MAX_BUCKETS, rb));
    unsigned int best_b = 1u;
    unsigned long long best = ~0ull;
    for (unsigned int b = 1; b <= maxb; ++b) {
        unsigned long long bytes = 0ull;
        (void) bell_bucketed(widths, order, b, src->rows, src->block_size, src->ell_cols, &bytes);
        if (bytes < best || (bytes == best && b < best_b)) {
            best = bytes;
            best_b = b;
        }
    }
    *bucket_count = best_b;
    *bucketed_bytes = best;
    return 1;
}

int csamd_build_bucketed_blocked_ell(const csamd_blocked_ell_view *src, unsigned int value_type, unsigned int requested_bucket_count,
                                     csamd_convert_workspace *ws, csamd_bucketed_blocked_ell *out) {
    if (!check_bell(src)) return 0;
    CSAMD_REQUIRE(ws != nullptr && out != nullptr, "build_bucketed_blocked_ell: null argument");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "build_bucketed_blocked_ell: unsupported value type %u", value_type);
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    csamd_bucketed_blocked_ell_clear(out);
    el_state *st = new el_state();
    out->storage = st;
    owned scratch;
    std::vector<unsigned int> widths, order;
    unsigned int *d_order = nullptr;
    hipStream_t s = (hipStream_t) ws->stream;
    const unsigned int bs = src->block_size, rows = src->rows, W = src->ell_cols / bs;
    CSAMD_REQUIRE(W <= 16000u, "build_bucketed_blocked_ell: %u column blocks per row block exceed the LDS map", W);
    if (!bell_prepare(src, ws, widths, order, &d_order, scratch)) {
        scratch.release();
        csamd_bucketed_blocked_ell_clear(out);
        return 0;
    }
    unsigned long long bytes = 0ull;
    const layout L = bell_bucketed(widths, order, requested_bucket_count, rows, bs, src->ell_cols, &bytes);
    const unsigned int nseg = (unsigned int) L.width.size();
    out->rows = rows;
    out->cols = src->cols;
    out->nnz = src->nnz;
    out->segment_count = nseg;
    out->bytes = bytes;
    out->segments = (csamd_blocked_ell_view *) st->mem.h(nseg * sizeof(csamd_blocked_ell_view));
    out->segment_row_offsets = (unsigned int *) st->mem.h((nseg + 1u) * sizeof(unsigned int));
    out->exec_to_canonical_rows = (unsigned int *) st->mem.d((size_t) rows * 4u);
    out->canonical_to_exec_rows = (unsigned int *) st->mem.d((size_t) rows * 4u);
    out->exec_to_canonical_cols = (unsigned int *) st->mem.d((size_t) src->cols * 4u);
    out->canonical_to_exec_cols = (unsigned int *) st->mem.d((size_t) src->cols * 4u);
    int ok = out->segments != nullptr && out->segment_row_offsets != nullptr && out->exec_to_canonical_rows != nullptr &&
             out->canonical_to_exec_rows != nullptr && out->exec_to_canonical_cols != nullptr && out->canonical_to_exec_cols != nullptr;
    seg_table T;
    std::memset(&T, 0, sizeof(T));
    T.count = nseg;
    unsigned int row_cursor = 0u;
    for (unsigned int g = 0; ok && g < nseg; ++g) {
        unsigned int seg_rows = 0u, seg_nnz = 0u;
        for (unsigned int q = L.begin[g]; q < L.begin[g + 1]; ++q) {
            const unsigned int rbi = order[q];
            const unsigned int r_in = std::min(bs, rows - rbi * bs);
            seg_rows += r_in;
            seg_nnz += r_in * widths[rbi] * bs;
        }
        out->segment_row_offsets[g] = row_cursor;
        T.begin[g] = L.begin[g];
        T.row0[g] = row_cursor;
        T.width[g] = L.width[g];
        csamd_blocked_ell_view &v = out->segments[g];
        std::memset(&v, 0, sizeof(v));
        v.rows = seg_rows;
        v.cols = src->cols;
        v.nnz = seg_nnz;
        v.block_size = bs;
        v.ell_cols = L.width[g] * bs;
        const size_t rbs = (seg_rows + bs - 1u) / bs;
        v.blockColIdx = (unsigned int *) st->mem.d(rbs * L.width[g] * 4u);
        v.val = (csamd_half *) st->mem.d((size_t) seg_rows * v.ell_cols * vb);
        ok = v.blockColIdx != nullptr && v.val != nullptr;
        T.idx[g] = v.blockColIdx;
        T.val[g] = v.val;
        row_cursor += seg_rows;
    }
    if (ok) {
        out->segment_row_offsets[nseg] = row_cursor;
        T.begin[nseg] = (unsigned int) order.size();
        T.row0[nseg] = row_cursor;
        const unsigned int rb = (unsigned int) order.size();
        const size_t lds = (size_t) std::max(W, 1u) * 4u;
        if (rb != 0u) {
            switch (vb) {
                case 1: bell_fill_kernel<uint8_t><<<rb, EL_FILL_THREADS, lds, s>>>(*src, d_order, rb, T, out->exec_to_canonical_rows, out->canonical_to_exec_rows); break;
                case 2: bell_fill_kernel<uint16_t><<<rb, EL_FILL_THREADS, lds, s>>>(*src, d_order, rb, T, out->exec_to_canonical_rows, out->canonical_to_exec_rows); break;
                default: bell_fill_kernel<uint32_t><<<rb, EL_FILL_THREADS, lds, s>>>(*src, d_order, rb, T, out->exec_to_canonical_rows, out->canonical_to_exec_rows); break;
            }
            ok = hip_ok(hipGetLastError(), "bell_fill_kernel");
        }
        if (ok && src->cols != 0u) {
            iota_kernel<<<grid_for(src->cols, 256u, 1024u), 256, 0, s>>>(out->exec_to_canonical_cols, src->cols);
            iota_kernel<<<grid_for(src->cols, 256u, 1024u), 256, 0, s>>>(out->canonical_to_exec_cols, src->cols);
            ok = hip_ok(hipGetLastError(), "identity column maps");
        }
        if (ok) ok = hip_ok(hipStreamSynchronize(s), "bucketed blocked_ell sync");
    }
    scratch.release();
    if (!ok) {
        if (csamd_last_error()[0] == '\0') set_error("build_bucketed_blocked_ell: allocation failed");
        csamd_bucketed_blocked_ell_clear(out);
        return 0;
    }
    return 1;
}

int csamd_choose_bucket_count_sliced_ell(const csamd_sliced_ell_view *src, csamd_convert_workspace *ws, unsigned int *bucket_count,
                                         unsigned long long *bucketed_bytes) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr && bucket_count != nullptr && bucketed_bytes != nullptr,
                  "choose_bucket_count_sliced: null argument");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    owned mem;
    std::vector<unsigned int> widths, order;
    unsigned int *d_order = nullptr;
    const int ok = sell_prepare(src, ws, widths, order, &d_order, mem);
    mem.release();
    if (!ok) return 0;
    const unsigned long long total = sell_total_slots_host(src);
    CSAMD_REQUIRE(total != EL_INF, "choose_bucket_count_sliced: slice metadata copy failed");
    const unsigned long long original = sell_packed_bytes(src->slice_count, total);
    std::vector<unsigned int> sw(order.size());
    for (size_t i = 0; i < order.size(); ++i) sw[i] = widths[order[i]];
    const unsigned int maxb = std::max(1u, std::min(EL_MAX_BUCKETS, src->rows));
    unsigned int best_b = 1u;
    unsigned long long best = ~0ull;
    for (unsigned int b = 1; b <= maxb; ++b) {
        unsigned long long bytes = 0ull;
        (void) sell_bucketed(sw, b, original, &bytes);
        if (bytes < best || (bytes == best && b < best_b)) {
            best = bytes;
            best_b = b;
        }
    }
    *bucket_count = best_b;
    *bucketed_bytes = best;
    return 1;
}

int csamd_build_bucketed_sliced_ell(const csamd_sliced_ell_view *src, unsigned int value_type, unsigned int requested_bucket_count,
                                    csamd_convert_workspace *ws, csamd_bucketed_sliced_ell *out) {
    CSAMD_REQUIRE(src != nullptr && ws != nullptr && out != nullptr, "build_bucketed_sliced_ell: null argument");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "build_bucketed_sliced_ell: unsupported value type %u", value_type);
    CSAMD_REQUIRE(src->rows == 0u || src->slice_count != 0u, "build_bucketed_sliced_ell: rows without slices");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    csamd_bucketed_sliced_ell_clear(out);
    el_state *st = new el_state();
    out->storage = st;
    owned scratch;
    std::vector<unsigned int> widths, order;
    unsigned int *d_order = nullptr;
    hipStream_t s = (hipStream_t) ws->stream;
    if (!sell_prepare(src, ws, widths, order, &d_order, scratch)) {
        scratch.release();
        csamd_bucketed_sliced_ell_clear(out);
        return 0;
    }
    const unsigned long long total = sell_total_slots_host(src);
    const unsigned long long original = sell_packed_bytes(src->slice_count, total);
    std::vector<unsigned int> sw(order.size());
    for (size_t i = 0; i < order.size(); ++i) sw[i] = widths[order[i]];
    unsigned long long bytes = 0ull;
    const layout L = sell_bucketed(sw, requested_bucket_count, original, &bytes);
    const unsigned int nseg = (unsigned int) L.width.size();
    const unsigned int rows = src->rows;
    out->rows = rows;
    out->cols = src->cols;
    out->nnz = src->nnz;
    out->segment_count = nseg;
    out->bytes = bytes;
    out->canonical_slice_count = src->slice_count;
    out->segments = (csamd_sliced_ell_view *) st->mem.h(nseg * sizeof(csamd_sliced_ell_view));
    out->segment_row_offsets = (unsigned int *) st->mem.h((nseg + 1u) * sizeof(unsigned int));
    out->exec_to_canonical_rows = (unsigned int *) st->mem.d((size_t) rows * 4u);
    out->canonical_to_exec_rows = (unsigned int *) st->mem.d((size_t) rows * 4u);
    out->canonical_slice_row_offsets = (unsigned int *) st->mem.d(((size_t) src->slice_count + 1u) * 4u);
    out->canonical_slice_widths = (unsigned int *) st->mem.d((size_t) src->slice_count * 4u);
    int ok = total != EL_INF && out->segments != nullptr && out->segment_row_offsets != nullptr && out->exec_to_canonical_rows != nullptr &&
             out->canonical_to_exec_rows != nullptr && out->canonical_slice_row_offsets != nullptr && out->canonical_slice_widths != nullptr;
    if (ok) {
        ok = hip_ok(hipMemcpyAsync(out->canonical_slice_row_offsets, src->slice_row_offsets, ((size_t) src->slice_count + 1u) * 4u,
                                   hipMemcpyDeviceToDevice, s), "copy slice offsets");
        if (ok && src->slice_count != 0u)
            ok = hip_ok(hipMemcpyAsync(out->canonical_slice_widths, src->slice_widths, (size_t) src->slice_count * 4u,
                                       hipMemcpyDeviceToDevice, s), "copy slice widths");
    }
    seg_table T;
    std::memset(&T, 0, sizeof(T));
    T.count = nseg;
    for (unsigned int g = 0; ok && g < nseg; ++g) {
        const unsigned int seg_rows = L.begin[g + 1] - L.begin[g];
        unsigned int seg_nnz = 0u;
        for (unsigned int q = L.begin[g]; q < L.begin[g + 1]; ++q) seg_nnz += sw[q];
        out->segment_row_offsets[g] = L.begin[g];
        T.begin[g] = L.begin[g];
        T.row0[g] = L.begin[g];
        T.width[g] = L.width[g];
        csamd_sliced_ell_view &v = out->segments[g];
        std::memset(&v, 0, sizeof(v));
        v.rows = seg_rows;
        v.cols = src->cols;
        v.nnz = seg_nnz;
        v.slice_count = seg_rows != 0u ? 1u : 0u;
        v.slice_rows = seg_rows;
        unsigned int *meta = (unsigned int *) st->mem.d(4u * 4u);
        const size_t slots = (size_t) seg_rows * L.width[g];
        v.col_idx = (unsigned int *) st->mem.d(slots * 4u);
        v.val = (csamd_half *) st->mem.d(slots * vb);
        ok = meta != nullptr && v.col_idx != nullptr && v.val != nullptr;
        if (ok) {
            const unsigned int h[4] = {0u, seg_rows, L.width[g], 0u};  // row offsets {0, rows}, width, slot offset 0
            ok = hip_ok(hipMemcpy(meta, h, sizeof(h), hipMemcpyHostToDevice), "H2D segment slice");
            v.slice_row_offsets = meta;
            v.slice_widths = meta + 2;
            v.slice_slot_offsets = meta + 3;
        }
        T.idx[g] = v.col_idx;
        T.val[g] = v.val;
    }
    if (ok) {
        out->segment_row_offsets[nseg] = rows;
        T.begin[nseg] = rows;
        T.row0[nseg] = rows;
        if (rows != 0u) {
            const unsigned int grid = grid_for((unsigned long long) rows * 64u, 256u, 4096u);
            switch (vb) {
                case 1: sell_fill_kernel<uint8_t><<<grid, 256, 0, s>>>(*src, d_order, T, out->exec_to_canonical_rows, out->canonical_to_exec_rows); break;
                case 2: sell_fill_kernel<uint16_t><<<grid, 256, 0, s>>>(*src, d_order, T, out->exec_to_canonical_rows, out->canonical_to_exec_rows); break;
                default: sell_fill_kernel<uint32_t><<<grid, 256, 0, s>>>(*src, d_order, T, out->exec_to_canonical_rows, out->canonical_to_exec_rows); break;
            }
            ok = hip_ok(hipGetLastError(), "sell_fill_kernel");
        }
        if (ok) ok = hip_ok(hipStreamSynchronize(s), "bucketed sliced_ell sync");
    }
    scratch.release();
    if (!ok) {
        if (csamd_last_error()[0] == '\0') set_error("build_bucketed_sliced_ell: allocation failed");
        csamd_bucketed_sliced_ell_clear(out);
        return 0;
    }
    return 1;
}

} // extern "C"
