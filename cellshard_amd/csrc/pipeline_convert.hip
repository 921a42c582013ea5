// Batched layout conversion for the partition pipeline: all Sliced-ELL
// partitions in three launches (slice widths, slot offsets, emit) and all
// Blocked-ELL partitions in one, over work lists that span the partitions
// (a work item finds its partition by binary search over per-partition
// bases), so the graph holds a fixed number of conversion nodes whatever
// the partition count.
//
// Layouts follow the single-partition converters: Sliced-ELL as
// sliced_ell_from_compressed (convert/sliced_ell_from_compressed.cuh:96-284:
// slice s holds rows [s S, s S + S), width = its longest row, rows laid out
// one after another, padding INVALID / 0), Blocked-ELL as
// blocked_ell_from_compressed (convert/blocked_ell_from_compressed.cuh:258-347:
// ascending unique block columns per row block, padded to the ell width with
// INVALID, values at row * ell_cols + slot * B + col % B, the last of
// repeated (row, col) entries kept, nnz = stored nonzero values).
#include "pipeline.hh"
#include "sell_write.hh"

namespace csamd {
namespace {

constexpr unsigned int PL_THREADS = 256u;

struct pl_sell {
    const unsigned int *ptr;
    const unsigned int *idx;
    const void *val;
    unsigned int rows, S, slices;
    unsigned int *offsets, *widths, *slot_offsets, *col_idx;
    void *out_val;
    unsigned long long slot_cap;
};

struct pl_bell {
    const unsigned int *ptr;
    const unsigned int *idx;
    const void *val;
    unsigned int rows, cols, lbs, W, nwords;
    unsigned int *bci;
    void *out_val;
};

// Last partition p with base[p] <= g (partitions of no work share the next one's base).
__device__ __forceinline__ unsigned int pl_find(const unsigned int *__restrict__ base, unsigned int n, unsigned int g) {
    unsigned int lo = 0u, hi = n - 1u;
    while (lo < hi) {
        const unsigned int mid = (lo + hi + 1u) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

template<typename S> __device__ __forceinline__ bool pl_nonzero(S v);
template<> __device__ __forceinline__ bool pl_nonzero<uint16_t>(uint16_t v) { return (v & 0x7fffu) != 0u; }
template<> __device__ __forceinline__ bool pl_nonzero<float>(float v) { return v != 0.0f; }
template<> __device__ __forceinline__ bool pl_nonzero<uint8_t>(uint8_t v) { return (v & 0x7fu) != 0u; }

__global__ void pl_fill_kernel(unsigned int *__restrict__ p, unsigned int n) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Slice s of its partition: row offset and width (the longest row).
__global__ void pl_sell_slices_kernel(const pl_sell *__restrict__ tab, unsigned int n, const unsigned int *__restrict__ slice_base,
                                      unsigned int total) {
    for (unsigned int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const unsigned int p = pl_find(slice_base, n, g);
        const pl_sell &P = tab[p];
        const unsigned int s = g - slice_base[p];
        const unsigned int r0 = s * P.S, r1 = r0 + P.S < P.rows ? r0 + P.S : P.rows;
        const CSAMD_GLOBAL unsigned int *ptr = gmem(P.ptr);
        unsigned int m = 0u;
        for (unsigned int r = r0; r < r1; ++r) m = max(m, ptr[r + 1u] - ptr[r]);
        P.widths[s] = m;
        P.offsets[s] = r0;
        if (s + 1u == P.slices) P.offsets[P.slices] = P.rows;
    }
}

// Slot offsets of one partition per workgroup (exclusive scan of heights x widths).
__global__ void __launch_bounds__(1024) pl_sell_scan_kernel(const pl_sell *__restrict__ tab, unsigned int n,
                                                            unsigned int *__restrict__ error) {
    __shared__ unsigned int smem[16];
    for (unsigned int p = blockIdx.x; p < n; p += gridDim.x) {
        const pl_sell &P = tab[p];
        unsigned long long carry = 0ull;
        for (unsigned int s0 = 0; s0 < P.slices; s0 += 1024u) {
            const unsigned int s = s0 + threadIdx.x;
            unsigned int c = 0u;
            if (s < P.slices) {
                const unsigned int h = min(P.S, P.rows - s * P.S);
                const unsigned long long v = (unsigned long long) h * P.widths[s];
                c = v > 0xffffffffull ? 0xffffffffu : (unsigned int) v;
            }
            unsigned int total = 0u;
            const unsigned int ex = block_exclusive_scan_u32(c, smem, &total);
            const unsigned long long o = carry + ex;
            if (s < P.slices) {
                P.slot_offsets[s] = (unsigned int) o;
                if (o > 0xffffffffull) atomicOr(error, 2u);
            }
            carry += total;
            __syncthreads();
        }
        if (threadIdx.x == 0u && carry > P.slot_cap) atomicOr(error, 2u);
    }
}

// One wave per row: its entries, then padding to the slice width.
template<typename S>
__global__ void pl_sell_emit_kernel(const pl_sell *__restrict__ tab, unsigned int n, const unsigned int *__restrict__ row_base,
                                    unsigned int total_rows) {
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (unsigned int g = blockIdx.x * nw + wave; g < total_rows; g += gridDim.x * nw) {
        const unsigned int p = pl_find(row_base, n, g);
        const pl_sell &P = tab[p];
        const unsigned int r = g - row_base[p];
        const unsigned int s = r / P.S;
        const unsigned int width = P.widths[s];
        const unsigned long long base = (unsigned long long) P.slot_offsets[s] + (unsigned long long) (r - s * P.S) * width;
        const unsigned int b = gmem(P.ptr)[r], e = gmem(P.ptr)[r + 1u];
        // slots past the capacity are not written (over capacity: flagged by the scan)
        sell_write_row<S>(gmem(P.idx), gmem((const S *) P.val), b, min(e - b, width), width, base, gmem(P.col_idx), gmem((S *) P.out_val),
                          sell_quads_ok<S>(P.col_idx, P.out_val), P.slot_cap);
    }
}

// One workgroup per row block: its block columns (LDS bitset, prefix
// popcounts), the padded value rows, then the entries (one wave per row).
// A row whose columns do not strictly ascend is replayed in order by one
// lane (the last repeat wins) and its nonzeros counted from the written row.
template<typename S>
__global__ void __launch_bounds__(PL_THREADS) pl_bell_kernel(const pl_bell *__restrict__ tab, unsigned int n,
                                                             const unsigned int *__restrict__ rb_base, unsigned int total,
                                                             unsigned int *__restrict__ counters, unsigned int *__restrict__ error) {
    extern __shared__ unsigned int lds[];
    __shared__ unsigned int smem[16];
    const unsigned int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6, nwaves = PL_THREADS / 64u;
    for (unsigned int g = blockIdx.x; g < total; g += gridDim.x) {
        const unsigned int p = pl_find(rb_base, n, g);
        const pl_bell &P = tab[p];
        const unsigned int rb = g - rb_base[p], lbs = P.lbs, bs = 1u << lbs, nwords = P.nwords, W = P.W;
        const unsigned int r0 = rb << lbs, r1 = min(P.rows, r0 + bs);
        const unsigned int ell = W << lbs;
        unsigned int *bits = lds, *prefix = lds + nwords;
        const CSAMD_GLOBAL unsigned int *ptr = gmem(P.ptr);
        const CSAMD_GLOBAL unsigned int *idx = gmem(P.idx);
        const CSAMD_GLOBAL S *val = gmem((const S *) P.val);
        CSAMD_GLOBAL S *out = gmem((S *) P.out_val);
        for (unsigned int w = tid; w < nwords; w += PL_THREADS) bits[w] = 0u;
        __syncthreads();
        const unsigned int eb = ptr[r0], ee = ptr[r1];
        // the row block's entries, PL_BITS_U per thread per step with the
        // loads issued before the atomics (one memory round trip per step)
        constexpr unsigned int PL_BITS_U = 16u;
        for (unsigned int i0 = eb; i0 < ee; i0 += PL_THREADS * PL_BITS_U) {
            unsigned int c[PL_BITS_U];
#pragma unroll
            for (unsigned int u = 0; u < PL_BITS_U; ++u) {
                const unsigned int i = i0 + u * PL_THREADS + tid;
                c[u] = i < ee ? idx[i] : CSAMD_INVALID_COL;
            }
#pragma unroll
            for (unsigned int u = 0; u < PL_BITS_U; ++u)
                if (c[u] < P.cols) atomicOr(&bits[(c[u] >> lbs) >> 5], 1u << ((c[u] >> lbs) & 31u));
        }
        __syncthreads();
        unsigned int carry = 0u;
        for (unsigned int w0 = 0; w0 < nwords; w0 += PL_THREADS) {
            const unsigned int w = w0 + tid;
            unsigned int total_w = 0u;
            const unsigned int ex = block_exclusive_scan_u32(w < nwords ? (unsigned int) __popc(bits[w]) : 0u, smem, &total_w);
            if (w < nwords) prefix[w] = carry + ex;
            carry += total_w;
            __syncthreads();
        }
        const unsigned int count = carry;
        if (tid == 0u && count > W) atomicOr(error, 1u);
        CSAMD_GLOBAL unsigned int *bci = gmem(P.bci) + (size_t) rb * W;
        for (unsigned int w = tid; w < nwords; w += PL_THREADS) {
            unsigned int m = bits[w], k = prefix[w];
            while (m != 0u) {
                const unsigned int b = (unsigned int) __ffs(m) - 1u;
                if (k < W) bci[k] = w * 32u + b;
                ++k;
                m &= m - 1u;
            }
        }
        for (unsigned int k = count + tid; k < W; k += PL_THREADS) bci[k] = CSAMD_INVALID_COL;
        const unsigned long long zb = (unsigned long long) r0 * ell, zn = (unsigned long long) (r1 - r0) * ell;
        constexpr unsigned int Z = 16u / sizeof(S);  // values per 16-byte store
        if (((zb * sizeof(S)) & 15u) == 0u && ((uintptr_t) P.out_val & 15u) == 0u) {
            CSAMD_GLOBAL u32x4 *z4 = (CSAMD_GLOBAL u32x4 *) (out + zb);
            const unsigned long long nz4 = zn / Z;
            for (unsigned long long i = tid; i < nz4; i += PL_THREADS) z4[i] = u32x4{0u, 0u, 0u, 0u};
            for (unsigned long long i = nz4 * Z + tid; i < zn; i += PL_THREADS) out[zb + i] = S(0);
        } else {
            for (unsigned long long i = tid; i < zn; i += PL_THREADS) out[zb + i] = S(0);
        }
        __syncthreads();  // zeros written before the entries
        unsigned int nz = 0u;
        for (unsigned int r = r0 + wave; r < r1; r += nwaves) {
            const unsigned int b = ptr[r], e = ptr[r + 1u];
            const unsigned long long ro = (unsigned long long) r * ell;
            auto pos_of = [&](unsigned int c, unsigned int *pos) -> bool {
                if (c >= P.cols) return false;
                const unsigned int blk = c >> lbs;
                const unsigned int slot = prefix[blk >> 5] + (unsigned int) __popc(bits[blk >> 5] & ((1u << (blk & 31u)) - 1u));
                if (slot >= W) return false;
                *pos = (slot << lbs) + (c & (bs - 1u));
                return true;
            };
            // Scatter, assuming the row strictly ascends, PL_U entries per
            // lane per step with every load issued before any use (one
            // memory round trip per step, not per entry); a step that finds
            // a column not above the previous one marks the row for the
            // ordered replay below, which rewrites all of its entries.
            constexpr unsigned int PL_U = 16u;
            bool simple = true;
            unsigned int rnz = 0u;
            for (unsigned int i0 = b; i0 < e; i0 += 64u * PL_U) {
                unsigned int c[PL_U], cn[PL_U];
                S v[PL_U];
#pragma unroll
                for (unsigned int u = 0; u < PL_U; ++u) {
                    const unsigned int i = i0 + u * 64u + lane;
                    c[u] = i < e ? idx[i] : CSAMD_INVALID_COL;
                    cn[u] = i + 1u < e ? idx[i + 1u] : CSAMD_INVALID_COL;
                    v[u] = i < e ? val[i] : S(0);
                }
                bool bad = false;
#pragma unroll
                for (unsigned int u = 0; u < PL_U; ++u) {
                    const unsigned int i = i0 + u * 64u + lane;
                    bad |= i + 1u < e && c[u] >= cn[u];
                    unsigned int pos;
                    if (i < e && pos_of(c[u], &pos)) {
                        out[ro + pos] = v[u];
                        rnz += pl_nonzero<S>(v[u]) ? 1u : 0u;
                    }
                }
                if (__ballot(bad) != 0ull) simple = false;
            }
            if (simple) {
                nz += rnz;
            } else {
                if (lane == 0u)
                    for (unsigned int i = b; i < e; ++i) {
                        unsigned int pos;
                        if (pos_of(idx[i], &pos)) out[ro + pos] = val[i];
                    }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                for (unsigned int k = lane; k < ell; k += 64u) nz += pl_nonzero<S>(out[ro + k]) ? 1u : 0u;
            }
        }
        nz = wave_sum_u32(nz);
        if (lane == 0u && nz != 0u) atomicAdd(&counters[p], nz);
        __syncthreads();  // LDS reused by the next row block
    }
}

__global__ void pl_bell_nnz_kernel(const unsigned int *__restrict__ counters, unsigned int *const *__restrict__ outs, unsigned int n) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (outs[i] != nullptr) *gmem(outs[i]) = counters[i];
}

template<typename S> int pl_launch(const pl_state &st, hipStream_t s) {
    const unsigned int cu = (unsigned int) device_cu_count();
    if (st.n_sell != 0u) {
        const pl_sell *tab = (const pl_sell *) st.sell_tab;
        if (st.sell_slices != 0u) {
            pl_sell_slices_kernel<<<grid_for(st.sell_slices, PL_THREADS, cu * 8u), PL_THREADS, 0, s>>>(tab, st.n_sell, st.sell_slice_base,
                                                                                                      st.sell_slices);
            CSAMD_HIP(hipGetLastError(), "pl_sell_slices_kernel");
        }
        pl_sell_scan_kernel<<<st.n_sell, 1024, 0, s>>>(tab, st.n_sell, st.error);
        CSAMD_HIP(hipGetLastError(), "pl_sell_scan_kernel");
        if (st.sell_rows != 0u) {
            pl_sell_emit_kernel<S><<<grid_for((st.sell_rows + 3u) / 4u, 1u, cu * 32u), PL_THREADS, 0, s>>>(tab, st.n_sell, st.sell_row_base,
                                                                                                           st.sell_rows);
            CSAMD_HIP(hipGetLastError(), "pl_sell_emit_kernel");
        }
    }
    if (st.n_bell != 0u) {
        if (st.bell_row_blocks != 0u) {
            const size_t lds = (size_t) 2u * st.bell_nwords * 4u;
            pl_bell_kernel<S><<<grid_for(st.bell_row_blocks, 1u, cu * 8u), PL_THREADS, lds, s>>>(
                (const pl_bell *) st.bell_tab, st.n_bell, st.bell_rb_base, st.bell_row_blocks, st.bell_counters, st.error);
            CSAMD_HIP(hipGetLastError(), "pl_bell_kernel");
        }
        pl_bell_nnz_kernel<<<grid_for(st.n_bell, PL_THREADS, 64u), PL_THREADS, 0, s>>>(st.bell_counters, st.bell_nnz_out, st.n_bell);
        CSAMD_HIP(hipGetLastError(), "pl_bell_nnz_kernel");
    }
    return 1;
}

} // namespace

int pl_prepare(const csamd_pipeline_partition *parts, unsigned int n, unsigned int value_type, pl_state *st) {
    std::memset(st, 0, sizeof(*st));
    st->value_type = value_type;
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "pipeline: value type %u", value_type);
    std::vector<pl_sell> sells;
    std::vector<pl_bell> bells;
    std::vector<unsigned int> sbase(1, 0u), rbase(1, 0u), bbase(1, 0u);
    std::vector<unsigned int *> bnnz;
    unsigned long long slices = 0ull, rows = 0ull, rbs = 0ull;
    for (unsigned int i = 0; i < n; ++i) {
        const csamd_pipeline_partition &P = parts[i];
        const csamd_compressed_view &c = P.compressed;
        if (P.layout != CSAMD_LAYOUT_SLICED_ELL && P.layout != CSAMD_LAYOUT_BLOCKED_ELL) continue;
        CSAMD_REQUIRE(c.axis == CSAMD_COMPRESSED_BY_ROW && (c.rows == 0u || c.majorPtr != nullptr) &&
                          (c.nnz == 0u || (c.minorIdx != nullptr && c.val != nullptr)),
                      "pipeline: partition %u needs its device CSR", i);
        if (P.layout == CSAMD_LAYOUT_SLICED_ELL) {
            const csamd_sliced_ell_view &v = P.sliced_ell;
            CSAMD_REQUIRE(c.rows != 0u && P.slice_rows != 0u, "pipeline: Sliced-ELL partition %u needs rows and a slice height", i);
            const unsigned int S = P.slice_rows < c.rows ? P.slice_rows : c.rows;
            const unsigned int ns = (c.rows + S - 1u) / S;
            CSAMD_REQUIRE(v.slice_count == ns && v.slice_rows == S && v.rows == c.rows && v.slice_row_offsets != nullptr &&
                              v.slice_widths != nullptr && v.slice_slot_offsets != nullptr &&
                              (P.slots_capacity == 0ull || (v.col_idx != nullptr && v.val != nullptr)),
                          "pipeline: Sliced-ELL partition %u: view of %u slices of %u rows expected with its arrays", i, ns, S);
            sells.push_back(pl_sell{c.majorPtr, c.minorIdx, c.val, c.rows, S, ns, v.slice_row_offsets, v.slice_widths,
                                    v.slice_slot_offsets, v.col_idx, v.val, P.slots_capacity});
            slices += ns;
            rows += c.rows;
            CSAMD_REQUIRE(slices <= 0xffffffffull && rows <= 0xffffffffull, "pipeline: Sliced-ELL work beyond 2^32");
            sbase.push_back((unsigned int) slices);
            rbase.push_back((unsigned int) rows);
        } else {
            const csamd_blocked_ell_view &v = P.blocked_ell;
            const unsigned int bs = P.block_size;
            CSAMD_REQUIRE(bs == 8u || bs == 16u || bs == 32u, "pipeline: Blocked-ELL partition %u block size %u (8, 16 or 32)", i, bs);
            const unsigned int lbs = bs == 8u ? 3u : bs == 16u ? 4u : 5u;
            const unsigned int W = P.ell_width;
            const unsigned int nrb = (c.rows + bs - 1u) >> lbs;
            CSAMD_REQUIRE(v.block_size == bs && v.ell_cols == W * bs && v.rows == c.rows &&
                              (W == 0u || nrb == 0u || (v.blockColIdx != nullptr && v.val != nullptr)),
                          "pipeline: Blocked-ELL partition %u: view of block %u and ell_cols %u expected with its arrays", i, bs, W * bs);
            const unsigned int nwords = (((c.cols + bs - 1u) >> lbs) + 31u) / 32u;
            CSAMD_REQUIRE((size_t) 2u * nwords * 4u <= 60u * 1024u, "pipeline: Blocked-ELL partition %u has too many block columns", i);
            bells.push_back(pl_bell{c.majorPtr, c.minorIdx, c.val, c.rows, c.cols, lbs, W, nwords, v.blockColIdx, v.val});
            bnnz.push_back(P.d_nnz);
            if (nwords > st->bell_nwords) st->bell_nwords = nwords;
            rbs += nrb;
            CSAMD_REQUIRE(rbs <= 0xffffffffull, "pipeline: Blocked-ELL work beyond 2^32");
            bbase.push_back((unsigned int) rbs);
        }
    }
    st->n_sell = (unsigned int) sells.size();
    st->n_bell = (unsigned int) bells.size();
    st->sell_slices = (unsigned int) slices;
    st->sell_rows = (unsigned int) rows;
    st->bell_row_blocks = (unsigned int) rbs;
    // one block: tables, bases, counters, error word
    const size_t o_sell = 0u, o_bell = align_up(o_sell + sells.size() * sizeof(pl_sell), 256u);
    const size_t o_sb = align_up(o_bell + bells.size() * sizeof(pl_bell), 256u);
    const size_t o_rb = align_up(o_sb + sbase.size() * 4u, 256u), o_bb = align_up(o_rb + rbase.size() * 4u, 256u);
    const size_t o_cnt = align_up(o_bb + bbase.size() * 4u, 256u), o_out = align_up(o_cnt + (bells.size() + 1u) * 4u, 256u);
    const size_t o_err = align_up(o_out + bnnz.size() * sizeof(void *), 256u);
    st->bytes = o_err + 256u;
    CSAMD_HIP(hipMalloc(&st->block, st->bytes), "hipMalloc pipeline tables");
    char *d = (char *) st->block;
    st->sell_tab = d + o_sell;
    st->bell_tab = d + o_bell;
    st->sell_slice_base = (unsigned int *) (d + o_sb);
    st->sell_row_base = (unsigned int *) (d + o_rb);
    st->bell_rb_base = (unsigned int *) (d + o_bb);
    st->bell_counters = (unsigned int *) (d + o_cnt);
    st->bell_nnz_out = (unsigned int **) (d + o_out);
    st->error = (unsigned int *) (d + o_err);
    std::vector<char> h(st->bytes, 0);
    if (!sells.empty()) std::memcpy(h.data() + o_sell, sells.data(), sells.size() * sizeof(pl_sell));
    if (!bells.empty()) std::memcpy(h.data() + o_bell, bells.data(), bells.size() * sizeof(pl_bell));
    std::memcpy(h.data() + o_sb, sbase.data(), sbase.size() * 4u);
    std::memcpy(h.data() + o_rb, rbase.data(), rbase.size() * 4u);
    std::memcpy(h.data() + o_bb, bbase.data(), bbase.size() * 4u);
    if (!bnnz.empty()) std::memcpy(h.data() + o_out, bnnz.data(), bnnz.size() * sizeof(void *));
    if (!hip_ok(hipMemcpy(st->block, h.data(), st->bytes, hipMemcpyHostToDevice), "H2D pipeline tables")) {
        pl_free(st);
        return 0;
    }
    return 1;
}

int pl_enqueue(const pl_state &st, hipStream_t s) {
    // counters and the error word cleared by a kernel (graph-safe)
    const unsigned int nclear = (unsigned int) (((char *) st.bell_nnz_out - (char *) st.bell_counters) / 4u);
    pl_fill_kernel<<<1, 256, 0, s>>>(st.bell_counters, nclear);
    CSAMD_HIP(hipGetLastError(), "pl_fill_kernel");
    pl_fill_kernel<<<1, 64, 0, s>>>(st.error, 1u);
    CSAMD_HIP(hipGetLastError(), "pl_fill_kernel");
    switch (value_type_bytes(st.value_type)) {
        case 1: return pl_launch<uint8_t>(st, s);
        case 2: return pl_launch<uint16_t>(st, s);
        default: return pl_launch<float>(st, s);
    }
}

void pl_free(pl_state *st) {
    if (st->block != nullptr) (void) hipFree(st->block);
    std::memset(st, 0, sizeof(*st));
}

} // namespace csamd
