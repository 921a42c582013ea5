// Deterministic compressed transpose (CSR -> CSC or CSC -> CSR).
//
// Reference: src/convert/compressed_transpose_raw.cuh:14-122 (cuSPARSE
// csr2cscEx2, else count / scan / atomic scatter with nondeterministic order
// inside a column).  Here the output order is canonical: inside every output
// segment the source major index ascends (what csr2cscEx2 and scipy give).
//
// gfx950 plan:
//   K1 count: one 1024-thread workgroup per tile of TR source rows, a u32
//      LDS histogram over all output segments (<= 40k fit in 160 KB), one wave
//      per row; the same pass records, per row, where each column band starts
//      (rows with ascending columns) -> rowband[r][b];
//   K2 segment totals + exclusive scan -> out_ptr; K3 per-(tile, segment)
//      base offsets in place of the counts (column-major running sums);
//   K4 scatter: one wave per (tile, column band), LDS cursors for the band,
//      rows of the tile in ascending order, so every output segment receives
//      its entries in ascending source-row order.
#include "common.hh"

namespace csamd {
namespace {

constexpr unsigned int TR_ROWS = 1024u;      // source rows per tile
constexpr unsigned int TR_BAND = 4096u;      // output segments per band (16 KB of LDS cursors)
constexpr unsigned int TR_COUNT_THREADS = 1024u;

struct tr_layout {
    unsigned int tiles;
    unsigned int bands;
    size_t off_counts, off_rowband, off_coltot, off_flags, off_scan, total;
};

tr_layout make_layout(unsigned int cDim, unsigned int uDim) {
    tr_layout L;
    L.tiles = cDim == 0u ? 1u : (cDim + TR_ROWS - 1u) / TR_ROWS;
    L.bands = uDim == 0u ? 1u : (uDim + TR_BAND - 1u) / TR_BAND;
    size_t p = 0;
    L.off_flags = p;
    p = align_up(p + 256u, 256u);
    L.off_counts = p;
    p = align_up(p + (size_t) L.tiles * uDim * 4u, 256u);
    L.off_rowband = p;
    p = align_up(p + (size_t) cDim * (L.bands + 1u) * 4u, 256u);
    L.off_coltot = p;
    p = align_up(p + (size_t) (uDim + 1u) * 4u, 256u);
    L.off_scan = p;
    p = align_up(p + exclusive_scan_scratch_bytes(uDim), 256u);
    L.total = p;
    return L;
}

// K1 (LDS histogram path).
__global__ void __launch_bounds__(TR_COUNT_THREADS) tr_count_kernel(unsigned int cDim, unsigned int uDim, const unsigned int *__restrict__ ptr,
                                                                    const unsigned int *__restrict__ idx, unsigned int bands,
                                                                    int use_lds, unsigned int *__restrict__ counts,
                                                                    unsigned int *__restrict__ rowband, unsigned int *__restrict__ flags) {
    extern __shared__ unsigned int hist[];
    const unsigned int t = blockIdx.x;
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int r1 = r0 + TR_ROWS < cDim ? r0 + TR_ROWS : cDim;
    unsigned int *tile_counts = counts + (size_t) t * uDim;
    if (use_lds) {
        for (unsigned int c = threadIdx.x; c < uDim; c += blockDim.x) hist[c] = 0u;
        __syncthreads();
    }
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = lane_id();
    for (unsigned int r = r0 + wave; r < r1; r += nw) {
        const unsigned int b = ptr[r], e = ptr[r + 1u];
        unsigned int *rb = rowband + (size_t) r * (bands + 1u);
        if (b == e) {
            for (unsigned int k = lane; k <= bands; k += 64u) rb[k] = b;
            continue;
        }
        for (unsigned int i = b + lane; i < e; i += 64u) {
            const unsigned int c = idx[i];
            if (c >= uDim) {
                flags[1] = 1u;
                continue;
            }
            if (use_lds) atomicAdd(&hist[c], 1u);
            else atomicAdd(&tile_counts[c], 1u);
            const unsigned int band = c / TR_BAND;
            int prev = -1;
            if (i > b) {
                const unsigned int pc = idx[i - 1u];
                prev = pc < uDim ? (int) (pc / TR_BAND) : (int) bands;
                if (pc > c) flags[0] = 1u;  // columns not ascending
            }
            for (int k = prev + 1; k <= (int) band; ++k) rb[k] = i;
            if (i == e - 1u)
                for (unsigned int k = band + 1u; k <= bands; ++k) rb[k] = e;
        }
    }
    if (use_lds) {
        __syncthreads();
        for (unsigned int c = threadIdx.x; c < uDim; c += blockDim.x) tile_counts[c] = hist[c];
    }
}

// K2a: per-segment totals over tiles.
__global__ void tr_coltot_kernel(const unsigned int *__restrict__ counts, unsigned int tiles, unsigned int uDim,
                                 unsigned int *__restrict__ coltot) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < uDim; c += gridDim.x * blockDim.x) {
        unsigned int s = 0u;
        for (unsigned int t = 0; t < tiles; ++t) s += counts[(size_t) t * uDim + c];
        coltot[c] = s;
    }
}

// K3: counts[t][c] <- out_ptr[c] + sum_{t' < t} counts[t'][c].
__global__ void tr_offsets_kernel(unsigned int *__restrict__ counts, unsigned int tiles, unsigned int uDim,
                                  const unsigned int *__restrict__ out_ptr) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < uDim; c += gridDim.x * blockDim.x) {
        unsigned int run = out_ptr[c];
        for (unsigned int t = 0; t < tiles; ++t) {
            const size_t k = (size_t) t * uDim + c;
            const unsigned int n = counts[k];
            counts[k] = run;
            run += n;
        }
    }
}

// K4: one wave per (tile, band).  sorted != 0: rowband gives each row's band
// sub-segment; otherwise every row is scanned and filtered by band.
template<typename S>
__global__ void __launch_bounds__(64) tr_scatter_kernel(unsigned int cDim, unsigned int uDim, const unsigned int *__restrict__ ptr,
                                                        const unsigned int *__restrict__ idx, const S *__restrict__ val,
                                                        const unsigned int *__restrict__ offsets, const unsigned int *__restrict__ rowband,
                                                        unsigned int bands, int sorted, unsigned int *__restrict__ out_idx,
                                                        S *__restrict__ out_val) {
    __shared__ unsigned int cursor[TR_BAND];
    const unsigned int t = blockIdx.x / bands;
    const unsigned int band = blockIdx.x - t * bands;
    const unsigned int c0 = band * TR_BAND;
    const unsigned int c1 = c0 + TR_BAND < uDim ? c0 + TR_BAND : uDim;
    const unsigned int lane = lane_id();
    for (unsigned int k = lane; k < c1 - c0; k += 64u) cursor[k] = offsets[(size_t) t * uDim + c0 + k];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const unsigned int r0 = t * TR_ROWS;
    const unsigned int r1 = r0 + TR_ROWS < cDim ? r0 + TR_ROWS : cDim;
    for (unsigned int r = r0; r < r1; ++r) {
        unsigned int b, e;
        if (sorted) {
            const unsigned int *rb = rowband + (size_t) r * (bands + 1u);
            b = rb[band];
            e = rb[band + 1u];
        } else {
            b = ptr[r];
            e = ptr[r + 1u];
        }
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

template<typename S>
int launch_scatter(const tr_layout &L, unsigned int cDim, unsigned int uDim, const unsigned int *ptr, const unsigned int *idx,
                   const void *val, const unsigned int *offsets, const unsigned int *rowband, int sorted, unsigned int *out_idx,
                   void *out_val, hipStream_t s) {
    const unsigned long long blocks = (unsigned long long) L.tiles * L.bands;
    CSAMD_REQUIRE(blocks <= 0x7fffffffull, "transpose: too many tiles");
    tr_scatter_kernel<S><<<(unsigned int) blocks, 64, 0, s>>>(cDim, uDim, ptr, idx, (const S *) val, offsets, rowband, L.bands, sorted,
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
    unsigned int *rowband = (unsigned int *) (base + L.off_rowband);
    unsigned int *coltot = (unsigned int *) (base + L.off_coltot);
    void *scan_tmp = base + L.off_scan;
    if (!zero_bytes(flags, 256u, s)) return 0;
    int max_lds = 65536;
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) (void) hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    }
    const int use_lds = (size_t) uDim * 4u <= (size_t) max_lds ? 1 : 0;
    if (!use_lds && !zero_bytes(counts, (size_t) L.tiles * uDim * 4u, s)) return 0;
    {
        const size_t lds = use_lds ? (size_t) uDim * 4u : 0u;
        auto k = tr_count_kernel;
        if (use_lds) CSAMD_HIP(hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds), "set lds");
        k<<<L.tiles, TR_COUNT_THREADS, lds, s>>>(cDim, uDim, d_ptr, d_idx, L.bands, use_lds, counts, rowband, flags);
        CSAMD_HIP(hipGetLastError(), "tr_count_kernel");
    }
    tr_coltot_kernel<<<grid_for(uDim, 256u, 4096u), 256, 0, s>>>(counts, L.tiles, uDim, coltot);
    CSAMD_HIP(hipGetLastError(), "tr_coltot_kernel");
    if (!exclusive_scan_u32(coltot, uDim, d_out_ptr, nullptr, scan_tmp, exclusive_scan_scratch_bytes(uDim), s)) return 0;
    tr_offsets_kernel<<<grid_for(uDim, 256u, 4096u), 256, 0, s>>>(counts, L.tiles, uDim, d_out_ptr);
    CSAMD_HIP(hipGetLastError(), "tr_offsets_kernel");
    unsigned int hflags[2] = {0u, 0u};
    CSAMD_HIP(hipMemcpyAsync(hflags, flags, 8u, hipMemcpyDeviceToHost, s), "D2H transpose flags");
    CSAMD_HIP(hipStreamSynchronize(s), "transpose sync");
    CSAMD_REQUIRE(hflags[1] == 0u, "transpose: minor index out of range (>= %u)", uDim);
    const int sorted = hflags[0] == 0u ? 1 : 0;
    switch (vb) {
        case 1: return launch_scatter<uint8_t>(L, cDim, uDim, d_ptr, d_idx, d_val, counts, rowband, sorted, d_out_idx, d_out_val, s);
        case 2: return launch_scatter<uint16_t>(L, cDim, uDim, d_ptr, d_idx, d_val, counts, rowband, sorted, d_out_idx, d_out_val, s);
        default: return launch_scatter<float>(L, cDim, uDim, d_ptr, d_idx, d_val, counts, rowband, sorted, d_out_idx, d_out_val, s);
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
