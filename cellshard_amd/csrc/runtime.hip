// Runtime support: errors, workspaces, device-wide scans and stable radix sort.
#include "common.hh"

#include <cstdarg>
#include <mutex>

namespace csamd {

namespace {
thread_local char g_last_error[1024] = {0};
}

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_last_error[0] = '\0'; }

int ensure_device(int device) {
    int current = -1;
    if (hipGetDevice(&current) != hipSuccess) {
        set_error("hipGetDevice failed");
        return 0;
    }
    if (current == device) return 1;
    return hip_ok(hipSetDevice(device), "hipSetDevice");
}

int device_cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) cu = 256;
        cached[dev] = cu;
    }
    return cached[dev];
}

void *convert_scratch(csamd_convert_workspace *ws, size_t bytes) {
    if (ws == nullptr) {
        set_error("convert workspace is null");
        return nullptr;
    }
    if (bytes == 0u) bytes = 256u;
    if (bytes <= ws->scratch_bytes && ws->scratch != nullptr) return ws->scratch;
    if (ws->scratch != nullptr) {
        (void) hipFree(ws->scratch);
        ws->scratch = nullptr;
        ws->scratch_bytes = 0u;
    }
    bytes = align_up(bytes, 1u << 20);
    if (!hip_ok(hipMalloc(&ws->scratch, bytes), "hipMalloc convert scratch")) return nullptr;
    ws->scratch_bytes = bytes;
    return ws->scratch;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (reduce-then-scan).
// ---------------------------------------------------------------------------
namespace {

constexpr unsigned int SCAN_THREADS = 256u;
constexpr unsigned int SCAN_ITEMS = 16u;
constexpr unsigned int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
static_assert(SCAN_TILE == CSAMD_SCAN_TILE, "scan tile size");

// Each thread owns SCAN_ITEMS consecutive items of the tile.
__device__ __forceinline__ void load_items(const unsigned int *__restrict__ in, unsigned long long n, unsigned long long first,
                                           unsigned int v[SCAN_ITEMS]) {
#pragma unroll
    for (unsigned int i = 0; i < SCAN_ITEMS; ++i) v[i] = first + i < n ? in[first + i] : 0u;
}

__global__ void __launch_bounds__(SCAN_THREADS) scan_tile_sums_kernel(const unsigned int *__restrict__ in, unsigned int n,
                                                                      unsigned long long *__restrict__ tile_sums) {
    __shared__ unsigned long long red[SCAN_THREADS / 64u];
    const unsigned long long first = (unsigned long long) blockIdx.x * SCAN_TILE + (unsigned long long) threadIdx.x * SCAN_ITEMS;
    unsigned int v[SCAN_ITEMS];
    load_items(in, n, first, v);
    unsigned long long s = 0ull;
#pragma unroll
    for (unsigned int i = 0; i < SCAN_ITEMS; ++i) s += v[i];
    s = wave_sum_u64(s);
    if (lane_id() == 0u) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0ull;
        for (unsigned int w = 0; w < SCAN_THREADS / 64u; ++w) t += red[w];
        tile_sums[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of tile sums in place, total into tile_sums[count].
__global__ void __launch_bounds__(1024) scan_tile_offsets_kernel(unsigned long long *__restrict__ tile_sums, unsigned int count) {
    __shared__ unsigned long long smem[16];
    unsigned long long carry = 0ull;  // same value in every thread
    for (unsigned int start = 0; start < count; start += blockDim.x) {
        const unsigned int i = start + threadIdx.x;
        const unsigned long long v = i < count ? tile_sums[i] : 0ull;
        unsigned long long total = 0ull;
        const unsigned long long ex = block_exclusive_scan_u64(v, smem, &total);
        if (i < count) tile_sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) tile_sums[count] = carry;
}

template<typename OutT>
__global__ void __launch_bounds__(SCAN_THREADS) scan_apply_kernel(const unsigned int *__restrict__ in, unsigned int n,
                                                                  const unsigned long long *__restrict__ tile_offsets,
                                                                  unsigned int tile_count, OutT *__restrict__ out,
                                                                  unsigned int *__restrict__ overflow, int write_total) {
    __shared__ unsigned long long smem[SCAN_THREADS / 64u];
    const unsigned long long first = (unsigned long long) blockIdx.x * SCAN_TILE + (unsigned long long) threadIdx.x * SCAN_ITEMS;
    unsigned int v[SCAN_ITEMS];
    load_items(in, n, first, v);
    unsigned long long own = 0ull;
#pragma unroll
    for (unsigned int i = 0; i < SCAN_ITEMS; ++i) own += v[i];
    unsigned long long total = 0ull;
    unsigned long long o = tile_offsets[blockIdx.x] + block_exclusive_scan_u64(own, smem, &total);
    bool over = false;
#pragma unroll
    for (unsigned int i = 0; i < SCAN_ITEMS; ++i) {
        if (first + i < n) {
            out[first + i] = (OutT) o;
            over |= o > 0xffffffffull;
        }
        o += v[i];
    }
    if (over && overflow != nullptr) *overflow = 1u;
    if (write_total && blockIdx.x == tile_count - 1u && threadIdx.x == 0) {
        const unsigned long long t = tile_offsets[tile_count];
        out[n] = (OutT) t;
        if (overflow != nullptr && t > 0xffffffffull) *overflow = 1u;
    }
}

} // namespace

size_t exclusive_scan_scratch_bytes(unsigned int n) {
    const unsigned int tiles = (n + SCAN_TILE - 1u) / SCAN_TILE + 1u;
    return align_up((size_t) (tiles + 1u) * sizeof(unsigned long long), 256u);
}

template<typename OutT>
static int exclusive_scan_impl(const unsigned int *d_in, unsigned int n, OutT *d_out, unsigned int *d_overflow,
                               void *scratch, size_t scratch_bytes, hipStream_t stream, bool write_total) {
    const unsigned int tiles = n == 0u ? 1u : (n + SCAN_TILE - 1u) / SCAN_TILE;
    CSAMD_REQUIRE(scratch != nullptr && scratch_bytes >= exclusive_scan_scratch_bytes(n), "scan scratch too small");
    unsigned long long *tile_sums = (unsigned long long *) scratch;
    scan_tile_sums_kernel<<<tiles, SCAN_THREADS, 0, stream>>>(d_in, n, tile_sums);
    CSAMD_HIP(hipGetLastError(), "scan_tile_sums_kernel");
    scan_tile_offsets_kernel<<<1, 1024, 0, stream>>>(tile_sums, tiles);
    CSAMD_HIP(hipGetLastError(), "scan_tile_offsets_kernel");
    scan_apply_kernel<OutT><<<tiles, SCAN_THREADS, 0, stream>>>(d_in, n, tile_sums, tiles, d_out, d_overflow, write_total ? 1 : 0);
    CSAMD_HIP(hipGetLastError(), "scan_apply_kernel");
    return 1;
}

int scan_tile_offsets_u32(const unsigned int *d_in, unsigned int n, unsigned long long *d_tile_offsets, hipStream_t stream) {
    const unsigned int tiles = n == 0u ? 1u : (n + SCAN_TILE - 1u) / SCAN_TILE;
    scan_tile_sums_kernel<<<tiles, SCAN_THREADS, 0, stream>>>(d_in, n, d_tile_offsets);
    CSAMD_HIP(hipGetLastError(), "scan_tile_sums_kernel");
    scan_tile_offsets_kernel<<<1, 1024, 0, stream>>>(d_tile_offsets, tiles);
    CSAMD_HIP(hipGetLastError(), "scan_tile_offsets_kernel");
    return 1;
}

int exclusive_scan_u32_to_u64(const unsigned int *d_in, unsigned int n, unsigned long long *d_out, void *scratch,
                              size_t scratch_bytes, hipStream_t stream) {
    return exclusive_scan_impl<unsigned long long>(d_in, n, d_out, nullptr, scratch, scratch_bytes, stream, true);
}

int exclusive_scan_u32(const unsigned int *d_in, unsigned int n, unsigned int *d_out, unsigned int *d_overflow,
                       void *scratch, size_t scratch_bytes, hipStream_t stream, bool write_total) {
    return exclusive_scan_impl<unsigned int>(d_in, n, d_out, d_overflow, scratch, scratch_bytes, stream, write_total);
}

namespace {
__global__ void zero_bytes_kernel(unsigned char *__restrict__ p, size_t n) {
    const size_t head = (16u - ((uintptr_t) p & 15u)) & 15u;
    const size_t h = head < n ? head : n;
    const size_t tid = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
    const size_t stride = (size_t) gridDim.x * blockDim.x;
    for (size_t i = tid; i < h; i += stride) p[i] = 0u;
    uint4 *q = (uint4 *) (p + h);
    const size_t nv = (n - h) / 16u;
    for (size_t i = tid; i < nv; i += stride) q[i] = make_uint4(0u, 0u, 0u, 0u);
    for (size_t i = h + nv * 16u + tid; i < n; i += stride) p[i] = 0u;
}
} // namespace

int zero_bytes(void *p, size_t n, hipStream_t stream) {
    if (n == 0u) return 1;
    const unsigned long long blocks = (n / 16u + 255u) / 256u;
    const unsigned int cap = (unsigned int) device_cu_count() * 16u;
    zero_bytes_kernel<<<blocks < 1ull ? 1u : (blocks > cap ? cap : (unsigned int) blocks), 256, 0, stream>>>((unsigned char *) p, n);
    return hip_ok(hipGetLastError(), "zero_bytes_kernel");
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.
// Histogram layout hist[digit * tiles + tile] so one exclusive scan yields
// digit-major, tile-minor offsets (stability across tiles).  Inside a tile the
// rank follows element order: chunks of 256 in order, and inside a chunk the
// wave-level peer mask from 8 ballots plus per-wave digit counts.
// ---------------------------------------------------------------------------
namespace {

constexpr unsigned int RS_THREADS = 256u;
constexpr unsigned int RS_CHUNKS = 16u;
constexpr unsigned int RS_TILE = RS_THREADS * RS_CHUNKS;

template<typename K>
__global__ void __launch_bounds__(RS_THREADS) radix_hist_kernel(const K *__restrict__ keys, unsigned int n, unsigned int shift,
                                                                unsigned int tiles, unsigned int *__restrict__ hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long base = (unsigned long long) blockIdx.x * RS_TILE;
    for (unsigned int c = 0; c < RS_CHUNKS; ++c) {
        const unsigned long long idx = base + (unsigned long long) c * RS_THREADS + threadIdx.x;
        if (idx < n) {
            const unsigned int d = (unsigned int) ((keys[idx] >> shift) & 0xffu);
            atomicAdd(&h[d], 1u);
        }
    }
    __syncthreads();
    hist[(unsigned long long) threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

template<typename K>
__global__ void __launch_bounds__(RS_THREADS) radix_scatter_kernel(const K *__restrict__ keys_in, K *__restrict__ keys_out,
                                                                   const unsigned int *__restrict__ vals_in,
                                                                   unsigned int *__restrict__ vals_out, unsigned int n,
                                                                   unsigned int shift, unsigned int tiles,
                                                                   const unsigned int *__restrict__ offsets) {
    __shared__ unsigned int running[256];
    __shared__ unsigned int wcnt[RS_THREADS / 64u][256];
    const unsigned int wave = threadIdx.x >> 6;
    running[threadIdx.x] = offsets[(unsigned long long) threadIdx.x * tiles + blockIdx.x];
    for (unsigned int w = 0; w < RS_THREADS / 64u; ++w) wcnt[w][threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long base = (unsigned long long) blockIdx.x * RS_TILE;
    for (unsigned int c = 0; c < RS_CHUNKS; ++c) {
        const unsigned long long idx = base + (unsigned long long) c * RS_THREADS + threadIdx.x;
        const bool valid = idx < n;
        K key = 0;
        unsigned int val = 0u;
        unsigned int d = 0u;
        if (valid) {
            key = keys_in[idx];
            val = vals_in != nullptr ? vals_in[idx] : 0u;
            d = (unsigned int) ((key >> shift) & 0xffu);
        }
        unsigned long long peers = __ballot(valid);
        for (unsigned int b = 0; b < 8u; ++b) {
            const unsigned long long bal = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const unsigned int rank_in_wave = (unsigned int) __popcll(peers & lanemask_lt());
        if (valid && rank_in_wave == 0u) wcnt[wave][d] = (unsigned int) __popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned int before = 0u;
            for (unsigned int w = 0; w < wave; ++w) before += wcnt[w][d];
            const unsigned int pos = running[d] + before + rank_in_wave;
            keys_out[pos] = key;
            if (vals_out != nullptr) vals_out[pos] = val;
        }
        __syncthreads();
        {
            unsigned int add = 0u;
            for (unsigned int w = 0; w < RS_THREADS / 64u; ++w) {
                add += wcnt[w][threadIdx.x];
                wcnt[w][threadIdx.x] = 0u;
            }
            running[threadIdx.x] += add;
        }
        __syncthreads();
    }
}

template<typename K>
size_t radix_scratch_bytes(unsigned int n) {
    const unsigned int tiles = n == 0u ? 1u : (n + RS_TILE - 1u) / RS_TILE;
    size_t bytes = 0u;
    bytes += align_up((size_t) n * sizeof(K), 256u);                 // keys tmp
    bytes += align_up((size_t) n * sizeof(unsigned int), 256u);      // vals tmp
    bytes += align_up((size_t) 256u * tiles * sizeof(unsigned int), 256u);       // hist
    bytes += align_up((size_t) (256u * tiles + 1u) * sizeof(unsigned int), 256u); // offsets
    bytes += exclusive_scan_scratch_bytes(256u * tiles);
    return bytes;
}

template<typename K>
int radix_sort_impl(const K *d_keys_in, K *d_keys_out, const unsigned int *d_vals_in, unsigned int *d_vals_out,
                    unsigned int n, unsigned int key_bits, void *scratch, size_t scratch_bytes, hipStream_t stream) {
    if (n == 0u) return 1;
    CSAMD_REQUIRE(scratch != nullptr && scratch_bytes >= radix_scratch_bytes<K>(n), "radix sort scratch too small");
    const unsigned int tiles = (n + RS_TILE - 1u) / RS_TILE;
    char *p = (char *) scratch;
    K *keys_tmp = (K *) p;
    p += align_up((size_t) n * sizeof(K), 256u);
    unsigned int *vals_tmp = (unsigned int *) p;
    p += align_up((size_t) n * sizeof(unsigned int), 256u);
    unsigned int *hist = (unsigned int *) p;
    p += align_up((size_t) 256u * tiles * sizeof(unsigned int), 256u);
    unsigned int *offsets = (unsigned int *) p;
    p += align_up((size_t) (256u * tiles + 1u) * sizeof(unsigned int), 256u);
    void *scan_tmp = p;
    const size_t scan_bytes = exclusive_scan_scratch_bytes(256u * tiles);

    unsigned int passes = (key_bits + 7u) / 8u;
    if (passes > sizeof(K)) passes = sizeof(K);
    if (passes == 0u) {
        CSAMD_HIP(hipMemcpyAsync(d_keys_out, d_keys_in, (size_t) n * sizeof(K), hipMemcpyDeviceToDevice, stream), "radix copy keys");
        if (d_vals_out != nullptr && d_vals_in != nullptr)
            CSAMD_HIP(hipMemcpyAsync(d_vals_out, d_vals_in, (size_t) n * sizeof(unsigned int), hipMemcpyDeviceToDevice, stream), "radix copy vals");
        return 1;
    }
    const K *src_k = d_keys_in;
    const unsigned int *src_v = d_vals_in;
    for (unsigned int pass = 0; pass < passes; ++pass) {
        const bool to_out = ((passes - 1u - pass) & 1u) == 0u;
        K *dst_k = to_out ? d_keys_out : keys_tmp;
        unsigned int *dst_v = d_vals_out != nullptr ? (to_out ? d_vals_out : vals_tmp) : nullptr;
        const unsigned int shift = pass * 8u;
        radix_hist_kernel<K><<<tiles, RS_THREADS, 0, stream>>>(src_k, n, shift, tiles, hist);
        CSAMD_HIP(hipGetLastError(), "radix_hist_kernel");
        if (!exclusive_scan_u32(hist, 256u * tiles, offsets, nullptr, scan_tmp, scan_bytes, stream)) return 0;
        radix_scatter_kernel<K><<<tiles, RS_THREADS, 0, stream>>>(src_k, dst_k, src_v, dst_v, n, shift, tiles, offsets);
        CSAMD_HIP(hipGetLastError(), "radix_scatter_kernel");
        src_k = dst_k;
        src_v = dst_v;
    }
    return 1;
}

} // namespace

size_t radix_sort_pairs_u32_scratch_bytes(unsigned int n) { return radix_scratch_bytes<unsigned int>(n); }
size_t radix_sort_pairs_u64_scratch_bytes(unsigned int n) { return radix_scratch_bytes<unsigned long long>(n); }

int radix_sort_pairs_u32(const unsigned int *d_keys_in, unsigned int *d_keys_out, const unsigned int *d_vals_in,
                         unsigned int *d_vals_out, unsigned int n, unsigned int key_bits, void *scratch,
                         size_t scratch_bytes, hipStream_t stream) {
    return radix_sort_impl<unsigned int>(d_keys_in, d_keys_out, d_vals_in, d_vals_out, n, key_bits, scratch, scratch_bytes, stream);
}

int radix_sort_pairs_u64(const unsigned long long *d_keys_in, unsigned long long *d_keys_out, const unsigned int *d_vals_in,
                         unsigned int *d_vals_out, unsigned int n, unsigned int key_bits, void *scratch,
                         size_t scratch_bytes, hipStream_t stream) {
    return radix_sort_impl<unsigned long long>(d_keys_in, d_keys_out, d_vals_in, d_vals_out, n, key_bits, scratch, scratch_bytes, stream);
}

} // namespace csamd

// ---------------------------------------------------------------------------
namespace csamd {
namespace {
template<int VT>
__global__ void decode_values_kernel(const typename value_traits<VT>::storage *__restrict__ in, unsigned long long n,
                                     float *__restrict__ out) {
    for (unsigned long long i = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long) gridDim.x * blockDim.x)
        out[i] = value_traits<VT>::decode(in[i]);
}
} // namespace
} // namespace csamd

// C ABI: errors, device count, convert workspace, scan.
// ---------------------------------------------------------------------------
extern "C" {

const char *csamd_last_error(void) { return csamd::g_last_error; }

int csamd_abi_version(void) { return 1; }

int csamd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void) hipGetLastError();
        return 0;
    }
    return n;
}

int csamd_set_device(int device) { return csamd::hip_ok(hipSetDevice(device), "hipSetDevice"); }

void csamd_convert_init(csamd_convert_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_convert_clear(csamd_convert_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    if (ws->owns_stream != 0 && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    if (ws->scratch != nullptr) (void) hipFree(ws->scratch);
    csamd_convert_init(ws);
}

int csamd_convert_setup(csamd_convert_workspace *ws, int device, csamd_stream_t stream) {
    if (ws == nullptr) {
        csamd::set_error("convert setup: null workspace");
        return 0;
    }
    csamd_convert_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice convert setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags convert");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    return 1;
}

int csamd_exclusive_scan_u32(const unsigned int *d_counts, unsigned int n, unsigned int *d_ptr,
                             csamd_convert_workspace *ws, unsigned long long *out_total) {
    using namespace csamd;
    CSAMD_REQUIRE(ws != nullptr && d_ptr != nullptr && (n == 0u || d_counts != nullptr), "exclusive_scan_u32: bad arguments");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t stream = (hipStream_t) ws->stream;
    const size_t scan_bytes = exclusive_scan_scratch_bytes(n);
    const size_t total_bytes = scan_bytes + 256u + (size_t) (n + 1u) * sizeof(unsigned long long);
    char *scratch = (char *) convert_scratch(ws, total_bytes);
    if (scratch == nullptr) return 0;
    unsigned long long *wide = (unsigned long long *) (scratch + scan_bytes + 256u);
    if (!exclusive_scan_u32_to_u64(d_counts, n, wide, scratch, scan_bytes, stream)) return 0;
    unsigned long long total = 0ull;
    CSAMD_HIP(hipMemcpyAsync(&total, wide + n, sizeof(total), hipMemcpyDeviceToHost, stream), "scan total D2H");
    CSAMD_HIP(hipStreamSynchronize(stream), "scan sync");
    CSAMD_REQUIRE(total <= 0xffffffffull, "exclusive_scan_u32: total %llu exceeds u32", total);
    if (!exclusive_scan_u32(d_counts, n, d_ptr, nullptr, scratch, scan_bytes, stream)) return 0;
    CSAMD_HIP(hipStreamSynchronize(stream), "scan sync 2");
    if (out_total != nullptr) *out_total = total;
    return 1;
}

int csamd_decode_values(unsigned int value_type, const void *d_values, unsigned long long n, float *d_out, csamd_stream_t stream) {
    CSAMD_REQUIRE(csamd::value_type_decodable(value_type), "decode_values: unsupported value type %u", value_type);
    if (n == 0ull) return 1;
    CSAMD_REQUIRE(d_values != nullptr && d_out != nullptr, "decode_values: null argument");
    const unsigned int grid = csamd::grid_for(n, 256u, 8192u);
    hipStream_t s = (hipStream_t) stream;
    switch (value_type) {
        case CSAMD_VALUE_F16: csamd::decode_values_kernel<CSAMD_VALUE_F16><<<grid, 256, 0, s>>>((const uint16_t *) d_values, n, d_out); break;
        case CSAMD_VALUE_BF16: csamd::decode_values_kernel<CSAMD_VALUE_BF16><<<grid, 256, 0, s>>>((const uint16_t *) d_values, n, d_out); break;
        case CSAMD_VALUE_F32: csamd::decode_values_kernel<CSAMD_VALUE_F32><<<grid, 256, 0, s>>>((const float *) d_values, n, d_out); break;
        default: csamd::decode_values_kernel<CSAMD_VALUE_FP8_E4M3><<<grid, 256, 0, s>>>((const uint8_t *) d_values, n, d_out); break;
    }
    return csamd::hip_ok(hipGetLastError(), "decode_values_kernel");
}

} // extern "C"
