// Shared device/host helpers for the cellshard_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/cellshard_amd/cellshard_amd.h"

namespace csamd {

// ---------------------------------------------------------------------------
// Error reporting: thread-local last error, 1/0 return convention.
// ---------------------------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

inline int hip_ok(hipError_t err, const char *label) {
    if (err == hipSuccess) return 1;
    set_error("HIP error at %s: %s", label, hipGetErrorString(err));
    return 0;
}

#define CSAMD_HIP(expr, label) do { if (!::csamd::hip_ok((expr), (label))) return 0; } while (0)
#define CSAMD_REQUIRE(cond, ...) do { if (!(cond)) { ::csamd::set_error(__VA_ARGS__); return 0; } } while (0)

__host__ __device__ inline size_t align_up(size_t v, size_t a) { return (v + a - 1u) & ~(a - 1u); }

// Launch-size helpers.
inline unsigned int grid_for(unsigned long long n, unsigned int per_block, unsigned int cap = 65535u * 16u) {
    unsigned long long g = (n + per_block - 1ull) / per_block;
    if (g < 1ull) g = 1ull;
    if (g > cap) g = cap;
    return (unsigned int) g;
}

// Grow-only scratch inside a convert workspace.
void *convert_scratch(csamd_convert_workspace *ws, size_t bytes);
int ensure_device(int device);

// Number of compute units on the current device (cached per device).
int device_cu_count();

// Owner of a bucketed execution partition's memory (the `storage` field of
// csamd_bucketed_blocked_ell / csamd_bucketed_sliced_ell): device and host
// blocks that the csamd_bucketed_*_clear calls free.  exec_storage_dev and
// exec_storage_host return nullptr when the allocation fails.
void *exec_storage_new();
void *exec_storage_dev(void *storage, size_t bytes);
void *exec_storage_host(void *storage, size_t bytes);
// The reference's packed-byte estimates of one Blocked-ELL / Sliced-ELL
// segment with 2-byte values (the `bytes` field of the bucketed partitions).
unsigned long long exec_bell_packed_bytes(unsigned long long rows, unsigned long long ell_cols, unsigned long long block_size);
unsigned long long exec_sell_packed_bytes(unsigned long long slices, unsigned long long slots);

// ---------------------------------------------------------------------------
// Value decoding (all to fp32).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float f16_bits_to_float(uint16_t bits) {
    __half h;
    __builtin_memcpy(&h, &bits, 2);
    return __half2float(h);
}

__device__ __forceinline__ float bf16_bits_to_float(uint16_t bits) {
    return __uint_as_float(((unsigned int) bits) << 16);
}

// OCP FP8 E4M3 (e4m3fn): bias 7, no infinities, S.1111.111 = NaN.
__host__ __device__ __forceinline__ float fp8_e4m3_to_float(uint8_t bits) {
    const unsigned int sign = (unsigned int) (bits >> 7) << 31;
    const unsigned int exp = (bits >> 3) & 0xFu;
    const unsigned int man = bits & 0x7u;
    unsigned int out;
    if (exp == 0xFu && man == 0x7u) {
        out = sign | 0x7FC00000u;
    } else if (exp == 0u) {
        // subnormal: man * 2^-9
        float f = (float) man * 0.001953125f;
        unsigned int b;
        __builtin_memcpy(&b, &f, 4);
        out = sign | b;
    } else {
        out = sign | ((exp + 120u) << 23) | (man << 20);
    }
    float r;
    __builtin_memcpy(&r, &out, 4);
    return r;
}

template<int VT> struct value_traits;
template<> struct value_traits<CSAMD_VALUE_F16> {
    using storage = uint16_t;
    static __device__ __forceinline__ float decode(storage s) { return f16_bits_to_float(s); }
};
template<> struct value_traits<CSAMD_VALUE_BF16> {
    using storage = uint16_t;
    static __device__ __forceinline__ float decode(storage s) { return bf16_bits_to_float(s); }
};
template<> struct value_traits<CSAMD_VALUE_F32> {
    using storage = float;
    static __device__ __forceinline__ float decode(storage s) { return s; }
};
// gfx950's v_cvt_f32_fp8 reads OCP e4m3fn (bias 7, 0x7f / 0xff NaN, no
// infinities): one instruction per value, checked against the closed form
// on all 256 codes by the value-decode test.
__device__ __forceinline__ float fp8_e4m3_decode_hw(unsigned int word, int byte) {
    switch (byte) {
        case 0: return __builtin_amdgcn_cvt_f32_fp8((int) word, 0);
        case 1: return __builtin_amdgcn_cvt_f32_fp8((int) word, 1);
        case 2: return __builtin_amdgcn_cvt_f32_fp8((int) word, 2);
        default: return __builtin_amdgcn_cvt_f32_fp8((int) word, 3);
    }
}

template<> struct value_traits<CSAMD_VALUE_FP8_E4M3> {
    using storage = uint8_t;
    static __device__ __forceinline__ float decode(storage s) { return fp8_e4m3_decode_hw(s, 0); }
};

inline int value_type_bytes(unsigned int vt) {
    switch (vt) {
        case CSAMD_VALUE_F16: return 2;
        case CSAMD_VALUE_BF16: return 2;
        case CSAMD_VALUE_F32: return 4;
        case CSAMD_VALUE_F64: return 8;
        case CSAMD_VALUE_FP8_E4M3: return 1;
        case CSAMD_VALUE_FP8_E5M2: return 1;
        default: return 0;
    }
}

inline int value_type_decodable(unsigned int vt) {
    return vt == CSAMD_VALUE_F16 || vt == CSAMD_VALUE_BF16 || vt == CSAMD_VALUE_F32 || vt == CSAMD_VALUE_FP8_E4M3;
}

// ---------------------------------------------------------------------------
// Global-memory views of generic pointers.  Accesses through a generic
// pointer compile to flat_* instructions; a flat access may touch LDS, so the
// compiler counts it on both the vector-memory and the LDS counters and then
// waits for every outstanding load (vmcnt(0) lgkmcnt(0)) wherever it needs
// one of them, which voids any prefetch.  Through these views the same
// accesses are global_* and the waits count only what they need.
// ---------------------------------------------------------------------------
#define CSAMD_GLOBAL __attribute__((address_space(1)))
template<typename T> __device__ __forceinline__ CSAMD_GLOBAL T *gmem(T *p) {
    return (CSAMD_GLOBAL T *) p;
}
template<typename T> __device__ __forceinline__ const CSAMD_GLOBAL T *gmem(const T *p) {
    return (const CSAMD_GLOBAL T *) p;
}

// ---------------------------------------------------------------------------
// Wave64 helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned int lane_id() { return __lane_id(); }

__device__ __forceinline__ float wave_sum_f32(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max_f32(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ unsigned int wave_sum_u32(unsigned int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ unsigned int wave_max_u32(unsigned int v) {
    for (int o = 32; o > 0; o >>= 1) { unsigned int w = __shfl_xor(v, o, 64); v = w > v ? w : v; }
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Lanes strictly below this one.
__device__ __forceinline__ unsigned long long lanemask_lt() {
    const unsigned int l = lane_id();
    return l == 0u ? 0ull : (~0ull >> (64u - l));
}

// Inclusive prefix sum over the wave with DPP moves (no LDS round trips):
// within each row of 16 lanes, shifts right by 1, 2, 4 and 8 (lanes whose
// source lies outside the row add 0); then lane 15 of rows 0 and 2 is added
// to rows 1 and 3 (row_bcast:15), and lane 31 to rows 2 and 3
// (row_bcast:31).  Disabled rows and invalid sources read the `old` operand,
// 0 here.
__device__ __forceinline__ unsigned int wave_inclusive_scan_u32(unsigned int v) {
    v += (unsigned int) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (unsigned int) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (unsigned int) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (unsigned int) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (unsigned int) __builtin_amdgcn_update_dpp(0, (int) v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1, 3
    v += (unsigned int) __builtin_amdgcn_update_dpp(0, (int) v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2, 3
    return v;
}
__device__ __forceinline__ unsigned long long wave_inclusive_scan_u64(unsigned long long v) {
    const unsigned int l = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long t = __shfl_up(v, o, 64);
        if (l >= (unsigned int) o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one u32 per thread; blockDim.x multiple of 64,
// at most 1024.  `smem` must hold 16 entries.  Returns the exclusive prefix and
// writes the block total to *total.
__device__ __forceinline__ unsigned int block_exclusive_scan_u32(unsigned int v, unsigned int *smem, unsigned int *total) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int nwaves = blockDim.x >> 6;
    const unsigned int incl = wave_inclusive_scan_u32(v);
    if (lane_id() == 63u) smem[wave] = incl;
    __syncthreads();
    unsigned int wave_base = 0u, sum = 0u;
    for (unsigned int w = 0; w < nwaves; ++w) {
        const unsigned int s = smem[w];
        if (w < wave) wave_base += s;
        sum += s;
    }
    __syncthreads();
    *total = sum;
    return wave_base + incl - v;
}

__device__ __forceinline__ unsigned long long block_exclusive_scan_u64(unsigned long long v, unsigned long long *smem,
                                                                        unsigned long long *total) {
    const unsigned int wave = threadIdx.x >> 6;
    const unsigned int nwaves = blockDim.x >> 6;
    const unsigned long long incl = wave_inclusive_scan_u64(v);
    if (lane_id() == 63u) smem[wave] = incl;
    __syncthreads();
    unsigned long long wave_base = 0ull, sum = 0ull;
    for (unsigned int w = 0; w < nwaves; ++w) {
        const unsigned long long s = smem[w];
        if (w < wave) wave_base += s;
        sum += s;
    }
    __syncthreads();
    *total = sum;
    return wave_base + incl - v;
}

// ---------------------------------------------------------------------------
// Device-wide scans (reduce-then-scan; inputs are metadata-sized).
// ---------------------------------------------------------------------------
// Tiles of CSAMD_SCAN_TILE consecutive items: per-tile exclusive offsets
// into d_tile_offsets[0..tiles), total into d_tile_offsets[tiles]
// (tiles = max(1, ceil(n / CSAMD_SCAN_TILE))).
#define CSAMD_SCAN_TILE 4096u
int scan_tile_offsets_u32(const unsigned int *d_in, unsigned int n, unsigned long long *d_tile_offsets, hipStream_t stream);
// Exclusive scan of u32 counts into u64 offsets out[0..n] (out[n] = total).
int exclusive_scan_u32_to_u64(const unsigned int *d_in, unsigned int n, unsigned long long *d_out,
                              void *scratch, size_t scratch_bytes, hipStream_t stream);
size_t exclusive_scan_scratch_bytes(unsigned int n);
// Exclusive scan of u32 counts into u32 offsets out[0..n] computed through
// u64; *d_overflow (device u32) set to 1 if the total exceeds u32.
// With write_total false only out[0..n-1] is written.
int exclusive_scan_u32(const unsigned int *d_in, unsigned int n, unsigned int *d_out, unsigned int *d_overflow,
                       void *scratch, size_t scratch_bytes, hipStream_t stream, bool write_total = true);

// Zero-fill n bytes (device fill kernel; usable inside captured graphs).
int zero_bytes(void *p, size_t n, hipStream_t stream);

// Stable ascending radix sort of (u32 key, u32 value) pairs, keys below
// 2^key_bits.  scratch from radix_sort_scratch_bytes.
size_t radix_sort_pairs_u32_scratch_bytes(unsigned int n);
int radix_sort_pairs_u32(const unsigned int *d_keys_in, unsigned int *d_keys_out, const unsigned int *d_vals_in,
                         unsigned int *d_vals_out, unsigned int n, unsigned int key_bits, void *scratch,
                         size_t scratch_bytes, hipStream_t stream);
size_t radix_sort_pairs_u64_scratch_bytes(unsigned int n);
int radix_sort_pairs_u64(const unsigned long long *d_keys_in, unsigned long long *d_keys_out,
                         const unsigned int *d_vals_in, unsigned int *d_vals_out, unsigned int n,
                         unsigned int key_bits, void *scratch, size_t scratch_bytes, hipStream_t stream);

} // namespace csamd
