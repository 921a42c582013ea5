// Unit loads shared by the row-streaming kernels (label reduce, group reduce).
//
// A unit is 4 consecutive entries of a row-major sparse layout: one 16-byte
// index load and one 4-value load (8 bytes for 2-byte values, 16 for f32, 4
// for fp8), both non-temporal.  Loaded words stay packed in registers until
// they are consumed: any instruction that reads them at load time would make
// the wave wait for the load right there and void the prefetch.
#pragma once

#include "common.hh"

namespace csamd {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template<int VT> struct unit_vals;
template<> struct unit_vals<CSAMD_VALUE_F16> {
    using raw = u32x2;
    static __device__ __forceinline__ raw vload(const uint16_t *v, unsigned int i0) {
        return __builtin_nontemporal_load(gmem((const u32x2 *) (v + i0)));
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
        return __builtin_nontemporal_load(gmem((const u32x2 *) (v + i0)));
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
        return __builtin_nontemporal_load(gmem((const u32x4 *) (v + i0)));
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
        return __builtin_nontemporal_load(gmem((const unsigned int *) (v + i0)));
    }
    static __device__ __forceinline__ raw sload(const uint8_t *v, unsigned int i0, unsigned int last) {
        return (unsigned int) v[min(i0, last)] | ((unsigned int) v[min(i0 + 1u, last)] << 8) |
               ((unsigned int) v[min(i0 + 2u, last)] << 16) | ((unsigned int) v[min(i0 + 3u, last)] << 24);
    }
    static __device__ __forceinline__ float get(const raw &r, unsigned int k) { return fp8_e4m3_decode_hw(r, (int) k); }
};


__device__ __forceinline__ unsigned int lane_of(const u32x4 &w, unsigned int k) {
    return k == 0u ? w.x : k == 1u ? w.y : k == 2u ? w.z : w.w;
}

// Value e of a 16-byte chunk.
template<typename S> __device__ __forceinline__ S chunk_val(const u32x4 &q, unsigned int e);
template<> __device__ __forceinline__ uint16_t chunk_val<uint16_t>(const u32x4 &q, unsigned int e) {
    const unsigned int w = lane_of(q, e >> 1);
    return (uint16_t) ((e & 1u) ? (w >> 16) : (w & 0xffffu));
}
template<> __device__ __forceinline__ float chunk_val<float>(const u32x4 &q, unsigned int e) { return __uint_as_float(lane_of(q, e)); }
template<> __device__ __forceinline__ uint8_t chunk_val<uint8_t>(const u32x4 &q, unsigned int e) {
    return (uint8_t) (lane_of(q, e >> 2) >> ((e & 3u) * 8u));
}

template<typename S, unsigned int U> struct bell_row_buf {
    u32x4 q[U];
    unsigned int bc[U];
};

// Walk the 16-byte chunks of one Blocked-ELL value row with the whole wave,
// U chunks per lane per step (chunk c0 + 64 u + lane), through two register
// buffers in turn: the next step's value and block-column loads are in
// flight while visit(c, q, bc) runs for the current step, u = 0, 1, ... in
// order.  Loads past the row's last chunk re-read it (visit sees c >= nc
// for them), so every load is unconditional.  Chunks per block must be a
// power of two (2^cpb_log2).
template<typename S, unsigned int U, typename Visit>
__device__ __forceinline__ void bell_stream_row(const csamd_blocked_ell_view &src, unsigned int width, unsigned int cpb_log2,
                                                unsigned int row, Visit visit) {
    constexpr unsigned int PER = 16u / sizeof(S);
    constexpr unsigned int STEP = 64u * U;
    const unsigned int nc = src.ell_cols / PER;
    if (nc == 0u) return;
    const unsigned int lane = lane_id();
    const CSAMD_GLOBAL u32x4 *vrow = (const CSAMD_GLOBAL u32x4 *) gmem((const S *) src.val + (unsigned long long) row * src.ell_cols);
    const CSAMD_GLOBAL unsigned int *bcrow = gmem(src.blockColIdx + (unsigned long long) (row / src.block_size) * width);
    auto load = [&](unsigned int c0, bell_row_buf<S, U> &B) {
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) {
            const unsigned int c = min(c0 + 64u * u + lane, nc - 1u);
            B.q[u] = vrow[c];
            B.bc[u] = bcrow[c >> cpb_log2];
        }
    };
    bell_row_buf<S, U> A, B;
    load(0u, A);
    for (unsigned int c0 = 0u; c0 < nc; c0 += 2u * STEP) {
        load(c0 + STEP, B);
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) visit(c0 + 64u * u + lane, A.q[u], A.bc[u]);
        if (c0 + STEP >= nc) break;
        load(c0 + 2u * STEP, A);
#pragma unroll
        for (unsigned int u = 0; u < U; ++u) visit(c0 + STEP + 64u * u + lane, B.q[u], B.bc[u]);
    }
}

} // namespace csamd
