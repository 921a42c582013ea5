// One wave writes one row of a Sliced-ELL partition: the row's slots
// [base, base + width) take its first n entries from (idx, val) + b, then
// padding (CSAMD_INVALID_COL / 0).  Used by the single-partition and the
// batched CSR -> Sliced-ELL emit kernels.
//
// The slots from the row's first 4-aligned one go in quads, one per lane: a
// 16-byte column store and one store of the four values (8 bytes for 2-byte
// values), so a wave stores 1 KB of columns per instruction instead of 256
// bytes; the up to 3 slots before and after the quads are written one per
// lane.  Quads need column and value arrays aligned to 16 and 4 x
// sizeof(S) bytes (`quads`); otherwise every slot is written on its own.
#pragma once

#include "stream.hh"

namespace csamd {

template<typename S> struct sell_quad;
template<> struct sell_quad<uint16_t> {
    static __device__ __forceinline__ void store(CSAMD_GLOBAL uint16_t *p, const uint16_t (&v)[4]) {
        u32x2 w;
        w.x = (unsigned int) v[0] | ((unsigned int) v[1] << 16);
        w.y = (unsigned int) v[2] | ((unsigned int) v[3] << 16);
        *(CSAMD_GLOBAL u32x2 *) p = w;
    }
};
template<> struct sell_quad<float> {
    static __device__ __forceinline__ void store(CSAMD_GLOBAL float *p, const float (&v)[4]) {
        u32x4 w;
        w.x = __float_as_uint(v[0]);
        w.y = __float_as_uint(v[1]);
        w.z = __float_as_uint(v[2]);
        w.w = __float_as_uint(v[3]);
        *(CSAMD_GLOBAL u32x4 *) p = w;
    }
};
template<> struct sell_quad<uint8_t> {
    static __device__ __forceinline__ void store(CSAMD_GLOBAL uint8_t *p, const uint8_t (&v)[4]) {
        *(CSAMD_GLOBAL unsigned int *) p =
            (unsigned int) v[0] | ((unsigned int) v[1] << 8) | ((unsigned int) v[2] << 16) | ((unsigned int) v[3] << 24);
    }
};

template<typename S>
__device__ __forceinline__ bool sell_quads_ok(const void *col_idx, const void *val) {
    return ((uintptr_t) col_idx & 15u) == 0u && ((uintptr_t) val & (4u * sizeof(S) - 1u)) == 0u;
}

// Slots at or past `limit` (the arrays' capacity) are not written.
template<typename S>
__device__ __forceinline__ void sell_write_row(const CSAMD_GLOBAL unsigned int *idx, const CSAMD_GLOBAL S *val, unsigned int b,
                                               unsigned int n, unsigned int width, unsigned long long base,
                                               CSAMD_GLOBAL unsigned int *oc, CSAMD_GLOBAL S *ov, bool quads,
                                               unsigned long long limit) {
    const unsigned int lane = lane_id();
    auto one = [&](unsigned int j) {
        oc[base + j] = j < n ? idx[b + j] : CSAMD_INVALID_COL;
        ov[base + j] = j < n ? val[b + j] : S(0);
    };
    if (!quads || base + width > limit) {
        for (unsigned int j = lane; j < width && base + j < limit; j += 64u) one(j);
        return;
    }
    const unsigned int head = (unsigned int) ((4ull - (base & 3ull)) & 3ull);
    const unsigned int h = head < width ? head : width;
    if (lane < h) one(lane);
    const unsigned int nq = (width - h) >> 2;
    // QU quads per lane per step, every load of a step issued before its
    // stores (entry indices clamped to the row's last entry; entries past
    // n are replaced by padding when stored)
    constexpr unsigned int QU = 4u;
    const unsigned int last = n != 0u ? n - 1u : 0u;
    // quads [0, qe) hold at least one entry; steps wholly past them store
    // padding without loading (the second loop)
    const unsigned int qe = n > h ? min(nq, (n - h + 3u) >> 2) : 0u;
    unsigned int q0 = 0u;
    for (; q0 < qe; q0 += 64u * QU) {
        unsigned int ci[QU][4];
        S vi[QU][4];
        if (n != 0u) {
#pragma unroll
            for (unsigned int u = 0; u < QU; ++u)
#pragma unroll
                for (unsigned int t = 0; t < 4u; ++t) {
                    const unsigned int jj = min(h + 4u * (q0 + 64u * u + lane) + t, last);
                    ci[u][t] = idx[b + jj];
                    vi[u][t] = val[b + jj];
                }
        }
#pragma unroll
        for (unsigned int u = 0; u < QU; ++u) {
            const unsigned int q = q0 + 64u * u + lane;
            if (q >= nq) continue;
            const unsigned int j = h + 4u * q;  // base + j is a multiple of 4
            u32x4 c;
            S v[4];
#pragma unroll
            for (unsigned int t = 0; t < 4u; ++t) {
                const bool in = j + t < n;
                c[t] = in ? ci[u][t] : CSAMD_INVALID_COL;
                v[t] = in ? vi[u][t] : S(0);
            }
            *(CSAMD_GLOBAL u32x4 *) (oc + base + j) = c;
            sell_quad<S>::store(ov + base + j, v);
        }
    }
    for (; q0 < nq; q0 += 64u * QU) {
        const u32x4 c = {CSAMD_INVALID_COL, CSAMD_INVALID_COL, CSAMD_INVALID_COL, CSAMD_INVALID_COL};
        const S v[4] = {S(0), S(0), S(0), S(0)};
#pragma unroll
        for (unsigned int u = 0; u < QU; ++u) {
            const unsigned int q = q0 + 64u * u + lane;
            if (q >= nq) continue;
            const unsigned int j = h + 4u * q;
            *(CSAMD_GLOBAL u32x4 *) (oc + base + j) = c;
            sell_quad<S>::store(ov + base + j, v);
        }
    }
    const unsigned int t0 = h + 4u * nq;
    if (lane < width - t0) one(t0 + lane);
}

// One wave pads slots [base + from, base + width) of a Sliced-ELL row
// (CSAMD_INVALID_COL / 0): quads from the first 4-aligned slot when the
// arrays allow them (`quads`, as sell_quads_ok), single slots otherwise
// and at both ends.
template<typename S>
__device__ __forceinline__ void sell_pad_row(CSAMD_GLOBAL unsigned int *oc, CSAMD_GLOBAL S *ov, unsigned long long base,
                                             unsigned int from, unsigned int width, bool quads) {
    const unsigned int lane = lane_id();
    if (from >= width) return;
    auto one = [&](unsigned int j) {
        oc[base + j] = CSAMD_INVALID_COL;
        ov[base + j] = S(0);
    };
    if (!quads) {
        for (unsigned int j = from + lane; j < width; j += 64u) one(j);
        return;
    }
    const unsigned int head = (unsigned int) ((4ull - ((base + from) & 3ull)) & 3ull);
    const unsigned int h = head < width - from ? head : width - from;
    if (lane < h) one(from + lane);
    const unsigned int j0 = from + h;  // base + j0 is a multiple of 4
    const unsigned int nq = (width - j0) >> 2;
    const u32x4 c = {CSAMD_INVALID_COL, CSAMD_INVALID_COL, CSAMD_INVALID_COL, CSAMD_INVALID_COL};
    const S v[4] = {S(0), S(0), S(0), S(0)};
    for (unsigned int q = lane; q < nq; q += 64u) {
        *(CSAMD_GLOBAL u32x4 *) (oc + base + j0 + 4u * q) = c;
        sell_quad<S>::store(ov + base + j0 + 4u * q, v);
    }
    const unsigned int t0 = j0 + 4u * nq;
    if (lane < width - t0) one(t0 + lane);
}

} // namespace csamd
