// Synthetic count matrices for measurement and tests (not a reference path).
//
// Row r: k_r = clamp(round(exp(ln(median) + sigma * z_r)), 0, cols), z_r ~ N(0,1).
// Columns: weighted sampling of k_r distinct columns without replacement with
// weights w_c (Efraimidis-Spirakis: key_c = log(u_rc) / w_c, take the k_r
// largest keys; equivalent to successive weighted draws), emitted ascending.
// Values: zero-truncated NB2 counts, mean mu_c, dispersion alpha (r = 1/alpha),
// drawn by inverting the CDF above P(0).
// All randomness is a counter-based hash of (seed, row, col, stream), so the
// output depends only on the parameters.
//
// gfx950 mapping: one 256-thread workgroup per row; the k-th largest key is
// found by a 4-pass 8-bit radix select (LDS histograms), keys are recomputed
// instead of stored, and the ascending emit uses wave ballots.
#include "common.hh"

namespace csamd {
namespace {

constexpr unsigned int SY_THREADS = 256u;
constexpr unsigned int SY_WAVES = SY_THREADS / 64u;

__host__ __device__ __forceinline__ unsigned int lowbias32(unsigned int x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ unsigned int hash4(unsigned int seed, unsigned int a, unsigned int b, unsigned int stream) {
    return lowbias32(lowbias32(lowbias32(seed ^ (stream * 0x9e3779b9u)) ^ a) + b * 0x85ebca6bu);
}

// uniform in (0, 1]: in float, (2^24 - 1) + 0.5 rounds to 2^24, so the
// largest draw is 1.0 (used where that is harmless: log keys, Box-Muller)
__device__ __forceinline__ float unit_open(unsigned int h) { return ((float) (h >> 8) + 0.5f) * (1.0f / 16777216.0f); }
// uniform in (0, 1), exactly: the count sampler inverts a CDF that never
// reaches 1 in floating point, so a draw of 1.0 would run its search to the
// underflow of the probability terms (counts of ~1e5, +inf in fp16)
__device__ __forceinline__ double unit_open_d(unsigned int h) { return ((double) (h >> 8) + 0.5) * (1.0 / 16777216.0); }

__device__ __forceinline__ unsigned int ordered_bits(float f) {
    const unsigned int b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ unsigned int sample_key(unsigned int seed, unsigned int row, unsigned int col, const float *__restrict__ inv_w) {
    const float u = unit_open(hash4(seed, row, col, 1u));
    return ordered_bits(__logf(u) * inv_w[col]);
}

__global__ void synth_row_nnz_kernel(unsigned int rows, unsigned int cols, unsigned int seed, float log_median, float sigma,
                                     unsigned int *__restrict__ row_nnz) {
    for (unsigned int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        const float u1 = unit_open(hash4(seed, r, 0u, 7u));
        const float u2 = unit_open(hash4(seed, r, 1u, 7u));
        const float z = sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
        const float x = expf(log_median + sigma * z);
        float k = rintf(x);
        if (!(k > 0.0f)) k = 0.0f;
        row_nnz[r] = k >= (float) cols ? cols : (unsigned int) k;
    }
}

__device__ __forceinline__ unsigned int nb_zero_truncated(unsigned int seed, unsigned int row, unsigned int col, float mu, float r_disp) {
    const double m = mu > 1e-12f ? (double) mu : 1e-12;
    const double rr = (double) r_disp;
    const double p = rr / (rr + m);
    const double q = 1.0 - p;
    const double p0 = pow(p, rr);
    const double u = unit_open_d(hash4(seed, row, col, 2u));
    const double t = p0 + u * (1.0 - p0);
    double pk = p0, cdf = p0;
    unsigned int k = 0u;
    const unsigned int cap = 1u << 20;
    do {
        ++k;
        pk *= q * ((double) k - 1.0 + rr) / (double) k;
        cdf += pk;
    } while (cdf < t && k < cap && pk > 0.0);
    return k;
}

template<typename S> __device__ __forceinline__ S encode_count(unsigned int k);
template<> __device__ __forceinline__ float encode_count<float>(unsigned int k) { return (float) k; }
template<> __device__ __forceinline__ uint16_t encode_count<uint16_t>(unsigned int k) {
    const __half h = __float2half_rn((float) k);
    uint16_t b;
    __builtin_memcpy(&b, &h, 2);
    return b;
}

template<typename S>
__global__ void __launch_bounds__(SY_THREADS) synth_fill_kernel(unsigned int rows, unsigned int cols, unsigned int seed, float r_disp,
                                                                const float *__restrict__ inv_w, const float *__restrict__ col_mean,
                                                                const unsigned int *__restrict__ ptr, unsigned int *__restrict__ idx,
                                                                S *__restrict__ val) {
    __shared__ unsigned int hist[256];
    __shared__ unsigned int sel_digit, sel_rem;
    __shared__ unsigned int wave_gt[SY_WAVES], wave_eq[SY_WAVES];
    const unsigned int wave = threadIdx.x >> 6, lane = lane_id();
    // columns of wave w: [w*span, min(cols, (w+1)*span))
    const unsigned int span = (((cols + SY_WAVES - 1u) / SY_WAVES) + 63u) & ~63u;
    const unsigned int cw0 = wave * span;
    const unsigned int cw1 = cw0 + span < cols ? cw0 + span : cols;
    for (unsigned int row = blockIdx.x; row < rows; row += gridDim.x) {
        const unsigned int base = ptr[row];
        const unsigned int k = ptr[row + 1u] - base;
        if (k == 0u) continue;
        unsigned int thr = 0u, take_eq = 0u;
        const bool all = k >= cols;
        if (!all) {
            unsigned int prefix = 0u, mask = 0u, rem = k;
            for (int pass = 3; pass >= 0; --pass) {
                const unsigned int shift = (unsigned int) pass * 8u;
                hist[threadIdx.x] = 0u;
                __syncthreads();
                for (unsigned int c = threadIdx.x; c < cols; c += SY_THREADS) {
                    const unsigned int key = sample_key(seed, row, c, inv_w);
                    if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 0xffu], 1u);
                }
                __syncthreads();
                // suffix counts from the top digit down
                const unsigned int d = 255u - threadIdx.x;
                const unsigned int h = hist[d];
                unsigned int tot = 0u;
                __shared__ unsigned int scan_tmp[16];
                const unsigned int above = block_exclusive_scan_u32(h, scan_tmp, &tot);
                if (above < rem && rem <= above + h) {
                    sel_digit = d;
                    sel_rem = rem - above;
                }
                __syncthreads();
                prefix |= sel_digit << shift;
                mask |= 0xffu << shift;
                rem = sel_rem;
                __syncthreads();
            }
            thr = prefix;
            take_eq = rem;
        }
        // per-wave totals
        unsigned int ngt = 0u, neq = 0u;
        for (unsigned int c0 = cw0; c0 < cw1; c0 += 64u) {
            const unsigned int c = c0 + lane;
            bool gt = false, eq = false;
            if (c < cw1) {
                if (all) gt = true;
                else {
                    const unsigned int key = sample_key(seed, row, c, inv_w);
                    gt = key > thr;
                    eq = key == thr;
                }
            }
            ngt += (unsigned int) __popcll(__ballot(gt));
            neq += (unsigned int) __popcll(__ballot(eq));
        }
        if (lane == 0u) {
            wave_gt[wave] = ngt;
            wave_eq[wave] = neq;
        }
        __syncthreads();
        unsigned int eq_before = 0u, eq_in_prior = 0u, out_before = 0u;
        for (unsigned int w = 0; w < wave; ++w) eq_in_prior += wave_eq[w];
        for (unsigned int w = 0; w < wave; ++w) {
            unsigned int eq_sel_w = 0u;
            // equal keys taken by wave w: clamp(take_eq - eq_before_w, 0, eq_w)
            const unsigned int avail = take_eq > eq_before ? take_eq - eq_before : 0u;
            eq_sel_w = wave_eq[w] < avail ? wave_eq[w] : avail;
            out_before += wave_gt[w] + eq_sel_w;
            eq_before += wave_eq[w];
        }
        (void) eq_in_prior;
        unsigned int eq_run = eq_before, out_run = out_before;
        for (unsigned int c0 = cw0; c0 < cw1; c0 += 64u) {
            const unsigned int c = c0 + lane;
            bool gt = false, eq = false;
            if (c < cw1) {
                if (all) gt = true;
                else {
                    const unsigned int key = sample_key(seed, row, c, inv_w);
                    gt = key > thr;
                    eq = key == thr;
                }
            }
            const unsigned long long beq = __ballot(eq);
            const unsigned int my_eq_rank = eq_run + (unsigned int) __popcll(beq & lanemask_lt());
            const bool sel = gt || (eq && my_eq_rank < take_eq);
            const unsigned long long bsel = __ballot(sel);
            if (sel) {
                const unsigned int pos = out_run + (unsigned int) __popcll(bsel & lanemask_lt());
                idx[base + pos] = c;
                val[base + pos] = encode_count<S>(nb_zero_truncated(seed, row, c, col_mean[c], r_disp));
            }
            eq_run += (unsigned int) __popcll(beq);
            out_run += (unsigned int) __popcll(bsel);
        }
        __syncthreads();
    }
}

// log1p(count / row_total * scale): per-row totals, then transform.
__global__ void row_totals_kernel(unsigned int rows, const unsigned int *__restrict__ ptr, const float *__restrict__ counts,
                                  float *__restrict__ totals) {
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (unsigned int r = blockIdx.x * nw + wave; r < rows; r += gridDim.x * nw) {
        double s = 0.0;
        for (unsigned int i = ptr[r] + lane_id(); i < ptr[r + 1u]; i += 64u) s += (double) counts[i];
        s = wave_sum_f64(s);
        if (lane_id() == 0u) totals[r] = (float) s;
    }
}

__global__ void log1p_kernel(unsigned int rows, const unsigned int *__restrict__ ptr, const unsigned int *__restrict__ idx,
                             const float *__restrict__ counts, const float *__restrict__ totals, float scale, float *__restrict__ out,
                             unsigned int *__restrict__ col_absmax_bits, unsigned int cols) {
    const unsigned int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (unsigned int r = blockIdx.x * nw + wave; r < rows; r += gridDim.x * nw) {
        const float t = totals[r];
        for (unsigned int i = ptr[r] + lane_id(); i < ptr[r + 1u]; i += 64u) {
            const float v = t > 0.0f ? log1pf(counts[i] / t * scale) : 0.0f;
            out[i] = v;
            const unsigned int c = idx[i];
            if (col_absmax_bits != nullptr && c < cols) atomicMax(col_absmax_bits + c, __float_as_uint(fabsf(v)));
        }
    }
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    unsigned int b = __float_as_uint(f);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((b >> 16) | 0x40u);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t) (b >> 16);
}

// OCP e4m3fn, round to nearest even, saturating to +-448.
__device__ __forceinline__ uint8_t f32_to_e4m3_rne(float f) {
    const unsigned int sign = (__float_as_uint(f) >> 24) & 0x80u;
    float a = fabsf(f);
    if (!(a == a)) return (uint8_t) (sign | 0x7fu);
    if (a >= 448.0f) return (uint8_t) (sign | 0x7eu);
    if (a < 0.0009765625f) {  // below half the smallest subnormal (2^-9 / 2) -> 0 or min subnormal
        // subnormal grid step 2^-9
        const float q = rintf(a * 512.0f);
        return (uint8_t) (sign | (unsigned int) q);
    }
    int e;
    frexpf(a, &e);             // a = m * 2^e, m in [0.5, 1)
    int exp_unb = e - 1;       // a = 1.x * 2^exp_unb
    if (exp_unb < -6) {        // subnormal range: step 2^-9
        const unsigned int q = (unsigned int) rintf(a * 512.0f);
        if (q >= 8u) return (uint8_t) (sign | (1u << 3));
        return (uint8_t) (sign | q);
    }
    const float scale = ldexpf(1.0f, 3 - exp_unb);  // mantissa with 3 fraction bits
    unsigned int q = (unsigned int) rintf(a * scale);  // in [8, 16]
    if (q == 16u) {
        q = 8u;
        exp_unb += 1;
    }
    const unsigned int ebits = (unsigned int) (exp_unb + 7);
    if (ebits > 15u || (ebits == 15u && (q & 7u) == 7u)) return (uint8_t) (sign | 0x7eu);
    return (uint8_t) (sign | (ebits << 3) | (q & 7u));
}

__global__ void encode_kernel(unsigned long long n, const float *__restrict__ src, const unsigned int *__restrict__ idx,
                              unsigned int value_type, const float *__restrict__ col_scale, void *__restrict__ out) {
    for (unsigned long long i = blockIdx.x * (unsigned long long) blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long) gridDim.x * blockDim.x) {
        const float v = src[i];
        if (value_type == CSAMD_VALUE_BF16) ((uint16_t *) out)[i] = f32_to_bf16_rne(v);
        else if (value_type == CSAMD_VALUE_F16) {
            const __half h = __float2half_rn(v);
            uint16_t b;
            __builtin_memcpy(&b, &h, 2);
            ((uint16_t *) out)[i] = b;
        } else if (value_type == CSAMD_VALUE_FP8_E4M3) {
            const float s = col_scale[idx[i]];
            ((uint8_t *) out)[i] = f32_to_e4m3_rne(s > 0.0f ? v / s : 0.0f);
        } else {
            ((float *) out)[i] = v;
        }
    }
}

__global__ void col_scale_kernel(unsigned int cols, const unsigned int *__restrict__ absmax_bits, float *__restrict__ scale) {
    for (unsigned int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
        const float m = __uint_as_float(absmax_bits[c]);
        scale[c] = m > 0.0f ? m / 448.0f : 1.0f;
    }
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

int csamd_synth_row_nnz(const csamd_synth_params *p, unsigned int *d_row_nnz, csamd_stream_t stream) {
    CSAMD_REQUIRE(p != nullptr && (p->rows == 0u || d_row_nnz != nullptr), "synth_row_nnz: null argument");
    CSAMD_REQUIRE(p->row_nnz_median > 0.0f, "synth_row_nnz: median must be positive");
    if (p->rows == 0u) return 1;
    synth_row_nnz_kernel<<<grid_for(p->rows, 256u, 8192u), 256, 0, (hipStream_t) stream>>>(p->rows, p->cols, p->seed,
                                                                                           logf(p->row_nnz_median), p->row_nnz_sigma, d_row_nnz);
    return hip_ok(hipGetLastError(), "synth_row_nnz_kernel");
}

int csamd_synth_fill(const csamd_synth_params *p, const unsigned int *d_ptr, unsigned int *d_idx, void *d_val, csamd_stream_t stream) {
    CSAMD_REQUIRE(p != nullptr && d_ptr != nullptr, "synth_fill: null argument");
    CSAMD_REQUIRE(p->d_col_inv_weight != nullptr && p->d_col_mean != nullptr, "synth_fill: column tables missing");
    CSAMD_REQUIRE(p->nb_alpha > 0.0f, "synth_fill: nb_alpha must be positive");
    CSAMD_REQUIRE(p->value_type == CSAMD_VALUE_F16 || p->value_type == CSAMD_VALUE_F32, "synth_fill: value type must be f16 or f32");
    if (p->rows == 0u || p->cols == 0u) return 1;
    const unsigned int grid = grid_for(p->rows, 1u, (unsigned int) device_cu_count() * 8u);
    // d_col_inv_weight carries 1 / w_c (inverse weights) for the key computation.
    if (p->value_type == CSAMD_VALUE_F16)
        synth_fill_kernel<uint16_t><<<grid, SY_THREADS, 0, (hipStream_t) stream>>>(p->rows, p->cols, p->seed, 1.0f / p->nb_alpha,
                                                                                 p->d_col_inv_weight, p->d_col_mean, d_ptr, d_idx,
                                                                                 (uint16_t *) d_val);
    else
        synth_fill_kernel<float><<<grid, SY_THREADS, 0, (hipStream_t) stream>>>(p->rows, p->cols, p->seed, 1.0f / p->nb_alpha,
                                                                              p->d_col_inv_weight, p->d_col_mean, d_ptr, d_idx,
                                                                              (float *) d_val);
    return hip_ok(hipGetLastError(), "synth_fill_kernel");
}

int csamd_synth_log1p_normalize(unsigned int rows, unsigned int cols, const unsigned int *d_ptr, const unsigned int *d_idx,
                                const float *d_counts, float scale, unsigned int value_type, void *d_val_out, float *d_feature_scale,
                                csamd_convert_workspace *ws) {
    CSAMD_REQUIRE(ws != nullptr && d_ptr != nullptr, "synth_log1p_normalize: null argument");
    CSAMD_REQUIRE(value_type == CSAMD_VALUE_F32 || value_type == CSAMD_VALUE_F16 || value_type == CSAMD_VALUE_BF16 ||
                      value_type == CSAMD_VALUE_FP8_E4M3,
                  "synth_log1p_normalize: unsupported value type %u", value_type);
    CSAMD_REQUIRE(value_type != CSAMD_VALUE_FP8_E4M3 || d_feature_scale != nullptr, "synth_log1p_normalize: fp8 needs feature scales");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    unsigned int nnz = 0u;
    CSAMD_HIP(hipMemcpyAsync(&nnz, d_ptr + rows, 4u, hipMemcpyDeviceToHost, s), "D2H nnz");
    CSAMD_HIP(hipStreamSynchronize(s), "sync");
    const size_t tot_bytes = align_up((size_t) rows * 4u, 256u), amax_bytes = align_up((size_t) cols * 4u, 256u);
    char *scr = (char *) convert_scratch(ws, tot_bytes + amax_bytes + align_up((size_t) nnz * 4u, 256u));
    if (scr == nullptr) return 0;
    float *totals = (float *) scr;
    unsigned int *amax = (unsigned int *) (scr + tot_bytes);
    float *tmp = (float *) (scr + tot_bytes + amax_bytes);
    if (!zero_bytes(amax, (size_t) cols * 4u, s)) return 0;
    const unsigned int g = grid_for((rows + 3u) / 4u, 1u, (unsigned int) device_cu_count() * 16u);
    row_totals_kernel<<<g, 256, 0, s>>>(rows, d_ptr, d_counts, totals);
    CSAMD_HIP(hipGetLastError(), "row_totals_kernel");
    log1p_kernel<<<g, 256, 0, s>>>(rows, d_ptr, d_idx, d_counts, totals, scale, tmp, amax, cols);
    CSAMD_HIP(hipGetLastError(), "log1p_kernel");
    if (value_type == CSAMD_VALUE_FP8_E4M3) {
        col_scale_kernel<<<grid_for(cols, 256u, 1024u), 256, 0, s>>>(cols, amax, d_feature_scale);
        CSAMD_HIP(hipGetLastError(), "col_scale_kernel");
    }
    encode_kernel<<<grid_for(nnz, 256u, 16384u), 256, 0, s>>>(nnz, tmp, d_idx, value_type, d_feature_scale, d_val_out);
    CSAMD_HIP(hipGetLastError(), "encode_kernel");
    CSAMD_HIP(hipStreamSynchronize(s), "log1p sync");
    return 1;
}

} // extern "C"
