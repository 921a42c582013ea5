/*
 * CPU oracle for the cellshard_amd hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never calls it.
 *
 * Each function restates the CellShard algorithm it names (paths relative to
 * the CellShard source tree).  Pinning (tests/test_oracle_*.py):
 *   group_reduce_*       mask_groups_runtime_test.cu fixtures (F3-corrected) +
 *                        dense numpy recomputation on random inputs
 *   label_reduce_*       dense numpy recomputation (np.add.at) on random inputs
 *   blocked_ell_*        the reference's own CPU converter compiled from
 *                        src/convert/blocked_ell_from_compressed.cuh (oracle/_ref)
 *                        + dense round trip
 *   sliced_ell_*         closed-form layout rules + dense round trip + the
 *                        mask_groups test's Sliced-ELL fixture
 *   transpose            scipy.sparse tocsc() (canonical ascending order)
 *   bucket / rebuild     numpy stable argsort, closed-form offsets, invariants
 *   bell_bucket_keys     closed-form hash recomputed in Python
 *   reoptimize           mask_groups_runtime_test.cu:179-198 fixture
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define INVALID_COL 0xffffffffu

enum { VT_F16 = 3, VT_F32 = 4, VT_BF16 = 6, VT_FP8_E4M3 = 7 };

/* ------------------------------------------------------------------------ */
/* value decoding                                                            */
/* ------------------------------------------------------------------------ */
static float bits_to_float(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

float oracle_half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t) (h >> 15) << 31;
    const uint32_t exp = (h >> 10) & 0x1fu;
    const uint32_t man = h & 0x3ffu;
    if (exp == 0u) {
        if (man == 0u) return bits_to_float(sign);
        float f = (float) man * (1.0f / 16777216.0f); /* man * 2^-24 */
        return sign ? -f : f;
    }
    if (exp == 31u) return bits_to_float(sign | 0x7f800000u | (man << 13));
    return bits_to_float(sign | ((exp + 112u) << 23) | (man << 13));
}

float oracle_fp8_e4m3_to_float(uint8_t b) {
    const int neg = (b >> 7) & 1;
    const uint32_t exp = (b >> 3) & 0xfu;
    const uint32_t man = b & 0x7u;
    float f;
    if (exp == 0xfu && man == 0x7u) return NAN;
    if (exp == 0u) f = ldexpf((float) man, -9);
    else f = ldexpf(1.0f + (float) man / 8.0f, (int) exp - 7);
    return neg ? -f : f;
}

static int vt_bytes(uint32_t vt) {
    switch (vt) {
        case VT_F16: case VT_BF16: return 2;
        case VT_F32: return 4;
        case VT_FP8_E4M3: return 1;
        default: return 0;
    }
}

static float decode(uint32_t vt, const void *val, size_t i) {
    switch (vt) {
        case VT_F16: return oracle_half_to_float(((const uint16_t *) val)[i]);
        case VT_BF16: return bits_to_float(((uint32_t) ((const uint16_t *) val)[i]) << 16);
        case VT_F32: return ((const float *) val)[i];
        case VT_FP8_E4M3: return oracle_fp8_e4m3_to_float(((const uint8_t *) val)[i]);
        default: return 0.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* Reference-ABI masked grouped row reduction                                 */
/* src/runtime/mask_groups.cu:65-105 (participation, empty row),             */
/* :145-194 (per-entry rule, keep predicate, percentages).                   */
/* Sums are exact FP64; keep/percentages use fp32 like the kernel.           */
/* ------------------------------------------------------------------------ */
typedef struct {
    double sum;
    float vmax;
    uint32_t detected;
    double groups[32];
} row_state;

typedef struct {
    uint32_t cols;
    const uint8_t *feature_keep;
    const uint32_t *fgm;
    uint32_t G;
} entry_ctx;

static void entry(row_state *s, const entry_ctx *x, uint32_t col, float v) {
    if (col < x->cols && v != 0.0f && (x->feature_keep == NULL || x->feature_keep[col] != 0u)) {
        s->sum += (double) v;
        if (x->fgm != NULL && x->G != 0u) {
            uint32_t m = x->fgm[col];
            while (m != 0u) {
                const uint32_t g = (uint32_t) __builtin_ctz(m);
                if (g < x->G) s->groups[g] += (double) v;
                m &= m - 1u;
            }
        }
        if (v > s->vmax) s->vmax = v;
        s->detected += 1u;
    }
}

typedef struct {
    double *row_totals;
    uint32_t *detected;
    float *max_values;
    uint8_t *keep;
    double *group_counts;
    float *group_pct;
} reduce_out;

static void finish(uint32_t row, const row_state *s, uint32_t G, float min_total, uint32_t min_det, float max_frac,
                   uint32_t frac_idx, const reduce_out *o) {
    const float rs = (float) s->sum;
    o->row_totals[row] = s->sum;
    o->detected[row] = s->detected;
    o->max_values[row] = s->vmax;
    {
        const float gsel = frac_idx < G ? (float) s->groups[frac_idx] : 0.0f;
        const float frac = rs > 0.0f ? gsel / rs : 0.0f;
        o->keep[row] = (uint8_t) (rs >= min_total && s->detected >= min_det && frac <= max_frac);
    }
    for (uint32_t g = 0; g < G; ++g) {
        const float gs = (float) s->groups[g];
        o->group_counts[(size_t) row * G + g] = s->groups[g];
        o->group_pct[(size_t) row * G + g] = rs > 0.0f ? (100.0f * gs) / rs : 0.0f;
    }
}

static void empty_row(uint32_t row, uint32_t G, const reduce_out *o) {
    o->row_totals[row] = 0.0;
    o->detected[row] = 0u;
    o->max_values[row] = 0.0f;
    o->keep[row] = 0u;
    for (uint32_t g = 0; g < G; ++g) {
        o->group_counts[(size_t) row * G + g] = 0.0;
        o->group_pct[(size_t) row * G + g] = 0.0f;
    }
}

int oracle_group_reduce_csr(uint32_t rows, uint32_t cols, const uint32_t *ptr, const uint32_t *idx, const void *val, uint32_t vt,
                            const uint8_t *row_keep, const uint8_t *feature_keep, const uint32_t *fgm, uint32_t G,
                            float min_total, uint32_t min_det, float max_frac, uint32_t frac_idx, double *row_totals,
                            uint32_t *detected, float *max_values, uint8_t *keep, double *group_counts, float *group_pct) {
    const entry_ctx x = {cols, feature_keep, fgm, G};
    const reduce_out o = {row_totals, detected, max_values, keep, group_counts, group_pct};
    if (G > 32u || vt_bytes(vt) == 0) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        row_state s;
        memset(&s, 0, sizeof(s));
        if (row_keep != NULL && row_keep[r] == 0u) {
            empty_row(r, G, &o);
            continue;
        }
        for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) entry(&s, &x, idx[i], decode(vt, val, i));
        finish(r, &s, G, min_total, min_det, max_frac, frac_idx, &o);
    }
    return 1;
}

/* slice lookup: mask_groups.cu:235-249 */
static uint32_t sell_slice(uint32_t row, uint32_t slice_count, uint32_t slice_rows, const uint32_t *offsets) {
    uint32_t slice = 0u;
    if (slice_rows == 32u) {
        slice = row >> 5;
        if (slice >= slice_count) slice = slice_count - 1u;
    } else if (slice_rows != 0u) {
        slice = row / slice_rows;
        if (slice >= slice_count) slice = slice_count - 1u;
    } else {
        while (slice + 1u < slice_count && row >= offsets[slice + 1u]) ++slice;
    }
    return slice;
}

int oracle_group_reduce_sell(uint32_t rows, uint32_t cols, uint32_t slice_count, uint32_t slice_rows, const uint32_t *offsets,
                             const uint32_t *widths, const uint32_t *slot_offsets, const uint32_t *col_idx, const void *val,
                             uint32_t vt, const uint8_t *row_keep, const uint8_t *feature_keep, const uint32_t *fgm, uint32_t G,
                             float min_total, uint32_t min_det, float max_frac, uint32_t frac_idx, double *row_totals,
                             uint32_t *detected, float *max_values, uint8_t *keep, double *group_counts, float *group_pct) {
    const entry_ctx x = {cols, feature_keep, fgm, G};
    const reduce_out o = {row_totals, detected, max_values, keep, group_counts, group_pct};
    if (G > 32u || vt_bytes(vt) == 0) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        row_state s;
        memset(&s, 0, sizeof(s));
        if (row_keep != NULL && row_keep[r] == 0u) {
            empty_row(r, G, &o);
            continue;
        }
        if (slice_count != 0u) {
            const uint32_t sl = sell_slice(r, slice_count, slice_rows, offsets);
            const uint32_t w = widths[sl];
            const size_t base = (size_t) slot_offsets[sl] + (size_t) (r - offsets[sl]) * w;
            for (uint32_t k = 0; k < w; ++k) entry(&s, &x, col_idx[base + k], decode(vt, val, base + k));
        }
        finish(r, &s, G, min_total, min_det, max_frac, frac_idx, &o);
    }
    return 1;
}

int oracle_group_reduce_bell(uint32_t rows, uint32_t cols, uint32_t block_size, uint32_t ell_cols, const uint32_t *bci,
                             const void *val, uint32_t vt, const uint8_t *row_keep, const uint8_t *feature_keep,
                             const uint32_t *fgm, uint32_t G, float min_total, uint32_t min_det, float max_frac,
                             uint32_t frac_idx, double *row_totals, uint32_t *detected, float *max_values, uint8_t *keep,
                             double *group_counts, float *group_pct) {
    const entry_ctx x = {cols, feature_keep, fgm, G};
    const reduce_out o = {row_totals, detected, max_values, keep, group_counts, group_pct};
    const uint32_t width = block_size != 0u ? ell_cols / block_size : 0u;
    if (G > 32u || vt_bytes(vt) == 0) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        row_state s;
        memset(&s, 0, sizeof(s));
        if (row_keep != NULL && row_keep[r] == 0u) {
            empty_row(r, G, &o);
            continue;
        }
        const uint32_t rb = block_size != 0u ? r / block_size : 0u;
        for (uint32_t e = 0; e < ell_cols; ++e) {
            const uint32_t slot = block_size != 0u ? e / block_size : 0u;
            const uint32_t lib = block_size != 0u ? e % block_size : 0u;
            const uint32_t bc = width != 0u ? bci[(size_t) rb * width + slot] : INVALID_COL;
            const uint32_t gene = bc != INVALID_COL ? bc * block_size + lib : cols;
            entry(&s, &x, gene, decode(vt, val, (size_t) r * ell_cols + e));
        }
        finish(r, &s, G, min_total, min_det, max_frac, frac_idx, &o);
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Row-label grouped reduction (project extension, no reference function;   */
/* masking rules from mask_groups.cu:65-71,155).                             */
/* ------------------------------------------------------------------------ */
uint32_t oracle_kept_features(uint32_t cols, const uint8_t *feature_keep, uint32_t *ids) {
    uint32_t f = 0u;
    for (uint32_t c = 0; c < cols; ++c)
        if (feature_keep == NULL || feature_keep[c] != 0u) {
            if (ids != NULL) ids[f] = c;
            ++f;
        }
    return f;
}

typedef struct {
    uint32_t cols, G, F;
    const uint32_t *labels;
    const uint8_t *row_keep;
    const uint8_t *feature_keep;
    const float *scale;
    uint32_t *rank; /* column -> kept rank or INVALID */
    double *sum, *sumsq;
    uint32_t *nnz;
} label_ctx;

static int label_begin(label_ctx *L, uint32_t rows, uint32_t *group_rows) {
    L->rank = (uint32_t *) malloc(((size_t) L->cols + 1u) * sizeof(uint32_t));
    if (L->rank == NULL) return 0;
    L->F = 0u;
    for (uint32_t c = 0; c < L->cols; ++c) L->rank[c] = (L->feature_keep == NULL || L->feature_keep[c] != 0u) ? L->F++ : INVALID_COL;
    memset(L->sum, 0, (size_t) L->G * L->F * sizeof(double));
    memset(L->sumsq, 0, (size_t) L->G * L->F * sizeof(double));
    memset(L->nnz, 0, (size_t) L->G * L->F * sizeof(uint32_t));
    memset(group_rows, 0, (size_t) L->G * sizeof(uint32_t));
    for (uint32_t r = 0; r < rows; ++r) {
        const uint32_t g = L->labels[r];
        if ((L->row_keep == NULL || L->row_keep[r] != 0u) && g < L->G) group_rows[g] += 1u;
    }
    return 1;
}

static void label_entry(const label_ctx *L, uint32_t g, uint32_t c, float raw) {
    if (c >= L->cols) return;
    const uint32_t f = L->rank[c];
    if (f == INVALID_COL) return;
    /* the per-feature scale multiplies in FP64, so an entry counts iff both
     * the stored value and its feature's scale are nonzero */
    const double d = L->scale != NULL ? (double) raw * (double) L->scale[c] : (double) raw;
    if (d != 0.0) {
        const size_t o = (size_t) g * L->F + f;
        L->sum[o] += d;
        L->sumsq[o] += d * d;
        L->nnz[o] += 1u;
    }
}

static int row_takes_part(const label_ctx *L, uint32_t r, uint32_t *g) {
    *g = L->labels[r];
    return (L->row_keep == NULL || L->row_keep[r] != 0u) && *g < L->G;
}

int oracle_label_reduce_csr(uint32_t rows, uint32_t cols, const uint32_t *ptr, const uint32_t *idx, const void *val, uint32_t vt,
                            const float *scale, const uint32_t *labels, const uint8_t *row_keep, const uint8_t *feature_keep,
                            uint32_t G, double *sum, double *sumsq, uint32_t *nnz, uint32_t *group_rows) {
    label_ctx L = {cols, G, 0u, labels, row_keep, feature_keep, scale, NULL, sum, sumsq, nnz};
    if (vt_bytes(vt) == 0 || !label_begin(&L, rows, group_rows)) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        uint32_t g;
        if (!row_takes_part(&L, r, &g)) continue;
        for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) label_entry(&L, g, idx[i], decode(vt, val, i));
    }
    free(L.rank);
    return 1;
}

int oracle_label_reduce_sell(uint32_t rows, uint32_t cols, uint32_t slice_count, uint32_t slice_rows, const uint32_t *offsets,
                             const uint32_t *widths, const uint32_t *slot_offsets, const uint32_t *col_idx, const void *val,
                             uint32_t vt, const float *scale, const uint32_t *labels, const uint8_t *row_keep,
                             const uint8_t *feature_keep, uint32_t G, double *sum, double *sumsq, uint32_t *nnz,
                             uint32_t *group_rows) {
    label_ctx L = {cols, G, 0u, labels, row_keep, feature_keep, scale, NULL, sum, sumsq, nnz};
    if (vt_bytes(vt) == 0 || !label_begin(&L, rows, group_rows)) return 0;
    for (uint32_t r = 0; r < rows && slice_count != 0u; ++r) {
        uint32_t g;
        if (!row_takes_part(&L, r, &g)) continue;
        const uint32_t sl = sell_slice(r, slice_count, slice_rows, offsets);
        const uint32_t w = widths[sl];
        const size_t base = (size_t) slot_offsets[sl] + (size_t) (r - offsets[sl]) * w;
        for (uint32_t k = 0; k < w; ++k) label_entry(&L, g, col_idx[base + k], decode(vt, val, base + k));
    }
    free(L.rank);
    return 1;
}

int oracle_label_reduce_bell(uint32_t rows, uint32_t cols, uint32_t block_size, uint32_t ell_cols, const uint32_t *bci,
                             const void *val, uint32_t vt, const float *scale, const uint32_t *labels, const uint8_t *row_keep,
                             const uint8_t *feature_keep, uint32_t G, double *sum, double *sumsq, uint32_t *nnz,
                             uint32_t *group_rows) {
    label_ctx L = {cols, G, 0u, labels, row_keep, feature_keep, scale, NULL, sum, sumsq, nnz};
    const uint32_t width = block_size != 0u ? ell_cols / block_size : 0u;
    if (vt_bytes(vt) == 0 || block_size == 0u || !label_begin(&L, rows, group_rows)) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        uint32_t g;
        if (!row_takes_part(&L, r, &g)) continue;
        const uint32_t rb = r / block_size;
        for (uint32_t e = 0; e < ell_cols; ++e) {
            const uint32_t slot = e / block_size;
            const uint32_t bc = bci[(size_t) rb * width + slot];
            if (bc == INVALID_COL) continue;
            label_entry(&L, g, bc * block_size + e % block_size, decode(vt, val, (size_t) r * ell_cols + e));
        }
    }
    free(L.rank);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* CSR -> Blocked-ELL: src/convert/blocked_ell_from_compressed.cuh:32-91     */
/* (shape), :157-190 (tuner), :258-347 (conversion).  Entries with a column  */
/* >= cols are skipped (the reference indexes past its marks array there).  */
/* ------------------------------------------------------------------------ */
static int cmp_u32(const void *a, const void *b) {
    const uint32_t x = *(const uint32_t *) a, y = *(const uint32_t *) b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int oracle_blocked_ell_shape(uint32_t rows, uint32_t cols, const uint32_t *ptr, const uint32_t *idx, uint32_t bs,
                             uint32_t *out_width, uint64_t *out_total, double *out_fill) {
    if (bs == 0u) return 0;
    *out_width = 0u;
    *out_total = 0u;
    *out_fill = 0.0;
    if (rows == 0u || cols == 0u) return 1;
    const uint32_t col_blocks = (cols + bs - 1u) / bs, row_blocks = (rows + bs - 1u) / bs;
    uint32_t *marks = (uint32_t *) calloc(col_blocks, sizeof(uint32_t));
    if (marks == NULL) return 0;
    uint32_t epoch = 1u, width = 0u;
    uint64_t total = 0u;
    for (uint32_t rb = 0; rb < row_blocks; ++rb) {
        uint32_t unique = 0u;
        const uint32_t r0 = rb * bs, r1 = r0 + bs < rows ? r0 + bs : rows;
        for (uint32_t r = r0; r < r1; ++r)
            for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) {
                if (idx[i] >= cols) continue;
                const uint32_t b = idx[i] / bs;
                if (marks[b] != epoch) {
                    marks[b] = epoch;
                    ++unique;
                }
            }
        total += unique;
        if (unique > width) width = unique;
        if (++epoch == 0u) {
            memset(marks, 0, (size_t) col_blocks * 4u);
            epoch = 1u;
        }
    }
    free(marks);
    *out_width = width;
    *out_total = total;
    *out_fill = (row_blocks == 0u || width == 0u) ? 1.0 : (double) total / ((double) row_blocks * (double) width);
    return 1;
}

int oracle_choose_blocked_ell_block_size(uint32_t rows, uint32_t cols, const uint32_t *ptr, const uint32_t *idx,
                                         const uint32_t *candidates, uint32_t n, uint32_t value_bytes, uint32_t *out_bs,
                                         double *out_fill, uint64_t *out_padded) {
    uint32_t best_bs = 0u;
    double best_fill = 0.0;
    uint64_t best_pad = 0u;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t bs = candidates[i];
        uint32_t w;
        uint64_t total;
        double fill;
        if (bs == 0u) continue;
        if (!oracle_blocked_ell_shape(rows, cols, ptr, idx, bs, &w, &total, &fill)) return 0;
        const uint64_t pad = (uint64_t) rows * w * bs * value_bytes;
        if (best_bs == 0u || pad + 1u < best_pad || (pad == best_pad && fill > best_fill + 1.0e-9) ||
            (pad == best_pad && fill + 1.0e-9 >= best_fill && bs > best_bs)) {
            best_bs = bs;
            best_fill = fill;
            best_pad = pad;
        }
    }
    if (best_bs == 0u) return 0;
    *out_bs = best_bs;
    *out_fill = best_fill;
    *out_padded = best_pad;
    return 1;
}

static int nonzero_bits(uint32_t vb, const uint8_t *p) {
    if (vb == 1u) return (p[0] & 0x7fu) != 0u;
    if (vb == 2u) {
        uint16_t h;
        memcpy(&h, p, 2);
        return (h & 0x7fffu) != 0u;
    }
    float f;
    memcpy(&f, p, 4);
    return f != 0.0f;
}

/* dst arrays: bci [row_blocks * width], val [rows * width * bs * vb] (caller
 * allocated for the width from oracle_blocked_ell_shape). */
int oracle_blocked_ell_from_compressed(uint32_t rows, uint32_t cols, const uint32_t *ptr, const uint32_t *idx, const void *val,
                                       uint32_t vb, uint32_t bs, uint32_t width, uint32_t *bci, void *dval, uint32_t *out_nnz) {
    if (bs == 0u || (vb != 1u && vb != 2u && vb != 4u)) return 0;
    const uint32_t row_blocks = (rows + bs - 1u) / bs, col_blocks = (cols + bs - 1u) / bs;
    const size_t ell_cols = (size_t) width * bs;
    for (size_t i = 0; i < (size_t) row_blocks * width; ++i) bci[i] = INVALID_COL;
    memset(dval, 0, (size_t) rows * ell_cols * vb);
    uint32_t *marks = (uint32_t *) calloc(col_blocks + 1u, sizeof(uint32_t));
    uint32_t *slot_map = (uint32_t *) calloc(col_blocks + 1u, sizeof(uint32_t));
    uint32_t *blocks = (uint32_t *) calloc(width + 1u, sizeof(uint32_t));
    if (marks == NULL || slot_map == NULL || blocks == NULL) {
        free(marks);
        free(slot_map);
        free(blocks);
        return 0;
    }
    uint32_t epoch = 1u;
    int64_t nnz = 0;
    const uint8_t *sv = (const uint8_t *) val;
    uint8_t *dv = (uint8_t *) dval;
    for (uint32_t rb = 0; rb < row_blocks; ++rb) {
        uint32_t unique = 0u;
        const uint32_t r0 = rb * bs, r1 = r0 + bs < rows ? r0 + bs : rows;
        if (++epoch == 0u) {
            memset(marks, 0, (size_t) (col_blocks + 1u) * 4u);
            epoch = 1u;
        }
        for (uint32_t r = r0; r < r1; ++r)
            for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) {
                if (idx[i] >= cols) continue;
                const uint32_t b = idx[i] / bs;
                if (marks[b] != epoch) {
                    marks[b] = epoch;
                    if (unique >= width) {
                        free(marks);
                        free(slot_map);
                        free(blocks);
                        return 0;
                    }
                    blocks[unique++] = b;
                }
            }
        qsort(blocks, unique, sizeof(uint32_t), cmp_u32);
        for (uint32_t s = 0; s < unique; ++s) {
            bci[(size_t) rb * width + s] = blocks[s];
            slot_map[blocks[s]] = s;
        }
        for (uint32_t r = r0; r < r1; ++r)
            for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) {
                if (idx[i] >= cols) continue;
                const uint32_t b = idx[i] / bs;
                uint8_t *d = dv + ((size_t) r * ell_cols + (size_t) slot_map[b] * bs + idx[i] % bs) * vb;
                const int prev = nonzero_bits(vb, d), next = nonzero_bits(vb, sv + (size_t) i * vb);
                if (!prev && next) ++nnz;
                else if (prev && !next) --nnz;
                memcpy(d, sv + (size_t) i * vb, vb);
            }
    }
    free(marks);
    free(slot_map);
    free(blocks);
    *out_nnz = (uint32_t) nnz;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Sliced-ELL: layout src/convert/sliced_ell_from_blocked_ell.cuh:111-142,   */
/* slot offsets runtime/device/sharded_device.cuh:609-620, tuner :64-109 and */
/* :196-221, CSR/CSC emit sliced_ell_from_compressed.cuh:96-164, BELL emit    */
/* sliced_ell_from_blocked_ell.cuh:36-62,233-277.                            */
/* ------------------------------------------------------------------------ */
int oracle_sliced_ell_layout(const uint32_t *row_nnz, uint32_t rows, uint32_t slice_rows, uint32_t *offsets, uint32_t *widths,
                             uint32_t *slot_offsets, uint32_t *out_count, uint64_t *out_total) {
    if (slice_rows == 0u) return 0;
    const uint32_t S = rows == 0u ? 0u : (slice_rows < rows ? slice_rows : rows);
    const uint32_t count = rows == 0u ? 0u : (rows + S - 1u) / S;
    uint64_t total = 0u;
    offsets[0] = 0u;
    for (uint32_t s = 0; s < count; ++s) {
        const uint32_t r0 = s * S, r1 = r0 + S < rows ? r0 + S : rows;
        uint32_t m = 0u;
        for (uint32_t r = r0; r < r1; ++r)
            if (row_nnz[r] > m) m = row_nnz[r];
        offsets[s + 1u] = r1;
        widths[s] = m;
        slot_offsets[s] = (uint32_t) total;
        total += (uint64_t) (r1 - r0) * m;
    }
    *out_count = count;
    *out_total = total;
    return total <= 0xffffffffull;
}

int oracle_compressed_row_nnz(uint32_t rows, uint32_t cols, uint32_t axis, const uint32_t *ptr, const uint32_t *idx, uint32_t *row_nnz) {
    if (axis == 0u) {
        for (uint32_t r = 0; r < rows; ++r) row_nnz[r] = ptr[r + 1u] - ptr[r];
        return 1;
    }
    memset(row_nnz, 0, (size_t) rows * 4u);
    for (uint32_t c = 0; c < cols; ++c)
        for (uint32_t i = ptr[c]; i < ptr[c + 1u]; ++i) {
            if (idx[i] >= rows) return 0;
            row_nnz[idx[i]] += 1u;
        }
    return 1;
}

int oracle_sliced_ell_from_compressed(uint32_t rows, uint32_t cols, uint32_t axis, const uint32_t *ptr, const uint32_t *idx,
                                      const void *val, uint32_t vb, uint32_t slice_count, const uint32_t *offsets,
                                      const uint32_t *widths, const uint32_t *slot_offsets, uint64_t total, uint32_t *col_idx,
                                      void *dval) {
    const uint8_t *sv = (const uint8_t *) val;
    uint8_t *dv = (uint8_t *) dval;
    for (uint64_t k = 0; k < total; ++k) col_idx[k] = INVALID_COL;
    memset(dval, 0, (size_t) total * vb);
    if (axis == 0u) {
        for (uint32_t s = 0; s < slice_count; ++s)
            for (uint32_t r = offsets[s]; r < offsets[s + 1u]; ++r) {
                const size_t base = (size_t) slot_offsets[s] + (size_t) (r - offsets[s]) * widths[s];
                uint32_t k = 0u;
                for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i, ++k) {
                    col_idx[base + k] = idx[i];
                    memcpy(dv + (base + k) * vb, sv + (size_t) i * vb, vb);
                }
            }
        return 1;
    }
    /* CSC source: scatter by row, columns ascending */
    uint32_t *cursor = (uint32_t *) malloc((size_t) (rows + 1u) * 4u);
    if (cursor == NULL) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        const uint32_t s = sell_slice(r, slice_count, 0u, offsets);
        cursor[r] = (uint32_t) (slot_offsets[s] + (size_t) (r - offsets[s]) * widths[s]);
    }
    for (uint32_t c = 0; c < cols; ++c)
        for (uint32_t i = ptr[c]; i < ptr[c + 1u]; ++i) {
            const uint32_t r = idx[i];
            if (r >= rows) {
                free(cursor);
                return 0;
            }
            col_idx[cursor[r]] = c;
            memcpy(dv + (size_t) cursor[r] * vb, sv + (size_t) i * vb, vb);
            cursor[r] += 1u;
        }
    free(cursor);
    return 1;
}

int oracle_blocked_ell_row_nnz(uint32_t rows, uint32_t cols, uint32_t bs, uint32_t ell_cols, const uint32_t *bci, const void *val,
                               uint32_t vb, uint32_t *row_nnz) {
    const uint32_t width = bs != 0u ? ell_cols / bs : 0u;
    if (bs == 0u) return 0;
    for (uint32_t r = 0; r < rows; ++r) {
        uint32_t n = 0u;
        for (uint32_t s = 0; s < width; ++s) {
            const uint32_t b = bci[(size_t) (r / bs) * width + s];
            if (b == INVALID_COL) continue;
            for (uint32_t l = 0; l < bs; ++l) {
                if (b * bs + l >= cols) continue;
                if (nonzero_bits(vb, (const uint8_t *) val + ((size_t) r * ell_cols + (size_t) s * bs + l) * vb)) ++n;
            }
        }
        row_nnz[r] = n;
    }
    return 1;
}

int oracle_sliced_ell_from_blocked_ell(uint32_t rows, uint32_t cols, uint32_t bs, uint32_t ell_cols, const uint32_t *bci,
                                       const void *val, uint32_t vb, uint32_t slice_count, const uint32_t *offsets,
                                       const uint32_t *widths, const uint32_t *slot_offsets, uint64_t total, uint32_t *col_idx,
                                       void *dval) {
    const uint32_t width = bs != 0u ? ell_cols / bs : 0u;
    const uint8_t *sv = (const uint8_t *) val;
    uint8_t *dv = (uint8_t *) dval;
    if (bs == 0u) return 0;
    for (uint64_t k = 0; k < total; ++k) col_idx[k] = INVALID_COL;
    memset(dval, 0, (size_t) total * vb);
    for (uint32_t s = 0; s < slice_count; ++s)
        for (uint32_t r = offsets[s]; r < offsets[s + 1u]; ++r) {
            const size_t base = (size_t) slot_offsets[s] + (size_t) (r - offsets[s]) * widths[s];
            size_t k = 0u;
            for (uint32_t sl = 0; sl < width; ++sl) {
                const uint32_t b = bci[(size_t) (r / bs) * width + sl];
                if (b == INVALID_COL) continue;
                for (uint32_t l = 0; l < bs; ++l) {
                    const uint32_t col = b * bs + l;
                    const uint8_t *p = sv + ((size_t) r * ell_cols + (size_t) sl * bs + l) * vb;
                    if (col >= cols || !nonzero_bits(vb, p)) continue;
                    col_idx[base + k] = col;
                    memcpy(dv + (base + k) * vb, p, vb);
                    ++k;
                }
            }
        }
    return 1;
}

int oracle_choose_sliced_ell_slice_rows(const uint32_t *row_nnz, uint32_t rows, uint32_t nnz, const uint32_t *candidates,
                                        uint32_t n, uint32_t *out_slice_rows, uint32_t *out_count, uint64_t *out_padded,
                                        double *out_score) {
    uint32_t best_s = 0u, best_count = 0u;
    uint64_t best_pad_bytes = 0u, best_pad = 0u;
    double best_score = INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        if (candidates[i] == 0u) continue;
        uint32_t S = 0u, count = 0u;
        uint64_t padded = 0u, pad_bytes = 0u;
        double score = 1.0;
        if (rows != 0u) {
            S = candidates[i] < rows ? candidates[i] : rows;
            for (uint32_t r = 0; r < rows; r += S) {
                const uint32_t r1 = r + S < rows ? r + S : rows;
                uint32_t m = 0u;
                for (uint32_t k = r; k < r1; ++k)
                    if (row_nnz[k] > m) m = row_nnz[k];
                padded += (uint64_t) (r1 - r) * m;
                ++count;
            }
            pad_bytes = padded * 6u + ((uint64_t) count * 2u + 1u) * 4u;
            score = (nnz == 0u ? 1.0 : (double) padded / (double) nnz) + 0.25 * ((double) count / (double) rows);
        }
        if (best_s == 0u || score + 1.0e-12 < best_score || (score <= best_score + 1.0e-12 && pad_bytes < best_pad_bytes) ||
            (score <= best_score + 1.0e-12 && pad_bytes == best_pad_bytes && S > best_s)) {
            best_s = S;
            best_count = count;
            best_pad = padded;
            best_pad_bytes = pad_bytes;
            best_score = score;
        }
    }
    if (best_s == 0u && rows != 0u) return 0;
    *out_slice_rows = best_s;
    *out_count = best_count;
    *out_padded = best_pad;
    *out_score = best_score;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Transpose: counting sort, ascending source index inside each segment.     */
/* Semantics of src/convert/compressed_transpose_raw.cuh:14-122 with the     */
/* canonical (csr2cscEx2) order.                                             */
/* ------------------------------------------------------------------------ */
int oracle_transpose(uint32_t cdim, uint32_t udim, const uint32_t *ptr, const uint32_t *idx, const void *val, uint32_t vb,
                     uint32_t *out_ptr, uint32_t *out_idx, void *out_val) {
    const uint8_t *sv = (const uint8_t *) val;
    uint8_t *dv = (uint8_t *) out_val;
    memset(out_ptr, 0, (size_t) (udim + 1u) * 4u);
    for (uint32_t r = 0; r < cdim; ++r)
        for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) {
            if (idx[i] >= udim) return 0;
            out_ptr[idx[i] + 1u] += 1u;
        }
    for (uint32_t c = 0; c < udim; ++c) out_ptr[c + 1u] += out_ptr[c];
    uint32_t *cur = (uint32_t *) malloc((size_t) (udim + 1u) * 4u);
    if (cur == NULL) return 0;
    memcpy(cur, out_ptr, (size_t) (udim + 1u) * 4u);
    for (uint32_t r = 0; r < cdim; ++r)
        for (uint32_t i = ptr[r]; i < ptr[r + 1u]; ++i) {
            const uint32_t p = cur[idx[i]]++;
            out_idx[p] = r;
            memcpy(dv + (size_t) p * vb, sv + (size_t) i * vb, vb);
        }
    free(cur);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Bucketing: src/bucket/major_nnz_raw.cuh:27-120 (stable ascending nnz      */
/* order, equal-count offsets from fill_equal_count_bucket_offsets.cuh),     */
/* log2 offsets (project rule), rebuild :188-250 with the documented meaning. */
/* ------------------------------------------------------------------------ */
int oracle_bucket_plan(const uint32_t *ptr, uint32_t n, uint32_t requested, uint32_t *nnz, uint32_t *nnz_sorted, uint32_t *order,
                       uint32_t *eq_offsets, uint32_t *out_buckets) {
    uint32_t maxv = 0u;
    for (uint32_t i = 0; i < n; ++i) {
        nnz[i] = ptr[i + 1u] - ptr[i];
        if (nnz[i] > maxv) maxv = nnz[i];
    }
    /* stable counting sort by value */
    uint32_t *cnt = (uint32_t *) calloc((size_t) maxv + 2u, sizeof(uint32_t));
    if (cnt == NULL) return 0;
    for (uint32_t i = 0; i < n; ++i) cnt[nnz[i] + 1u] += 1u;
    for (uint32_t v = 0; v <= maxv; ++v) cnt[v + 1u] += cnt[v];
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t p = cnt[nnz[i]]++;
        order[p] = i;
        nnz_sorted[p] = nnz[i];
    }
    free(cnt);
    uint32_t b = n == 0u ? 1u : (requested < 1u ? 1u : (requested > n ? n : requested));
    if (n == 0u) {
        eq_offsets[0] = 0u;
        *out_buckets = b;
        return 1;
    }
    const uint32_t size = n / b, rem = n % b;
    uint32_t off = 0u;
    for (uint32_t k = 0; k < b; ++k) {
        eq_offsets[k] = off;
        off += size + (k < rem ? 1u : 0u);
    }
    eq_offsets[b] = n;
    *out_buckets = b;
    return 1;
}

/* bucket(n) = clamp(bit_width(n) - min_log2, 0, B - 1) */
uint32_t oracle_log2_bucket(uint32_t v, uint32_t B, uint32_t min_log2) {
    uint32_t bw = 0u;
    while (bw < 32u && (v >> bw) != 0u) ++bw;
    const int b = (int) bw - (int) min_log2;
    if (b < 0) return 0u;
    if ((uint32_t) b > B - 1u) return B - 1u;
    return (uint32_t) b;
}

int oracle_log2_offsets(const uint32_t *nnz_sorted, uint32_t n, uint32_t B, uint32_t min_log2, uint32_t *offsets) {
    for (uint32_t b = 0; b <= B; ++b) offsets[b] = 0u;
    uint32_t *counts = (uint32_t *) calloc(B, sizeof(uint32_t));
    if (counts == NULL) return 0;
    for (uint32_t i = 0; i < n; ++i) counts[oracle_log2_bucket(nnz_sorted[i], B, min_log2)] += 1u;
    for (uint32_t b = 0; b < B; ++b) offsets[b + 1u] = offsets[b] + counts[b];
    free(counts);
    return 1;
}

int oracle_rebuild(const uint32_t *ptr, const uint32_t *idx, const void *val, uint32_t vb, uint32_t n, const uint32_t *order,
                   uint32_t *dst_ptr, uint32_t *dst_idx, void *dst_val, uint32_t *inverse) {
    const uint8_t *sv = (const uint8_t *) val;
    uint8_t *dv = (uint8_t *) dst_val;
    dst_ptr[0] = 0u;
    for (uint32_t d = 0; d < n; ++d) {
        const uint32_t s = order[d];
        const uint32_t b = ptr[s], len = ptr[s + 1u] - b;
        dst_ptr[d + 1u] = dst_ptr[d] + len;
        memcpy(dst_idx + dst_ptr[d], idx + b, (size_t) len * 4u);
        memcpy(dv + (size_t) dst_ptr[d] * vb, sv + (size_t) b * vb, (size_t) len * vb);
        if (inverse != NULL) inverse[s] = d;
    }
    return 1;
}

/* Blocked-ELL row-block keys: src/bucket/operators/blocked_ell_bucket.cuh:77-96 */
void oracle_bell_bucket_keys(uint32_t row_blocks, uint32_t width, const uint32_t *bci, uint64_t *keys) {
    for (uint32_t rb = 0; rb < row_blocks; ++rb) {
        uint32_t valid = 0u, h = 2166136261u;
        for (uint32_t s = 0; s < width; ++s) {
            const uint32_t c = bci[(size_t) rb * width + s];
            if (c == INVALID_COL) continue;
            ++valid;
            h ^= c + 0x9e3779b9u + (h << 6) + (h >> 2);
            h *= 16777619u;
        }
        keys[rb] = ((uint64_t) valid << 32) | (uint64_t) h;
    }
}

/* ------------------------------------------------------------------------ */
/* Masked Blocked-ELL -> CSR: src/runtime/mask_groups.cu:383-462,960-1034.   */
/* Rows in ascending order; entries in ascending ell position.               */
/* ------------------------------------------------------------------------ */
int oracle_reoptimize(uint32_t rows, uint32_t cols, uint32_t bs, uint32_t ell_cols, const uint32_t *bci, const uint16_t *val,
                      const uint8_t *row_keep, const uint8_t *feature_keep, const uint32_t *row_remap,
                      const uint32_t *feature_remap, uint32_t out_rows, uint32_t out_cols, uint32_t *major_ptr,
                      uint32_t *minor_idx, uint16_t *out_val, uint32_t *out_nnz) {
    const uint32_t width = bs != 0u ? ell_cols / bs : 0u;
    memset(major_ptr, 0, (size_t) (out_rows + 1u) * 4u);
    uint32_t *counts = (uint32_t *) calloc((size_t) out_rows + 1u, 4u);
    uint32_t *wr = (uint32_t *) calloc((size_t) out_rows + 1u, 4u);
    if (counts == NULL || wr == NULL) {
        free(counts);
        free(wr);
        return 0;
    }
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t r = 0; r < rows; ++r) {
            if (row_keep != NULL && row_keep[r] == 0u) continue;
            const uint32_t orow = row_remap != NULL ? row_remap[r] : r;
            if (orow >= out_rows) continue;
            uint32_t n = 0u;
            for (uint32_t e = 0; e < ell_cols; ++e) {
                const uint32_t slot = bs != 0u ? e / bs : 0u, lib = bs != 0u ? e % bs : 0u;
                const uint32_t b = width != 0u ? bci[(size_t) (bs != 0u ? r / bs : 0u) * width + slot] : INVALID_COL;
                const uint32_t col = b != INVALID_COL ? b * bs + lib : cols;
                const uint16_t h = val[(size_t) r * ell_cols + e];
                if (col < cols && oracle_half_to_float(h) != 0.0f && (feature_keep == NULL || feature_keep[col] != 0u)) {
                    const uint32_t oc = feature_remap != NULL ? feature_remap[col] : col;
                    if (oc < out_cols) {
                        if (pass == 1) {
                            const uint32_t w = major_ptr[orow] + wr[orow]++;
                            minor_idx[w] = oc;
                            out_val[w] = h;
                        }
                        ++n;
                    }
                }
            }
            if (pass == 0) counts[orow] = n; /* last writer wins, as the reference's count kernel */
        }
        if (pass == 0) {
            uint32_t run = 0u;
            for (uint32_t o = 0; o < out_rows; ++o) {
                major_ptr[o] = run;
                run += counts[o];
            }
            major_ptr[out_rows] = run;
        }
    }
    *out_nnz = major_ptr[out_rows];
    free(counts);
    free(wr);
    return 1;
}

/* ------------------------------------------------------------------------
 * Serving query (project extension, no reference counterpart).  Restated on
 * the row-major CSR, so it does not share the column layout the device
 * path streams.  Seeded labels / mask: lowbias32 hash (Wellons), labels by
 * multiply-high range reduction, mask by the top 24 hash bits.
 * ------------------------------------------------------------------------ */
uint32_t oracle_serve_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

uint32_t oracle_serve_label(uint32_t seed, uint32_t r, uint32_t G) {
    return (uint32_t) (((uint64_t) oracle_serve_hash(seed ^ (r * 0x9E3779B9u)) * G) >> 32);
}

int oracle_serve_keep(uint32_t seed, uint32_t r, uint32_t thr24) {
    return (oracle_serve_hash(seed ^ (r * 0x85EBCA6Bu)) >> 8) < thr24;
}

/* out is [k x G]: f64 for op 0 (sum) / 1 (sumsq), u32 for op 2 (nnz). */
int oracle_serve_query(uint32_t rows, uint32_t cols, const uint32_t *ptr, const uint32_t *idx, const void *val, uint32_t vt,
                       const uint32_t *features, uint32_t k, uint32_t G, uint32_t op, const uint32_t *labels,
                       const uint8_t *keep, uint32_t label_seed, uint32_t mask_seed, uint32_t thr24, void *out) {
    uint32_t *slot = (uint32_t *) malloc(((size_t) cols + 1u) * sizeof(uint32_t));
    if (slot == NULL) return 0;
    for (uint32_t c = 0; c < cols; ++c) slot[c] = INVALID_COL;
    for (uint32_t j = 0; j < k; ++j) slot[features[j]] = j;
    if (op == 2u) memset(out, 0, (size_t) k * G * sizeof(uint32_t));
    else memset(out, 0, (size_t) k * G * sizeof(double));
    for (uint32_t r = 0; r < rows; ++r) {
        const int kept = keep != NULL ? keep[r] != 0u : oracle_serve_keep(mask_seed, r, thr24);
        if (!kept) continue;
        const uint32_t g = labels != NULL ? labels[r] : oracle_serve_label(label_seed, r, G);
        if (g >= G) continue;
        for (uint32_t e = ptr[r]; e < ptr[r + 1u]; ++e) {
            const uint32_t c = idx[e];
            if (c >= cols || slot[c] == INVALID_COL) continue;
            const float v = decode(vt, val, e);
            if (v == 0.0f) continue;
            const size_t o = (size_t) slot[c] * G + g;
            if (op == 2u) ((uint32_t *) out)[o] += 1u;
            else ((double *) out)[o] += op == 1u ? (double) v * (double) v : (double) v;
        }
    }
    free(slot);
    return 1;
}
