// Streaming reference-ABI masked grouped row reduction over row-major
// layouts (compressed rows, Sliced-ELL) — the fast path behind
// compute_sparse_group_reduce / _compressed_fallback (mask_groups.hip).
#pragma once

#include "common.hh"

namespace csamd {

struct gr_outputs {
    float *row_totals;
    float *max_values;
    unsigned int *detected_features;
    unsigned char *row_keep;
    float *group_counts;       // null when group_count == 0
    float *group_percentages;
};

// A row-major source: compressed rows (ptr != null) or Sliced-ELL slices.
struct gr_source {
    unsigned int rows;
    unsigned int cols;
    const unsigned int *ptr;
    unsigned int slice_count;
    unsigned int slice_rows;            // uniform slice height, 0 if not uniform
    const unsigned int *slice_offsets;
    const unsigned int *slice_widths;
    const unsigned int *slot_offsets;
    const unsigned int *idx;
    const void *val;
    unsigned int value_type;
};

// A Blocked-ELL source: rows x ell_cols values, row-major; row r's slot s
// (block_size entries) holds columns block_col_idx[(r / block_size) * (ell_cols /
// block_size) + s] * block_size + [0, block_size), or none when that index is
// CSAMD_INVALID_COL.
struct gr_bell_source {
    unsigned int rows;
    unsigned int cols;
    unsigned int block_size;
    unsigned int ell_cols;
    const unsigned int *block_col_idx;
    const void *val;
    unsigned int value_type;
};

// Device scratch the streaming path needs for a source of `cols` columns
// (feature map + kept-feature group table + workgroup row bounds).
size_t gr_scratch_bytes(unsigned int cols);
// Whether the streaming path takes this source (unit-aligned arrays, a
// column count whose feature map fits in LDS, a decodable value type);
// otherwise the per-row kernels of mask_groups.hip run.
bool gr_eligible(const gr_source &src);
// Enqueue the reduction on `stream` (asynchronous).  fgm may be null when
// group_count == 0.  scratch must hold gr_scratch_bytes(src.cols) bytes,
// 256-byte aligned.
int gr_reduce_rows(const gr_source &src, const uint32_t *fgm, unsigned int group_count, const unsigned char *row_keep,
                   const unsigned char *feature_keep, const csamd_sparse_group_filter_params &filter, const gr_outputs &out,
                   void *scratch, size_t scratch_bytes, hipStream_t stream);

// Whether the Blocked-ELL streaming path takes this source: block size 8, 16
// or 32, values aligned for 8-value loads, a feature map that fits in LDS.
bool gr_bell_eligible(const gr_bell_source &src);
// As gr_reduce_rows, over a Blocked-ELL source (scratch: gr_scratch_bytes(cols)).
int gr_reduce_bell(const gr_bell_source &src, const uint32_t *fgm, unsigned int group_count, const unsigned char *row_keep,
                   const unsigned char *feature_keep, const csamd_sparse_group_filter_params &filter, const gr_outputs &out,
                   void *scratch, size_t scratch_bytes, hipStream_t stream);

} // namespace csamd
