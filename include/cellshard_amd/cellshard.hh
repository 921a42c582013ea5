// C++ drop-in layer over the cellshard_amd C ABI, with the reference's
// namespaces, type names and function signatures for the sparse hot path
// (cudaStream_t -> hipStream_t).  Header-only; link libcellshard_amd.so.
//
//   cellshard::sparse    compressed / blocked_ell / sliced_ell host structs
//                        (formats/{compressed,blocked_ell,sliced_ell}.cuh)
//   cellshard::device    compressed_view / blocked_ell_view / sliced_ell_view,
//                        partition_record, upload/release, allocation cache
//                        (runtime/device/sharded_device.cuh)
//   cellshard::runtime   mask_groups (runtime/mask_groups.cuh)
//   cellshard::convert   Blocked-ELL / Sliced-ELL converters and tuners,
//                        build_compressed_transpose_raw (src/convert/)
//   cellshard::bucket    major-nnz bucket plan / rebuild (src/bucket/major_nnz_raw.cuh)
//   cellshard::sharded   partition / shard metadata (runtime/layout/sharded.cuh)
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#ifndef CSAMD_HALF_T
#define CSAMD_HALF_T __half
#endif
#endif

#include "cellshard_amd.h"

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#if !defined(__HIPCC__)
#if defined(__has_include) && __has_include(<hip/hip_runtime_api.h>) && defined(__HIP_PLATFORM_AMD__)
#include <hip/hip_runtime_api.h>
#else
typedef struct ihipStream_t *hipStream_t;
typedef int hipError_t;
#endif
#endif

namespace cellshard {

namespace types {
using u32 = std::uint32_t;
using i32 = std::int32_t;
using u64 = std::uint64_t;
using dim_t = u32;
using nnz_t = u32;
using idx_t = u32;
using ptr_t = u32;
} // namespace types

namespace real {
using storage_t = csamd_half;
using compute_t = float;
using accum_t = float;
enum { value_f16 = CSAMD_VALUE_F16, value_f32 = CSAMD_VALUE_F32, value_f64 = CSAMD_VALUE_F64,
       value_bf16 = CSAMD_VALUE_BF16, value_fp8_e4m3 = CSAMD_VALUE_FP8_E4M3, value_fp8_e5m2 = CSAMD_VALUE_FP8_E5M2 };
} // namespace real

// Binary search over offsets[i] <= row < offsets[i+1] (core/offset_span.cuh:12).
template<typename OffsetT>
inline OffsetT find_offset_span(OffsetT row, const OffsetT *offsets, OffsetT count) {
    OffsetT lo = 0, hi = count;
    while (lo < hi) {
        const OffsetT mid = lo + ((hi - lo) >> 1);
        if (row < offsets[mid]) hi = mid;
        else if (row >= offsets[mid + 1]) lo = mid + 1;
        else return mid;
    }
    return count;
}

// ---------------------------------------------------------------------------
namespace sparse {

enum { compressed_by_row = CSAMD_COMPRESSED_BY_ROW, compressed_by_col = CSAMD_COMPRESSED_BY_COL };
enum { compressed_host_registered = 1u, blocked_ell_host_registered = 1u, sliced_ell_host_registered = 1u };
enum : std::uint32_t { blocked_ell_invalid_col = CSAMD_INVALID_COL, sliced_ell_invalid_col = CSAMD_INVALID_COL };

using compressed = csamd_host_compressed;
using blocked_ell = csamd_host_blocked_ell;
using sliced_ell = csamd_host_sliced_ell;
using coo = csamd_host_coo;

inline types::dim_t major_dim(const compressed *m) { return m->axis == compressed_by_col ? m->cols : m->rows; }
inline types::dim_t minor_dim(const compressed *m) { return m->axis == compressed_by_col ? m->rows : m->cols; }

inline void init(compressed *m, types::dim_t rows = 0, types::dim_t cols = 0, types::nnz_t nnz = 0,
                 types::u32 axis = compressed_by_row) {
    std::memset(m, 0, sizeof(*m));
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->axis = axis;
}
inline void init(blocked_ell *m, types::dim_t rows = 0, types::dim_t cols = 0, types::nnz_t nnz = 0, types::u32 block_size = 0,
                 types::u32 ell_cols = 0) {
    std::memset(m, 0, sizeof(*m));
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->block_size = block_size;
    m->ell_cols = ell_cols;
}
inline void init(sliced_ell *m, types::dim_t rows = 0, types::dim_t cols = 0, types::nnz_t nnz = 0) {
    std::memset(m, 0, sizeof(*m));
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
}

inline void clear(compressed *m) { csamd_host_compressed_clear(m); }
inline void clear(blocked_ell *m) { csamd_host_blocked_ell_clear(m); }
inline void clear(sliced_ell *m) { csamd_host_sliced_ell_clear(m); }

inline std::size_t align_up_(std::size_t v, std::size_t a) { return (v + a - 1u) / a * a; }

// formats/compressed.cuh:122-153 (one packed block: ptr | idx | val)
inline int allocate(compressed *m) {
    const std::size_t ptr_count = (std::size_t) major_dim(m) + 1u;
    const std::size_t minor_off = align_up_(ptr_count * 4u, 4u);
    const std::size_t val_off = align_up_(minor_off + (std::size_t) m->nnz * 4u, 2u);
    const std::size_t total = val_off + (std::size_t) m->nnz * 2u;
    const types::dim_t rows = m->rows, cols = m->cols, nnz = m->nnz;
    const types::u32 axis = m->axis;
    clear(m);
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->axis = axis;
    if (total == 0u) return 1;
    m->storage = std::malloc(total);
    if (m->storage == nullptr) return 0;
    m->majorPtr = (types::ptr_t *) m->storage;
    m->minorIdx = nnz != 0u ? (types::idx_t *) ((char *) m->storage + minor_off) : nullptr;
    m->val = nnz != 0u ? (real::storage_t *) ((char *) m->storage + val_off) : nullptr;
    return 1;
}

inline types::dim_t row_block_count(const blocked_ell *m) { return m->block_size == 0u ? 0u : (m->rows + m->block_size - 1u) / m->block_size; }
inline types::dim_t col_block_count(const blocked_ell *m) { return m->block_size == 0u ? 0u : (m->cols + m->block_size - 1u) / m->block_size; }
inline types::u32 ell_width_blocks(const blocked_ell *m) {
    return (m->block_size == 0u || m->ell_cols == 0u) ? 0u : m->ell_cols / m->block_size;
}
inline unsigned long pack_blocked_ell_aux(types::u32 block_size, unsigned long ell_width) {
    return (ell_width << 16u) | (unsigned long) block_size;
}

// formats/blocked_ell.cuh:134-158
inline int allocate(blocked_ell *m) {
    const std::size_t idx_bytes = (std::size_t) row_block_count(m) * ell_width_blocks(m) * 4u;
    const std::size_t val_off = align_up_(idx_bytes, 2u);
    const std::size_t total = val_off + (std::size_t) m->rows * m->ell_cols * 2u;
    blocked_ell keep = *m;
    keep.storage = nullptr;
    clear(m);
    *m = keep;
    m->flags = 0u;
    m->blockColIdx = nullptr;
    m->val = nullptr;
    if (total == 0u) return 1;
    m->storage = std::malloc(total);
    if (m->storage == nullptr) return 0;
    m->blockColIdx = idx_bytes != 0u ? (types::idx_t *) m->storage : nullptr;
    m->val = m->rows != 0u && m->ell_cols != 0u ? (real::storage_t *) ((char *) m->storage + val_off) : nullptr;
    return 1;
}

inline types::u32 total_slots(const sliced_ell *m) {
    types::u32 t = 0u;
    if (m == nullptr || m->slice_row_offsets == nullptr || m->slice_widths == nullptr) return 0u;
    for (types::u32 s = 0; s < m->slice_count; ++s) t += (m->slice_row_offsets[s + 1u] - m->slice_row_offsets[s]) * m->slice_widths[s];
    return t;
}
inline unsigned long pack_sliced_ell_aux(types::u32 slice_count, types::u32 slots) {
    return ((unsigned long) slots << 32u) | (unsigned long) slice_count;
}
inline types::u32 uniform_slice_rows(const sliced_ell *m) {
    if (m == nullptr || m->slice_count == 0u || m->slice_row_offsets == nullptr) return 0u;
    if (m->slice_count == 1u) return m->rows;
    const types::u32 step = m->slice_row_offsets[1] - m->slice_row_offsets[0];
    if (step == 0u) return 0u;
    for (types::u32 s = 1; s + 1u < m->slice_count; ++s)
        if (m->slice_row_offsets[s + 1u] - m->slice_row_offsets[s] != step) return 0u;
    if (m->slice_row_offsets[m->slice_count] - m->slice_row_offsets[m->slice_count - 1u] > step) return 0u;
    return step;
}

// formats/sliced_ell.cuh:70-79: first slot of slice `slice_id` (slices are
// stored back to back); 0 past the last slice boundary.
inline std::size_t slice_slot_base(const sliced_ell *m, types::u32 slice_id) {
    if (m == nullptr || m->slice_row_offsets == nullptr || m->slice_widths == nullptr || slice_id > m->slice_count) return 0u;
    std::size_t base = 0u;
    for (types::u32 s = 0; s < slice_id; ++s)
        base += (std::size_t) (m->slice_row_offsets[s + 1u] - m->slice_row_offsets[s]) * m->slice_widths[s];
    return base;
}
// formats/sliced_ell.cuh:81-88: the slice holding `row` (slice_count when
// none).  Offsets ascend, so a binary search gives the reference's answer.
inline types::u32 find_slice(const sliced_ell *m, types::dim_t row) {
    if (m == nullptr) return 0u;
    if (m->slice_row_offsets == nullptr || row >= m->rows) return m->slice_count;
    types::u32 lo = 0u, hi = m->slice_count;  // first slice whose end is > row
    while (lo < hi) {
        const types::u32 mid = lo + ((hi - lo) >> 1);
        if (row < m->slice_row_offsets[mid + 1u]) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}
// formats/sliced_ell.cuh:113-125: valid (column != invalid) slots of a row.
inline types::u32 row_nnz(const sliced_ell *m, types::dim_t row) {
    const types::u32 slice = find_slice(m, row);
    if (m == nullptr || slice >= m->slice_count) return 0u;
    const types::u32 width = m->slice_widths[slice];
    const std::size_t base = slice_slot_base(m, slice) + (std::size_t) (row - m->slice_row_offsets[slice]) * width;
    types::u32 n = 0u;
    for (types::u32 k = 0; k < width; ++k) n += m->col_idx[base + k] != sliced_ell_invalid_col ? 1u : 0u;
    return n;
}

// Host footprint of one materialized part (formats/compressed.cuh:69-74,
// blocked_ell.cuh:78-82, sliced_ell.cuh:100-111): the struct plus its arrays.
inline std::size_t bytes(const compressed *m) {
    return sizeof(*m) + (std::size_t) (major_dim(m) + 1u) * 4u + (std::size_t) m->nnz * 4u + (std::size_t) m->nnz * 2u;
}
inline std::size_t bytes(const blocked_ell *m) {
    return sizeof(*m) + (std::size_t) row_block_count(m) * ell_width_blocks(m) * 4u + (std::size_t) m->rows * m->ell_cols * 2u;
}
inline std::size_t bytes(const sliced_ell *m) {
    const std::size_t slots = total_slots(m);
    const std::size_t w_off = align_up_(((std::size_t) m->slice_count + 1u) * 4u, 4u);
    const std::size_t col_off = align_up_(w_off + (std::size_t) m->slice_count * 4u, 4u);
    const std::size_t val_off = align_up_(col_off + slots * 4u, 2u);
    return sizeof(*m) + val_off + slots * 2u;
}

// formats/sliced_ell.cuh:165-224
inline int allocate(sliced_ell *m, types::u32 slice_count, const types::u32 *offsets, const types::u32 *widths) {
    if (m == nullptr || offsets == nullptr || widths == nullptr) return 0;
    if (slice_count != 0u && offsets[slice_count] != m->rows) return 0;
    std::size_t total = 0u;
    for (types::u32 s = 0; s < slice_count; ++s) total += (std::size_t) (offsets[s + 1u] - offsets[s]) * widths[s];
    const std::size_t w_off = align_up_((std::size_t) (slice_count + 1u) * 4u, 4u);
    const std::size_t col_off = align_up_(w_off + (std::size_t) slice_count * 4u, 4u);
    const std::size_t val_off = align_up_(col_off + total * 4u, 2u);
    const std::size_t bytes = val_off + total * 2u;
    const types::dim_t rows = m->rows, cols = m->cols, nnz = m->nnz;
    clear(m);
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->slice_count = slice_count;
    if (bytes == 0u) return 1;
    m->storage = std::malloc(bytes);
    if (m->storage == nullptr) {
        m->slice_count = 0u;
        return 0;
    }
    char *b = (char *) m->storage;
    m->slice_row_offsets = (types::u32 *) b;
    m->slice_widths = (types::u32 *) (b + w_off);
    m->col_idx = (types::idx_t *) (b + col_off);
    m->val = (real::storage_t *) (b + val_off);
    std::memcpy(m->slice_row_offsets, offsets, (std::size_t) (slice_count + 1u) * 4u);
    if (slice_count != 0u) std::memcpy(m->slice_widths, widths, (std::size_t) slice_count * 4u);
    for (std::size_t k = 0; k < total; ++k) m->col_idx[k] = sliced_ell_invalid_col;
    if (total != 0u) std::memset((void *) m->val, 0, total * 2u);
    return 1;
}

} // namespace sparse

// ---------------------------------------------------------------------------
namespace device {

using compressed_view = csamd_compressed_view;
using blocked_ell_view = csamd_blocked_ell_view;
using sliced_ell_view = csamd_sliced_ell_view;

template<typename MatrixT>
struct partition_record : csamd_partition_record {};

using cached_allocation = csamd_cached_allocation;

template<typename MatrixT>
inline void zero_record(partition_record<MatrixT> *r) {
    std::memset(r, 0, sizeof(*r));
    r->device_id = -1;
}

// runtime/device/sharded_device.cuh:343-755: one packed device allocation,
// H2D copies of the arrays and of the view descriptor.  `cache` may be null.
inline hipError_t upload(const sparse::compressed *src, partition_record<sparse::compressed> *r, csamd_allocation_cache *cache = nullptr) {
    return csamd_upload_compressed(src, r, nullptr, cache, nullptr) ? (hipError_t) 0 : (hipError_t) 1;
}
inline hipError_t upload_async(const sparse::compressed *src, partition_record<sparse::compressed> *r, hipStream_t s,
                               csamd_allocation_cache *cache = nullptr) {
    return csamd_upload_compressed(src, r, nullptr, cache, (csamd_stream_t) s) ? (hipError_t) 0 : (hipError_t) 1;
}
inline hipError_t upload(const sparse::blocked_ell *src, partition_record<sparse::blocked_ell> *r, csamd_allocation_cache *cache = nullptr) {
    return csamd_upload_blocked_ell(src, r, nullptr, cache, nullptr) ? (hipError_t) 0 : (hipError_t) 1;
}
inline hipError_t upload_async(const sparse::blocked_ell *src, partition_record<sparse::blocked_ell> *r, hipStream_t s,
                               csamd_allocation_cache *cache = nullptr) {
    return csamd_upload_blocked_ell(src, r, nullptr, cache, (csamd_stream_t) s) ? (hipError_t) 0 : (hipError_t) 1;
}
inline hipError_t upload(const sparse::sliced_ell *src, partition_record<sparse::sliced_ell> *r, csamd_allocation_cache *cache = nullptr) {
    return csamd_upload_sliced_ell(src, r, nullptr, cache, nullptr) ? (hipError_t) 0 : (hipError_t) 1;
}
inline hipError_t upload_async(const sparse::sliced_ell *src, partition_record<sparse::sliced_ell> *r, hipStream_t s,
                               csamd_allocation_cache *cache = nullptr) {
    return csamd_upload_sliced_ell(src, r, nullptr, cache, (csamd_stream_t) s) ? (hipError_t) 0 : (hipError_t) 1;
}
template<typename MatrixT>
inline hipError_t release(partition_record<MatrixT> *r, csamd_allocation_cache *cache = nullptr) {
    return csamd_release_record(r, cache) ? (hipError_t) 0 : (hipError_t) 1;
}

} // namespace device

// ---------------------------------------------------------------------------
namespace runtime {

static constexpr unsigned int CELLSHARD_MAX_MASK_GROUPS = CSAMD_MAX_MASK_GROUPS;

using sparse_mask_layout = csamd_sparse_mask_layout;
static constexpr sparse_mask_layout sparse_mask_layout_unknown = CSAMD_LAYOUT_UNKNOWN;
static constexpr sparse_mask_layout sparse_mask_layout_blocked_ell = CSAMD_LAYOUT_BLOCKED_ELL;
static constexpr sparse_mask_layout sparse_mask_layout_sliced_ell = CSAMD_LAYOUT_SLICED_ELL;
static constexpr sparse_mask_layout sparse_mask_layout_compressed = CSAMD_LAYOUT_COMPRESSED;

using group_mask_config_view = csamd_group_mask_config_view;
using row_feature_mask_view = csamd_row_feature_mask_view;
using sparse_group_filter_params = csamd_sparse_group_filter_params;
using sparse_group_reduce_result = csamd_sparse_group_reduce_result;
using sparse_group_reduce_workspace = csamd_sparse_group_reduce_workspace;
using sparse_group_reduce_fleet_result = csamd_sparse_group_reduce_fleet_result;
using sparse_group_reduce_fleet_workspace = csamd_sparse_group_reduce_fleet_workspace;
using sparse_group_reduce_fleet_config = csamd_sparse_group_reduce_fleet_config;
using masked_sparse_reoptimize_config = csamd_masked_sparse_reoptimize_config;
using masked_sparse_reoptimize_metadata = csamd_masked_sparse_reoptimize_metadata;
using masked_sparse_reoptimize_result = csamd_masked_sparse_reoptimize_result;
using masked_sparse_reoptimize_workspace = csamd_masked_sparse_reoptimize_workspace;
using auto_reoptimize_prediction_input = csamd_auto_reoptimize_prediction_input;
using auto_reoptimize_prediction = csamd_auto_reoptimize_prediction;

inline int predict_masked_sparse_reoptimization(const auto_reoptimize_prediction_input *in, auto_reoptimize_prediction *out) {
    return csamd_predict_masked_sparse_reoptimization(in, out);
}

inline void init(sparse_group_reduce_workspace *w) { csamd_group_reduce_init(w); }
inline void init(sparse_group_reduce_fleet_workspace *f) { csamd_fleet_init(f); }
inline void init(masked_sparse_reoptimize_workspace *w) { csamd_reoptimize_init(w); }
inline void clear(sparse_group_reduce_workspace *w) { csamd_group_reduce_clear(w); }
inline void clear(sparse_group_reduce_fleet_workspace *f) { csamd_fleet_clear(f); }
inline void clear(masked_sparse_reoptimize_workspace *w) { csamd_reoptimize_clear(w); }

inline int setup(sparse_group_reduce_workspace *w, int device, hipStream_t stream = (hipStream_t) 0) {
    return csamd_group_reduce_setup(w, device, (csamd_stream_t) stream);
}
inline int setup_fleet(sparse_group_reduce_fleet_workspace *f, const sparse_group_reduce_fleet_config *config = nullptr) {
    return csamd_setup_fleet(f, config);
}
inline int setup(masked_sparse_reoptimize_workspace *w, int device, hipStream_t stream = (hipStream_t) 0) {
    return csamd_reoptimize_setup(w, device, (csamd_stream_t) stream);
}
inline int reserve(sparse_group_reduce_workspace *w, unsigned int rows, unsigned int cols, unsigned int values, unsigned int groups) {
    return csamd_group_reduce_reserve(w, rows, cols, values, groups);
}
inline int upload_feature_group_masks(sparse_group_reduce_workspace *w, unsigned int cols, const std::uint32_t *host_masks) {
    return csamd_upload_feature_group_masks(w, cols, host_masks);
}
inline int compute_sparse_group_reduce(const device::blocked_ell_view *src, sparse_group_reduce_workspace *w,
                                       const group_mask_config_view *groups, const row_feature_mask_view *masks,
                                       const sparse_group_filter_params *filter, sparse_group_reduce_result *out) {
    return csamd_compute_sparse_group_reduce_blocked_ell(src, w, groups, masks, filter, out);
}
inline int compute_sparse_group_reduce(const device::sliced_ell_view *src, sparse_group_reduce_workspace *w,
                                       const group_mask_config_view *groups, const row_feature_mask_view *masks,
                                       const sparse_group_filter_params *filter, sparse_group_reduce_result *out) {
    return csamd_compute_sparse_group_reduce_sliced_ell(src, w, groups, masks, filter, out);
}
inline int compute_sparse_group_reduce_compressed_fallback(const device::compressed_view *src, sparse_group_reduce_workspace *w,
                                                           const group_mask_config_view *groups, const row_feature_mask_view *masks,
                                                           const sparse_group_filter_params *filter, sparse_group_reduce_result *out) {
    return csamd_compute_sparse_group_reduce_compressed(src, w, groups, masks, filter, out);
}
inline int compute_sparse_group_reduce_fleet(const device::blocked_ell_view *src_by_slot, sparse_group_reduce_fleet_workspace *f,
                                             const group_mask_config_view *groups, const row_feature_mask_view *masks,
                                             const sparse_group_filter_params *filter, sparse_group_reduce_fleet_result *out) {
    return csamd_compute_sparse_group_reduce_fleet_blocked_ell(src_by_slot, f, groups, masks, filter, out);
}
inline int compute_sparse_group_reduce_fleet(const device::sliced_ell_view *src_by_slot, sparse_group_reduce_fleet_workspace *f,
                                             const group_mask_config_view *groups, const row_feature_mask_view *masks,
                                             const sparse_group_filter_params *filter, sparse_group_reduce_fleet_result *out) {
    return csamd_compute_sparse_group_reduce_fleet_sliced_ell(src_by_slot, f, groups, masks, filter, out);
}
inline int manual_reoptimize_masked_sparse(const device::blocked_ell_view *src, const row_feature_mask_view *masks,
                                           const masked_sparse_reoptimize_config *config, masked_sparse_reoptimize_workspace *w,
                                           masked_sparse_reoptimize_result *out) {
    return csamd_manual_reoptimize_masked_sparse(src, masks, config, w, out);
}

} // namespace runtime

// ---------------------------------------------------------------------------
namespace convert {

using blocked_ell_tune_result = csamd_blocked_ell_tune_result;
using sliced_ell_tune_result = csamd_sliced_ell_tune_result;

inline int choose_blocked_ell_block_size(const sparse::compressed *src, const unsigned int *candidates, unsigned int count,
                                         blocked_ell_tune_result *out) {
    return csamd_host_choose_blocked_ell_block_size(src, candidates, count, out);
}
inline int blocked_ell_from_compressed(const sparse::compressed *src, unsigned int block_size, sparse::blocked_ell *dst) {
    return csamd_host_blocked_ell_from_compressed(src, block_size, dst);
}
inline int blocked_ell_from_compressed_auto(const sparse::compressed *src, const unsigned int *candidates, unsigned int count,
                                            sparse::blocked_ell *dst, blocked_ell_tune_result *picked = nullptr) {
    return csamd_host_blocked_ell_from_compressed_auto(src, candidates, count, dst, picked);
}
inline int choose_sliced_ell_slice_rows_from_compressed(const sparse::compressed *src, const unsigned int *candidates,
                                                        unsigned int count, sliced_ell_tune_result *out) {
    return csamd_host_choose_sliced_ell_slice_rows(src, candidates, count, out);
}
inline int sliced_ell_from_compressed(const sparse::compressed *src, unsigned int slice_rows, sparse::sliced_ell *dst) {
    return csamd_host_sliced_ell_from_compressed(src, slice_rows, dst);
}
inline int sliced_ell_from_compressed_cuda(const sparse::compressed *src, unsigned int slice_rows, sparse::sliced_ell *dst,
                                           int device = 0, hipStream_t stream = (hipStream_t) 0) {
    (void) stream;
    if (!csamd_set_device(device)) return 0;
    return csamd_host_sliced_ell_from_compressed(src, slice_rows, dst);
}
inline int sliced_ell_from_compressed_auto(const sparse::compressed *src, const unsigned int *candidates, unsigned int count,
                                           sparse::sliced_ell *dst, sliced_ell_tune_result *picked = nullptr) {
    return csamd_host_sliced_ell_from_compressed_auto(src, candidates, count, dst, picked);
}
inline int sliced_ell_from_blocked_ell(const sparse::blocked_ell *src, unsigned int slice_rows, sparse::sliced_ell *dst) {
    return csamd_host_sliced_ell_from_blocked_ell(src, slice_rows, dst);
}
inline int sliced_ell_from_blocked_ell_cuda(const sparse::blocked_ell *src, unsigned int slice_rows, sparse::sliced_ell *dst,
                                            int device = 0, hipStream_t stream = (hipStream_t) 0) {
    (void) stream;
    if (!csamd_set_device(device)) return 0;
    return csamd_host_sliced_ell_from_blocked_ell(src, slice_rows, dst);
}

// src/convert/compressed_transpose_raw.cuh:14-27 (same parameters; d_heads is
// unused, d_scan_tmp must hold csamd_compressed_transpose_scratch_bytes()).
static inline int build_compressed_transpose_raw(types::dim_t cDim, types::dim_t uDim, types::nnz_t nnz, const types::ptr_t *d_cAxPtr,
                                                 const types::idx_t *d_uAxIdx, const real::storage_t *d_val, types::ptr_t *d_out_uAxPtr,
                                                 types::ptr_t *d_heads, types::idx_t *d_out_cAxIdx, real::storage_t *d_out_val,
                                                 void *d_scan_tmp, std::size_t scan_bytes, hipStream_t stream) {
    return csamd_build_compressed_transpose_raw(cDim, uDim, nnz, d_cAxPtr, d_uAxIdx, d_val, d_out_uAxPtr, d_heads, d_out_cAxIdx,
                                                d_out_val, d_scan_tmp, scan_bytes, (csamd_stream_t) stream);
}
// src/convert/blocked_ell_from_coo_cuda.cuh:612-966 (the workspace argument
// is accepted for the signature; the library keeps its own).
inline int choose_blocked_ell_block_size_from_coo_cuda(const sparse::coo *src, types::dim_t cols, const types::u32 *feature_to_global,
                                                       const unsigned int *candidates, unsigned int candidate_count, int device,
                                                       blocked_ell_tune_result *out, void *workspace = nullptr,
                                                       hipStream_t stream = (hipStream_t) 0) {
    (void) workspace;
    (void) stream;
    return csamd_choose_blocked_ell_block_size_from_coo(src, cols, feature_to_global, candidates, candidate_count, device, out);
}
inline int blocked_ell_from_coo_cuda(const sparse::coo *src, types::dim_t cols, const types::u32 *feature_to_global,
                                     unsigned int block_size, sparse::blocked_ell *dst, int device, void *workspace = nullptr,
                                     hipStream_t stream = (hipStream_t) 0) {
    (void) workspace;
    (void) stream;
    return csamd_blocked_ell_from_coo(src, cols, feature_to_global, block_size, dst, device);
}
inline int blocked_ell_from_coo_cuda_auto(const sparse::coo *src, types::dim_t cols, const types::u32 *feature_to_global,
                                          const unsigned int *candidates, unsigned int candidate_count, sparse::blocked_ell *dst,
                                          int device, blocked_ell_tune_result *picked = nullptr, void *workspace = nullptr,
                                          hipStream_t stream = (hipStream_t) 0) {
    (void) workspace;
    (void) stream;
    return csamd_blocked_ell_from_coo_auto(src, cols, feature_to_global, candidates, candidate_count, dst, device, picked);
}

static inline int build_transpose_cs_from_cs_raw(types::dim_t cDim, types::dim_t uDim, types::nnz_t nnz, const types::ptr_t *a,
                                                 const types::idx_t *b, const real::storage_t *c, types::ptr_t *d, types::ptr_t *e,
                                                 types::idx_t *f, real::storage_t *g, void *h, std::size_t i, hipStream_t s) {
    return build_compressed_transpose_raw(cDim, uDim, nnz, a, b, c, d, e, f, g, h, i, s);
}

// Masked Blocked-ELL -> CSR (convert/filtered_blocked_ell_to_compressed.cuh:125-415).
using filtered_blocked_ell_stats = csamd_filtered_blocked_ell_stats;
using filtered_blocked_ell_result = csamd_filtered_blocked_ell_result;
using filtered_blocked_ell_workspace = csamd_filtered_blocked_ell_workspace;

inline void init(filtered_blocked_ell_workspace *ws) { csamd_filtered_blocked_ell_init(ws); }
inline void clear(filtered_blocked_ell_workspace *ws) { csamd_filtered_blocked_ell_clear(ws); }
inline int setup(filtered_blocked_ell_workspace *ws, int device, hipStream_t stream = (hipStream_t) 0) {
    return csamd_filtered_blocked_ell_setup(ws, device, (csamd_stream_t) stream);
}
inline int filter_blocked_ell_to_compressed(const device::blocked_ell_view *src, types::dim_t output_rows,
                                                     types::dim_t output_cols, const unsigned char *d_keep_rows,
                                                     const unsigned char *d_keep_cols, const types::idx_t *d_row_remap,
                                                     const types::idx_t *d_col_remap, filtered_blocked_ell_workspace *ws,
                                                     filtered_blocked_ell_result *out) {
    return csamd_filter_blocked_ell_to_compressed(src, CSAMD_VALUE_F16, output_rows, output_cols, d_keep_rows, d_keep_cols,
                                                  d_row_remap, d_col_remap, ws, out);
}
inline int build_bucketed_filtered_blocked_ell_major_view(const device::blocked_ell_view *src, types::dim_t output_rows,
                                                                   types::dim_t output_cols, const unsigned char *d_keep_rows,
                                                                   const unsigned char *d_keep_cols, const types::idx_t *d_row_remap,
                                                                   const types::idx_t *d_col_remap,
                                                                   types::idx_t requested_bucket_count,
                                                                   filtered_blocked_ell_workspace *ws,
                                                                   filtered_blocked_ell_result *out) {
    return csamd_build_bucketed_filtered_blocked_ell_major_view(src, CSAMD_VALUE_F16, output_rows, output_cols, d_keep_rows,
                                                                d_keep_cols, d_row_remap, d_col_remap, requested_bucket_count, ws,
                                                                out);
}

} // namespace convert

// Shard repack: masked CSR -> bucketed CSR -> Blocked-ELL
// (repack/operators/sharded_blocked_ell.cuh:17-723).  The reference reads the
// shard's partitions from a sharded_device state; here the caller passes
// the partitions' views (host array, device arrays) in shard order.
namespace repack {
using sharded_blocked_ell_repack_config = csamd_sharded_blocked_ell_repack_config;
using sharded_blocked_ell_repack_stats = csamd_sharded_blocked_ell_repack_stats;
using sharded_blocked_ell_repack_result = csamd_sharded_blocked_ell_repack_result;
using sharded_blocked_ell_repack_workspace = csamd_sharded_blocked_ell_repack_workspace;

inline void init(sharded_blocked_ell_repack_workspace *ws) { csamd_repack_init(ws); }
inline void clear(sharded_blocked_ell_repack_workspace *ws) { csamd_repack_clear(ws); }
inline int setup(sharded_blocked_ell_repack_workspace *ws, int device, hipStream_t stream = (hipStream_t) 0) {
    return csamd_repack_setup(ws, device, (csamd_stream_t) stream);
}
inline int build_repacked_shard_blocked_ell(const device::blocked_ell_view *part_views, unsigned long part_count,
                                            unsigned long shard_id, const sharded_blocked_ell_repack_config *cfg,
                                            const unsigned char *d_keep_rows, const unsigned char *d_keep_cols,
                                            const types::idx_t *d_row_remap, const types::idx_t *d_col_remap,
                                            sharded_blocked_ell_repack_workspace *ws, sharded_blocked_ell_repack_result *out) {
    return csamd_build_repacked_blocked_ell(part_views, (unsigned int) part_count, CSAMD_VALUE_F16, cfg, d_keep_rows, d_keep_cols,
                                            d_row_remap, d_col_remap, shard_id, ws, out);
}
} // namespace repack

// ---------------------------------------------------------------------------
// Bucketed execution partitions (src/io/csh5/execution_builders_part.hh:
// 1206-1275, 1535-1596) over device views; segments stay on the device.
namespace execution {

using bucketed_blocked_ell_partition = csamd_bucketed_blocked_ell;
using bucketed_sliced_ell_partition = csamd_bucketed_sliced_ell;

inline int build_bucketed_execution_partition(bucketed_blocked_ell_partition *out, const csamd_blocked_ell_view *part,
                                              std::uint32_t requested_bucket_count, csamd_convert_workspace *ws,
                                              std::uint64_t *bucketed_bytes_out = nullptr, unsigned int value_type = CSAMD_VALUE_F16) {
    if (!csamd_build_bucketed_blocked_ell(part, value_type, requested_bucket_count, ws, out)) return 0;
    if (bucketed_bytes_out != nullptr) *bucketed_bytes_out = out->bytes;
    return 1;
}
inline int choose_bucket_count_for_part(const csamd_blocked_ell_view *part, csamd_convert_workspace *ws, std::uint32_t *bucket_count_out,
                                        std::uint64_t *bucketed_bytes_out) {
    unsigned long long b = 0ull;
    if (!csamd_choose_bucket_count_blocked_ell(part, ws, bucket_count_out, &b)) return 0;
    *bucketed_bytes_out = b;
    return 1;
}
inline int build_bucketed_sliced_execution_partition(bucketed_sliced_ell_partition *out, const csamd_sliced_ell_view *part,
                                                     std::uint32_t requested_bucket_count, csamd_convert_workspace *ws,
                                                     std::uint64_t *bucketed_bytes_out = nullptr,
                                                     unsigned int value_type = CSAMD_VALUE_F16) {
    if (!csamd_build_bucketed_sliced_ell(part, value_type, requested_bucket_count, ws, out)) return 0;
    if (bucketed_bytes_out != nullptr) *bucketed_bytes_out = out->bytes;
    return 1;
}
inline int choose_bucket_count_for_sliced_part(const csamd_sliced_ell_view *part, csamd_convert_workspace *ws,
                                               std::uint32_t *bucket_count_out, std::uint64_t *bucketed_bytes_out) {
    unsigned long long b = 0ull;
    if (!csamd_choose_bucket_count_sliced_ell(part, ws, bucket_count_out, &b)) return 0;
    *bucketed_bytes_out = b;
    return 1;
}
inline void clear(bucketed_blocked_ell_partition *p) { csamd_bucketed_blocked_ell_clear(p); }
inline void clear(bucketed_sliced_ell_partition *p) { csamd_bucketed_sliced_ell_clear(p); }

} // namespace execution

// ---------------------------------------------------------------------------
// CSPACK01 shard packs of execution partitions (docs/SPEC_CSPACK_V1.md:87-155;
// materialize / load: src/io/csh5/execution_runtime_part.hh:569-796).  The
// reference keeps these behind its .csh5 runtime; here they take a path, a
// shard id and the partitions directly, and loads stage to the device.
namespace pack {

using file = csamd_cspack_file;
using workspace = csamd_cspack_workspace;
constexpr unsigned long long any_shard = CSAMD_CSPACK_ANY_SHARD;

inline void init(workspace *ws) { csamd_cspack_init(ws); }
inline void clear(workspace *ws) { csamd_cspack_clear(ws); }
inline int setup(workspace *ws, int device, hipStream_t stream = nullptr) {
    return csamd_cspack_setup(ws, device, (csamd_stream_t) stream);
}
inline int materialize_blocked_ell_cspack(const char *path, std::uint64_t shard_id,
                                          const execution::bucketed_blocked_ell_partition *parts, std::uint32_t count,
                                          workspace *ws, unsigned int value_type = CSAMD_VALUE_F16) {
    return csamd_cspack_write_blocked_ell(path, shard_id, parts, count, value_type, ws);
}
inline int materialize_sliced_ell_cspack(const char *path, std::uint64_t shard_id,
                                         const execution::bucketed_sliced_ell_partition *parts, std::uint32_t count,
                                         workspace *ws, unsigned int value_type = CSAMD_VALUE_F16) {
    return csamd_cspack_write_sliced_ell(path, shard_id, parts, count, value_type, ws);
}
inline int open(const char *path, std::uint64_t shard_id, file *out) { return csamd_cspack_open(path, shard_id, out); }
inline void close(file *f) { csamd_cspack_close(f); }
inline int load_blocked_pack_partition_from_cspack(const file *f, std::uint64_t local_partition,
                                                   execution::bucketed_blocked_ell_partition *out, workspace *ws,
                                                   unsigned int value_type = CSAMD_VALUE_F16) {
    return csamd_cspack_stage_blocked_ell(f, local_partition, value_type, ws, out);
}
inline int load_sliced_bucketed_partition_from_cspack(const file *f, std::uint64_t local_partition,
                                                      execution::bucketed_sliced_ell_partition *out, workspace *ws,
                                                      unsigned int value_type = CSAMD_VALUE_F16) {
    return csamd_cspack_stage_sliced_ell(f, local_partition, value_type, ws, out);
}
// out: f->partition_count entries.
inline int load_blocked_pack_shard_from_cspack(const file *f, execution::bucketed_blocked_ell_partition *out, workspace *ws,
                                               unsigned int value_type = CSAMD_VALUE_F16) {
    return csamd_cspack_stage_shard_blocked_ell(f, value_type, ws, out);
}
// rebuild_canonical_sliced_partition on the device (d_col null: sizing only).
inline int rebuild_canonical_sliced_partition(const execution::bucketed_sliced_ell_partition *part, unsigned int *d_slot_offsets,
                                              unsigned int *d_col, void *d_val, std::uint64_t capacity, std::uint64_t *slots_out,
                                              csamd_convert_workspace *ws, unsigned int value_type = CSAMD_VALUE_F16) {
    unsigned long long n = 0ull;
    const int ok = csamd_canonical_sliced_ell_from_bucketed_sliced_ell(part, value_type, d_slot_offsets, d_col, d_val, capacity, &n, ws);
    *slots_out = n;
    return ok;
}
// bucketed_partition_to_canonical_coo as a device CSR in canonical row order
// (d_idx null: sizing only); blocked_ell_from_compressed of it is the
// reference's reconstruct_canonical_blocked_ell_part.
inline int canonical_compressed_from_bucketed_partition(const execution::bucketed_blocked_ell_partition *part, unsigned int *d_ptr,
                                                        unsigned int *d_idx, void *d_val, std::uint64_t capacity,
                                                        std::uint64_t *nnz_out, csamd_convert_workspace *ws,
                                                        unsigned int value_type = CSAMD_VALUE_F16, bool ptr_ready = false) {
    unsigned long long n = 0ull;
    const int ok = csamd_canonical_compressed_from_bucketed_blocked_ell(part, value_type, d_ptr, ptr_ready ? 1 : 0, d_idx, d_val,
                                                                        capacity, &n, ws);
    *nnz_out = n;
    return ok;
}

} // namespace pack

// ---------------------------------------------------------------------------
namespace bucket {

using major_nnz_bucket_plan_view = csamd_major_nnz_bucket_plan_view;

static inline types::idx_t clamp_bucket_count(types::dim_t major_dim, types::idx_t requested) {
    return csamd_clamp_bucket_count(major_dim, requested);
}
static inline int major_nnz_bucket_sort_scratch_bytes(types::dim_t major_dim, std::size_t *out) {
    return csamd_major_nnz_bucket_sort_scratch_bytes(major_dim, out);
}
static inline int major_nnz_bucket_scan_scratch_bytes(types::dim_t major_dim, std::size_t *out) {
    return csamd_major_nnz_bucket_scan_scratch_bytes(major_dim, out);
}
static inline int build_major_nnz_bucket_plan_raw(const types::ptr_t *d_major_ptr, types::dim_t major_dim, types::idx_t *d_major_nnz,
                                                  types::idx_t *d_major_nnz_sorted, types::idx_t *d_major_order_in,
                                                  types::idx_t *d_major_order_out, types::idx_t *d_bucket_offsets,
                                                  types::idx_t requested_bucket_count, void *d_sort_tmp, std::size_t sort_tmp_bytes,
                                                  hipStream_t stream) {
    return csamd_build_major_nnz_bucket_plan_raw(d_major_ptr, major_dim, d_major_nnz, d_major_nnz_sorted, d_major_order_in,
                                                 d_major_order_out, d_bucket_offsets, requested_bucket_count, d_sort_tmp,
                                                 sort_tmp_bytes, (csamd_stream_t) stream);
}
static inline int rebuild_compressed_major_order_raw(const types::ptr_t *d_src_major_ptr, const types::idx_t *d_src_minor_idx,
                                                     const real::storage_t *d_src_val, types::dim_t major_dim,
                                                     const types::idx_t *d_major_order, types::ptr_t *d_dst_major_ptr,
                                                     types::idx_t *d_dst_minor_idx, real::storage_t *d_dst_val,
                                                     types::idx_t *d_inverse_major_order, void *d_scan_tmp,
                                                     std::size_t scan_tmp_bytes, hipStream_t stream) {
    return csamd_rebuild_compressed_major_order_raw(d_src_major_ptr, d_src_minor_idx, d_src_val, CSAMD_VALUE_F16, major_dim,
                                                    d_major_order, d_dst_major_ptr, d_dst_minor_idx, d_dst_val,
                                                    d_inverse_major_order, d_scan_tmp, scan_tmp_bytes, (csamd_stream_t) stream);
}

// Column / row orders by rank barycenter (operators/blocked_ell_bipartite_optimize.cuh:361-500).
// The workspace and stream arguments are accepted for signature
// compatibility; the library manages its own buffers and stream.
inline int build_column_order_from_sampled_coo(const sparse::coo *sampled, const std::uint32_t *sample_row_rank, int device,
                                               std::vector<std::uint32_t> *exec_to_canonical_cols,
                                               std::vector<std::uint32_t> *canonical_to_exec_cols, void *workspace = nullptr,
                                               hipStream_t stream = (hipStream_t) 0) {
    (void) workspace;
    (void) stream;
    if (sampled == nullptr || exec_to_canonical_cols == nullptr || canonical_to_exec_cols == nullptr) return 0;
    exec_to_canonical_cols->assign(sampled->cols, 0u);
    canonical_to_exec_cols->assign(sampled->cols, 0u);
    return csamd_build_column_order_from_sampled_coo(sampled, sample_row_rank, device, exec_to_canonical_cols->data(),
                                                     canonical_to_exec_cols->data());
}
inline int build_row_order_from_coo(const sparse::coo *src, const std::uint32_t *canonical_to_exec_cols, int device,
                                    std::vector<std::uint32_t> *exec_to_canonical_rows,
                                    std::vector<std::uint32_t> *canonical_to_exec_rows, void *workspace = nullptr,
                                    hipStream_t stream = (hipStream_t) 0) {
    (void) workspace;
    (void) stream;
    if (src == nullptr || exec_to_canonical_rows == nullptr || canonical_to_exec_rows == nullptr) return 0;
    exec_to_canonical_rows->assign(src->rows, 0u);
    canonical_to_exec_rows->assign(src->rows, 0u);
    return csamd_build_row_order_from_coo(src, canonical_to_exec_cols, device, exec_to_canonical_rows->data(),
                                          canonical_to_exec_rows->data());
}

} // namespace bucket

// ---------------------------------------------------------------------------
// sharded<MatrixT>: partition / shard metadata (runtime/layout/sharded.cuh:30-344,
// runtime/host/sharded_host.cuh:14-548).  Layout identical to csamd_sharded;
// the metadata logic lives in the library (csamd_sharded_*).  Parts are the
// caller's `new MatrixT` objects: clear() / drop_*() destroy them with
// sparse::clear + delete, as the reference does.
template<typename MatrixT>
struct alignas(16) sharded {
    unsigned long rows;
    unsigned long cols;
    unsigned long nnz;
    unsigned long num_partitions;
    unsigned long partition_capacity;
    MatrixT **parts;
    unsigned long *partition_offsets;
    unsigned long *partition_rows;
    unsigned long *partition_nnz;
    unsigned long *partition_aux;
    unsigned long num_shards;
    unsigned long shard_capacity;
    unsigned long *shard_offsets;
    unsigned long *shard_parts;
};

template<typename MatrixT> struct part_format;
template<> struct part_format<sparse::compressed> { static constexpr unsigned int value = CSAMD_FORMAT_COMPRESSED; };
template<> struct part_format<sparse::blocked_ell> { static constexpr unsigned int value = CSAMD_FORMAT_BLOCKED_ELL; };
template<> struct part_format<sparse::sliced_ell> { static constexpr unsigned int value = CSAMD_FORMAT_SLICED_ELL; };

template<typename MatrixT>
inline csamd_sharded *c_sharded(sharded<MatrixT> *m) { return reinterpret_cast<csamd_sharded *>(m); }
template<typename MatrixT>
inline const csamd_sharded *c_sharded(const sharded<MatrixT> *m) { return reinterpret_cast<const csamd_sharded *>(m); }

template<typename MatrixT>
inline void init(sharded<MatrixT> *m) { csamd_sharded_init(c_sharded(m)); }

// sharded_host.cuh:14-19
template<typename MatrixT>
inline void destroy(MatrixT *m) {
    if (m == nullptr) return;
    sparse::clear(m);
    delete m;
}

template<typename MatrixT>
inline void clear(sharded<MatrixT> *m) {
    if (m->parts != nullptr)
        for (unsigned long i = 0; i < m->num_partitions; ++i) destroy(m->parts[i]);
    csamd_sharded_clear(c_sharded(m), 0u);
}
template<typename MatrixT>
inline int reserve_partitions(sharded<MatrixT> *m, unsigned long capacity) { return csamd_sharded_reserve_partitions(c_sharded(m), capacity); }
template<typename MatrixT>
inline int reserve_shards(sharded<MatrixT> *m, unsigned long capacity) { return csamd_sharded_reserve_shards(c_sharded(m), capacity); }
template<typename MatrixT>
inline void rebuild_shard_parts(sharded<MatrixT> *m) { csamd_sharded_rebuild_shard_parts(c_sharded(m)); }
template<typename MatrixT>
inline void rebuild_partition_offsets(sharded<MatrixT> *m) { csamd_sharded_rebuild_partition_offsets(c_sharded(m)); }
template<typename MatrixT>
inline int define_partitions(sharded<MatrixT> *m, unsigned long cols, unsigned long num_partitions, const unsigned long *partition_rows,
                             const unsigned long *partition_nnz, const unsigned long *partition_aux) {
    clear(m);
    return csamd_sharded_define_partitions(c_sharded(m), 0u, cols, num_partitions, partition_rows, partition_nnz, partition_aux);
}
template<typename MatrixT>
inline int set_shards_to_partitions(sharded<MatrixT> *m) { return csamd_sharded_set_shards_to_partitions(c_sharded(m)); }
// takes ownership of `part` (a `new MatrixT`)
template<typename MatrixT>
inline int append_partition(sharded<MatrixT> *m, MatrixT *part) {
    return csamd_sharded_append_partition(c_sharded(m), part_format<MatrixT>::value, part);
}
template<typename MatrixT>
inline int concatenate(sharded<MatrixT> *dst, sharded<MatrixT> *src) { return csamd_sharded_concatenate(c_sharded(dst), c_sharded(src)); }
template<typename MatrixT>
inline int set_equal_shards(sharded<MatrixT> *m, unsigned long count) { return csamd_sharded_set_equal_shards(c_sharded(m), count); }
// *out_offsets is malloc'ed; free() it
template<typename MatrixT>
inline int build_shard_offsets_by_rows(const sharded<MatrixT> *m, unsigned long target_rows_per_shard, unsigned long **out_offsets,
                                       unsigned long *out_count) {
    return csamd_sharded_build_shard_offsets_by_rows(c_sharded(m), target_rows_per_shard, out_offsets, out_count);
}
template<typename MatrixT>
inline int reshard(sharded<MatrixT> *m, unsigned long count, const unsigned long *offsets) { return csamd_sharded_reshard(c_sharded(m), count, offsets); }
template<typename MatrixT>
inline int set_shards_by_nnz(sharded<MatrixT> *m, unsigned long max_nnz) { return csamd_sharded_set_shards_by_nnz(c_sharded(m), max_nnz); }
template<typename MatrixT>
inline int set_shards_by_part_bytes(sharded<MatrixT> *m, std::size_t max_bytes) {
    return csamd_sharded_set_shards_by_part_bytes(c_sharded(m), part_format<MatrixT>::value, max_bytes);
}
template<typename MatrixT>
inline int drop_partition(sharded<MatrixT> *m, unsigned long partId) {
    if (partId >= m->num_partitions) return 0;
    destroy(m->parts[partId]);
    m->parts[partId] = nullptr;
    return 1;
}
template<typename MatrixT>
inline int drop_all_partitions(sharded<MatrixT> *m) {
    for (unsigned long i = 0; i < m->num_partitions; ++i)
        if (!drop_partition(m, i)) return 0;
    return 1;
}

template<typename MatrixT>
inline unsigned long partition_nnz(const MatrixT *m) { return m->nnz; }
template<typename MatrixT>
inline unsigned long partition_aux(const MatrixT *m) { return csamd_partition_aux(part_format<MatrixT>::value, m); }

template<typename MatrixT>
inline unsigned long find_partition(const sharded<MatrixT> *m, unsigned long row) { return csamd_sharded_find_partition(c_sharded(m), row); }
template<typename MatrixT>
inline unsigned long find_shard(const sharded<MatrixT> *m, unsigned long row) { return csamd_sharded_find_shard(c_sharded(m), row); }
template<typename MatrixT>
inline unsigned long first_row_in_partition(const sharded<MatrixT> *m, unsigned long p) {
    return (p >= m->num_partitions || m->partition_offsets == nullptr) ? m->rows : m->partition_offsets[p];
}
template<typename MatrixT>
inline unsigned long last_row_in_partition(const sharded<MatrixT> *m, unsigned long p) {
    return (p >= m->num_partitions || m->partition_offsets == nullptr) ? m->rows : m->partition_offsets[p + 1];
}
template<typename MatrixT>
inline int row_is_partition_boundary(const sharded<MatrixT> *m, unsigned long row) {
    return csamd_sharded_row_is_partition_boundary(c_sharded(m), row);
}
template<typename MatrixT>
inline unsigned long first_partition_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    return csamd_sharded_first_partition_in_shard(c_sharded(m), s);
}
template<typename MatrixT>
inline unsigned long last_partition_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    return csamd_sharded_last_partition_in_shard(c_sharded(m), s);
}
template<typename MatrixT>
inline unsigned long first_row_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    return (s >= m->num_shards || m->shard_offsets == nullptr) ? m->rows : m->shard_offsets[s];
}
template<typename MatrixT>
inline unsigned long last_row_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    return (s >= m->num_shards || m->shard_offsets == nullptr) ? m->rows : m->shard_offsets[s + 1];
}
template<typename MatrixT>
inline unsigned long partition_count_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    const unsigned long b = first_partition_in_shard(m, s), e = last_partition_in_shard(m, s);
    return b >= e ? 0ul : e - b;
}
template<typename MatrixT>
inline unsigned long rows_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    const unsigned long b = first_row_in_shard(m, s), e = last_row_in_shard(m, s);
    return b >= e ? 0ul : e - b;
}
template<typename MatrixT>
inline int partition_loaded(const sharded<MatrixT> *m, unsigned long p) { return p < m->num_partitions && m->parts[p] != nullptr; }
template<typename MatrixT>
inline int shard_loaded(const sharded<MatrixT> *m, unsigned long s) {
    if (s >= m->num_shards) return 0;
    for (unsigned long i = first_partition_in_shard(m, s); i < last_partition_in_shard(m, s); ++i)
        if (!partition_loaded(m, i)) return 0;
    return 1;
}
template<typename MatrixT>
inline unsigned long nnz_in_shard(const sharded<MatrixT> *m, unsigned long s) {
    if (s >= m->num_shards) return 0ul;
    unsigned long t = 0ul;
    for (unsigned long i = first_partition_in_shard(m, s); i < last_partition_in_shard(m, s); ++i) t += m->partition_nnz[i];
    return t;
}
template<typename MatrixT>
inline int drop_shard(sharded<MatrixT> *m, unsigned long s) {
    if (s >= m->num_shards) return 0;
    for (unsigned long i = first_partition_in_shard(m, s); i < last_partition_in_shard(m, s); ++i)
        if (!drop_partition(m, i)) return 0;
    return 1;
}
template<typename MatrixT>
inline std::size_t partition_bytes(const sharded<MatrixT> *m, unsigned long p) {
    return csamd_sharded_partition_bytes(c_sharded(m), part_format<MatrixT>::value, p);
}
template<typename MatrixT>
inline std::size_t shard_bytes(const sharded<MatrixT> *m, unsigned long s) {
    return csamd_sharded_shard_bytes(c_sharded(m), part_format<MatrixT>::value, s);
}
template<typename MatrixT>
inline std::size_t bytes(const sharded<MatrixT> *m) { return csamd_sharded_bytes(c_sharded(m), part_format<MatrixT>::value); }

// ---------------------------------------------------------------------------
// sharded_device<MatrixT>: device residency (runtime/device/sharded_device.cuh:43-50,
// 119-178, 1098-1722).  Layout identical to csamd_sharded_device.  The
// reference's `const shard_storage *files` (its .csh5 fetch) is a CSPACK01
// shard pack here (pack::file, may be null when the host parts are
// materialized); hipError_t 0 = success, 1 = failure (csamd_last_error()).
namespace device {

template<typename MatrixT>
struct sharded_device {
    unsigned long capacity;
    partition_record<MatrixT> *parts;
    unsigned long cache_count;
    unsigned long cache_capacity;
    cached_allocation *cache;
};

template<typename MatrixT>
inline csamd_sharded_device *c_state(sharded_device<MatrixT> *s) { return reinterpret_cast<csamd_sharded_device *>(s); }
inline hipError_t status_(int ok) { return ok ? (hipError_t) 0 : (hipError_t) 1; }

template<typename MatrixT>
inline void init(sharded_device<MatrixT> *s) { csamd_sharded_device_init(c_state(s)); }
template<typename MatrixT>
inline void clear(sharded_device<MatrixT> *s) { (void) csamd_sharded_device_clear(c_state(s)); }
template<typename MatrixT>
inline int reserve(sharded_device<MatrixT> *s, unsigned long capacity) { return csamd_sharded_device_reserve(c_state(s), capacity); }
template<typename MatrixT>
inline std::size_t device_partition_bytes(const sharded<MatrixT> *v, unsigned long p) {
    return csamd_sharded_device_partition_bytes(c_sharded(v), part_format<MatrixT>::value, p);
}
template<typename MatrixT>
inline std::size_t device_shard_bytes(const sharded<MatrixT> *v, unsigned long s) {
    return csamd_sharded_device_shard_bytes(c_sharded(v), part_format<MatrixT>::value, s);
}
template<typename MatrixT>
inline int set_shards_by_device_bytes(sharded<MatrixT> *v, std::size_t max_bytes) {
    return csamd_sharded_set_shards_by_device_bytes(c_sharded(v), part_format<MatrixT>::value, max_bytes);
}
template<typename MatrixT>
inline hipError_t upload_partition(sharded_device<MatrixT> *s, const sharded<MatrixT> *v, unsigned long partId, int deviceId) {
    return status_(csamd_sharded_device_upload_partition(c_state(s), c_sharded(v), part_format<MatrixT>::value, partId, deviceId, nullptr));
}
template<typename MatrixT>
inline hipError_t upload_partition_async(sharded_device<MatrixT> *s, const sharded<MatrixT> *v, unsigned long partId, int deviceId,
                                         hipStream_t stream) {
    return status_(csamd_sharded_device_upload_partition(c_state(s), c_sharded(v), part_format<MatrixT>::value, partId, deviceId,
                                                         (csamd_stream_t) stream));
}
template<typename MatrixT>
inline hipError_t release_partition(sharded_device<MatrixT> *s, unsigned long partId) {
    return status_(csamd_sharded_device_release_partition(c_state(s), partId));
}
template<typename MatrixT>
inline hipError_t upload_shard(sharded_device<MatrixT> *s, const sharded<MatrixT> *v, unsigned long shardId, int deviceId) {
    return status_(csamd_sharded_device_upload_shard(c_state(s), c_sharded(v), part_format<MatrixT>::value, shardId, deviceId, nullptr));
}
template<typename MatrixT>
inline hipError_t upload_shard_async(sharded_device<MatrixT> *s, const sharded<MatrixT> *v, unsigned long shardId, int deviceId,
                                     hipStream_t stream) {
    return status_(csamd_sharded_device_upload_shard(c_state(s), c_sharded(v), part_format<MatrixT>::value, shardId, deviceId,
                                                     (csamd_stream_t) stream));
}
template<typename MatrixT>
inline hipError_t release_shard(sharded_device<MatrixT> *s, const sharded<MatrixT> *v, unsigned long shardId) {
    return status_(csamd_sharded_device_release_shard(c_state(s), c_sharded(v), shardId));
}
template<typename MatrixT>
inline hipError_t stage_partition_async(sharded_device<MatrixT> *s, sharded<MatrixT> *v, const csamd_cspack_file *files,
                                        unsigned long partId, int deviceId, hipStream_t stream, int drop_host_after_upload) {
    if (!csamd_sharded_device_stage_partition(c_state(s), c_sharded(v), part_format<MatrixT>::value, files, partId, deviceId,
                                              (csamd_stream_t) stream, 0))
        return (hipError_t) 1;
    return status_(!drop_host_after_upload || v->parts[partId] == nullptr || drop_partition(v, partId));
}
template<typename MatrixT>
inline hipError_t stage_partition(sharded_device<MatrixT> *s, sharded<MatrixT> *v, const csamd_cspack_file *files, unsigned long partId,
                                  int deviceId, int drop_host_after_upload) {
    return stage_partition_async(s, v, files, partId, deviceId, (hipStream_t) nullptr, drop_host_after_upload);
}
template<typename MatrixT>
inline hipError_t stage_shard_async(sharded_device<MatrixT> *s, sharded<MatrixT> *v, const csamd_cspack_file *files, unsigned long shardId,
                                    int deviceId, hipStream_t stream, int drop_host_after_upload) {
    if (!csamd_sharded_device_stage_shard(c_state(s), c_sharded(v), part_format<MatrixT>::value, files, shardId, deviceId,
                                          (csamd_stream_t) stream, 0))
        return (hipError_t) 1;
    return status_(!drop_host_after_upload || drop_shard(v, shardId));
}
template<typename MatrixT>
inline hipError_t stage_shard(sharded_device<MatrixT> *s, sharded<MatrixT> *v, const csamd_cspack_file *files, unsigned long shardId,
                              int deviceId, int drop_host_after_upload) {
    return stage_shard_async(s, v, files, shardId, deviceId, (hipStream_t) nullptr, drop_host_after_upload);
}
// stages the incoming shard before the outgoing one is released (both resident for a moment)
template<typename MatrixT>
inline hipError_t swap_shard_async(sharded_device<MatrixT> *s, sharded<MatrixT> *v, const csamd_cspack_file *files, unsigned long outShardId,
                                   unsigned long inShardId, int deviceId, hipStream_t stream, int drop_host_after_upload,
                                   int drop_host_after_release) {
    hipError_t e = stage_shard_async(s, v, files, inShardId, deviceId, stream, drop_host_after_upload);
    if (e != (hipError_t) 0) return e;
    e = release_shard(s, v, outShardId);
    if (e != (hipError_t) 0) return e;
    return status_(!drop_host_after_release || drop_shard(v, outShardId));
}
template<typename MatrixT>
inline hipError_t swap_shard(sharded_device<MatrixT> *s, sharded<MatrixT> *v, const csamd_cspack_file *files, unsigned long outShardId,
                             unsigned long inShardId, int deviceId, int drop_host_after_upload, int drop_host_after_release) {
    return swap_shard_async(s, v, files, outShardId, inShardId, deviceId, (hipStream_t) nullptr, drop_host_after_upload,
                            drop_host_after_release);
}

} // namespace device

// ---------------------------------------------------------------------------
// Sharded CSR -> sharded Blocked-ELL repack (convert/blocked_ell_from_compressed.cuh:480-600).
namespace convert {
inline int repack_sharded_compressed_to_blocked_ell(const sharded<sparse::compressed> *src, unsigned int block_size,
                                                    unsigned long target_block_rows, sharded<sparse::blocked_ell> *dst) {
    return csamd_repack_sharded_compressed_to_blocked_ell(c_sharded(src), block_size, target_block_rows, c_sharded(dst));
}
inline int repack_sharded_compressed_to_blocked_ell_auto(const sharded<sparse::compressed> *src, const unsigned int *candidates,
                                                         unsigned int candidate_count, unsigned long target_block_rows,
                                                         sharded<sparse::blocked_ell> *dst, blocked_ell_tune_result *picked = nullptr) {
    return csamd_repack_sharded_compressed_to_blocked_ell_auto(c_sharded(src), candidates, candidate_count, target_block_rows,
                                                               c_sharded(dst), picked);
}
} // namespace convert

// ---------------------------------------------------------------------------
// Workspace-owned major-nnz bucketing (bucket/operators/major_nnz.cuh:11-213,
// sharded_major_nnz.cuh:35-273).  Result views alias the workspace.
namespace bucket {

using compressed_major_bucket_workspace = csamd_compressed_major_bucket_workspace;
using compressed_major_bucket_result = csamd_compressed_major_bucket_result;
using sharded_major_bucket_workspace = csamd_sharded_major_bucket_workspace;
using sharded_major_bucket_result = csamd_sharded_major_bucket_result;

inline void init(compressed_major_bucket_workspace *ws) { csamd_major_bucket_init(ws); }
inline void clear(compressed_major_bucket_workspace *ws) { csamd_major_bucket_clear(ws); }
// as the reference: the workspace creates its own non-blocking stream
inline int setup(compressed_major_bucket_workspace *ws, int device) { return csamd_major_bucket_setup(ws, device, nullptr); }
inline int reserve(compressed_major_bucket_workspace *ws, types::dim_t major_dim, types::nnz_t nnz, types::idx_t bucket_count) {
    return csamd_major_bucket_reserve(ws, major_dim, nnz, bucket_count);
}
inline int build_plan(const device::compressed_view *src, types::idx_t requested_bucket_count, compressed_major_bucket_workspace *ws,
                      major_nnz_bucket_plan_view *out) {
    return csamd_major_bucket_build_plan(src, requested_bucket_count, ws, out);
}
inline int rebuild(const device::compressed_view *src, const major_nnz_bucket_plan_view *plan, compressed_major_bucket_workspace *ws,
                   compressed_major_bucket_result *out) {
    return csamd_major_bucket_rebuild(src, plan, ws, out);
}
inline int build_bucketed_major_view(const device::compressed_view *src, types::idx_t requested_bucket_count,
                                     compressed_major_bucket_workspace *ws, compressed_major_bucket_result *out) {
    return csamd_build_bucketed_major_view(src, requested_bucket_count, ws, out);
}

inline void init(sharded_major_bucket_workspace *ws) { csamd_sharded_major_bucket_init(ws); }
inline void clear(sharded_major_bucket_workspace *ws) { csamd_sharded_major_bucket_clear(ws); }
inline int setup(sharded_major_bucket_workspace *ws, int device) { return csamd_sharded_major_bucket_setup(ws, device, nullptr); }
inline int reserve_parts(sharded_major_bucket_workspace *ws, unsigned long part_count) {
    return csamd_sharded_major_bucket_reserve_parts(ws, part_count);
}
inline int build_bucketed_shard_major_view(const device::sharded_device<sparse::compressed> *state, const sharded<sparse::compressed> *view,
                                           unsigned long shard_id, types::idx_t requested_bucket_count,
                                           sharded_major_bucket_workspace *ws, sharded_major_bucket_result *out) {
    return csamd_build_bucketed_shard_major_view(reinterpret_cast<const csamd_sharded_device *>(state), c_sharded(view), shard_id,
                                                 requested_bucket_count, ws, out);
}

} // namespace bucket

} // namespace cellshard
