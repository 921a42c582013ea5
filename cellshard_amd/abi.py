"""ctypes mirror of include/cellshard_amd/cellshard_amd.h.

Every struct is declared with the C field order; structs the C header marks
aligned(16) get tail padding so sizeof matches (tests/test_abi.py checks each
size and field offset against the C compiler's view).
"""
from __future__ import annotations

import ctypes as C

VALUE_F16 = 3
VALUE_F32 = 4
VALUE_F64 = 5
VALUE_BF16 = 6
VALUE_FP8_E4M3 = 7
VALUE_FP8_E5M2 = 8

COMPRESSED_BY_ROW = 0
COMPRESSED_BY_COL = 1
INVALID_COL = 0xFFFFFFFF
MAX_MASK_GROUPS = 32
MAX_LABEL_GROUPS = 4096

LAYOUT_UNKNOWN = 0
LAYOUT_BLOCKED_ELL = 1
LAYOUT_SLICED_ELL = 2
LAYOUT_COMPRESSED = 3

VALUE_BYTES = {VALUE_F16: 2, VALUE_BF16: 2, VALUE_F32: 4, VALUE_F64: 8, VALUE_FP8_E4M3: 1, VALUE_FP8_E5M2: 1}

c_u32 = C.c_uint32
c_u32_p = C.POINTER(C.c_uint32)
c_u8_p = C.POINTER(C.c_uint8)
c_void_p = C.c_void_p
c_size_t = C.c_size_t


def _aligned16(name: str, fields):
    """Build a ctypes Structure whose size is rounded up to 16 bytes."""
    base = type(name + "_unpadded", (C.Structure,), {"_fields_": fields})
    size = C.sizeof(base)
    pad = (-size) % 16
    if pad:
        fields = list(fields) + [("_tail_pad", C.c_uint8 * pad)]
    return type(name, (C.Structure,), {"_fields_": fields})


compressed_view = _aligned16("compressed_view", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("axis", c_u32),
    ("majorPtr", c_void_p), ("minorIdx", c_void_p), ("val", c_void_p),
])

blocked_ell_view = _aligned16("blocked_ell_view", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("block_size", c_u32), ("ell_cols", c_u32),
    ("blockColIdx", c_void_p), ("val", c_void_p),
])

sliced_ell_view = _aligned16("sliced_ell_view", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("slice_count", c_u32), ("slice_rows", c_u32),
    ("slice_row_offsets", c_void_p), ("slice_widths", c_void_p), ("slice_slot_offsets", c_void_p),
    ("col_idx", c_void_p), ("val", c_void_p),
])

group_mask_config_view = _aligned16("group_mask_config_view", [
    ("group_count", c_u32), ("group_names", c_void_p), ("feature_group_masks", c_void_p),
])

row_feature_mask_view = _aligned16("row_feature_mask_view", [
    ("row_keep", c_void_p), ("feature_keep", c_void_p), ("row_remap", c_void_p), ("feature_remap", c_void_p),
])

sparse_group_filter_params = _aligned16("sparse_group_filter_params", [
    ("min_total", C.c_float), ("min_detected_features", c_u32), ("max_group_fraction", C.c_float),
    ("fraction_group_index", c_u32),
])

sparse_group_reduce_result = _aligned16("sparse_group_reduce_result", [
    ("rows", c_u32), ("group_count", c_u32),
    ("row_totals", c_void_p), ("detected_features", c_void_p), ("max_values", c_void_p), ("row_keep", c_void_p),
    ("group_counts", c_void_p), ("group_percentages", c_void_p),
])

sparse_group_reduce_workspace = _aligned16("sparse_group_reduce_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("owns_stream", C.c_int),
    ("rows_capacity", c_u32), ("cols_capacity", c_u32), ("group_capacity", c_u32), ("values_capacity", c_u32),
    ("row_block", c_void_p), ("feature_block", c_void_p), ("scratch_block", c_void_p), ("scratch_bytes", c_size_t),
    ("row_totals", c_void_p), ("detected_features", c_void_p), ("max_values", c_void_p), ("row_keep", c_void_p),
    ("group_counts", c_void_p), ("group_percentages", c_void_p), ("feature_group_masks", c_void_p),
])

sparse_group_reduce_fleet_workspace = _aligned16("sparse_group_reduce_fleet_workspace", [
    ("slot_count", c_u32), ("device_ids", c_void_p), ("slots", c_void_p),
    ("devices", C.POINTER(sparse_group_reduce_workspace)), ("results", C.POINTER(sparse_group_reduce_result)),
])

sparse_group_reduce_fleet_result = _aligned16("sparse_group_reduce_fleet_result", [
    ("slot_count", c_u32), ("leader_index", c_u32), ("slot_results", C.POINTER(sparse_group_reduce_result)),
])

sparse_group_reduce_fleet_config = _aligned16("sparse_group_reduce_fleet_config", [
    ("device_ids", c_void_p), ("device_count", c_u32), ("enable_peer_access", c_u32), ("stream_flags", c_u32),
    ("ranked_nccl", c_void_p),
])

masked_sparse_reoptimize_config = _aligned16("masked_sparse_reoptimize_config", [
    ("output_rows", c_u32), ("output_cols", c_u32), ("requested_bucket_count", c_u32), ("output_block_size", c_u32),
    ("requested_layout", c_u32),
])

masked_sparse_reoptimize_metadata = _aligned16("masked_sparse_reoptimize_metadata", [
    ("kept_rows", c_u32), ("kept_features", c_u32), ("live_nnz", c_u32),
    ("bytes_before", C.c_uint64), ("bytes_after", C.c_uint64), ("output_layout", c_u32),
])

masked_sparse_reoptimize_result = _aligned16("masked_sparse_reoptimize_result", [
    ("compressed", compressed_view), ("blocked_ell", blocked_ell_view), ("metadata", masked_sparse_reoptimize_metadata),
])

masked_sparse_reoptimize_workspace = _aligned16("masked_sparse_reoptimize_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("owns_stream", C.c_int), ("private_workspace", c_void_p),
])

auto_reoptimize_prediction_input = _aligned16("auto_reoptimize_prediction_input", [
    ("density_change", C.c_float), ("row_drop_ratio", C.c_float), ("feature_drop_ratio", C.c_float),
    ("layout_padding_change", C.c_float), ("estimated_rebuild_cost_ms", C.c_float), ("expected_reuse_count", c_u32),
])

auto_reoptimize_prediction = _aligned16("auto_reoptimize_prediction", [
    ("should_reoptimize", c_u32), ("estimated_net_gain_ms", C.c_float),
])

label_reduce_query = _aligned16("label_reduce_query", [
    ("group_count", c_u32), ("value_type", c_u32), ("row_labels", c_void_p), ("row_keep", c_void_p),
    ("feature_keep", c_void_p), ("feature_scale", c_void_p), ("compute_moments", c_u32), ("reserved", c_u32),
])

label_reduce_result = _aligned16("label_reduce_result", [
    ("group_count", c_u32), ("kept_features", c_u32), ("kept_rows", c_u32), ("reserved", c_u32),
    ("sum", c_void_p), ("sumsq", c_void_p), ("nnz", c_void_p), ("group_rows", c_void_p),
    ("mean", c_void_p), ("var", c_void_p), ("kept_feature_ids", c_void_p),
])

label_reduce_workspace = _aligned16("label_reduce_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("owns_stream", C.c_int), ("private_workspace", c_void_p),
])

convert_workspace = _aligned16("convert_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("owns_stream", C.c_int), ("scratch", c_void_p), ("scratch_bytes", c_size_t),
])

blocked_ell_tune_result = _aligned16("blocked_ell_tune_result", [
    ("block_size", c_u32), ("fill_ratio", C.c_double), ("padded_bytes", c_size_t),
])

sliced_ell_tune_result = _aligned16("sliced_ell_tune_result", [
    ("slice_rows", c_u32), ("slice_count", c_u32), ("padded_slots", c_size_t), ("padded_bytes", c_size_t),
    ("storage_ratio", C.c_double), ("slice_penalty", C.c_double), ("score", C.c_double),
])

major_nnz_bucket_plan_view = _aligned16("major_nnz_bucket_plan_view", [
    ("major_dim", c_u32), ("bucket_count", c_u32), ("major_order", c_void_p), ("major_nnz_sorted", c_void_p),
    ("bucket_offsets", c_void_p), ("inverse_major_order", c_void_p),
])

filtered_blocked_ell_stats = _aligned16("filtered_blocked_ell_stats", [
    ("kept_rows", c_u32), ("output_rows", c_u32), ("output_cols", c_u32), ("live_nnz", c_u32),
    ("kept_row_slots", C.c_uint64), ("dead_slots", C.c_uint64), ("live_fill_ratio", C.c_double),
    ("dead_value_bytes", c_size_t),
])

compressed_major_bucket_result = _aligned16("compressed_major_bucket_result", [
    ("rebuilt", compressed_view), ("plan", major_nnz_bucket_plan_view),
])

filtered_blocked_ell_result = _aligned16("filtered_blocked_ell_result", [
    ("filtered", compressed_view), ("bucketed", compressed_major_bucket_result), ("stats", filtered_blocked_ell_stats),
    ("has_bucketed", C.c_int),
])

filtered_blocked_ell_workspace = _aligned16("filtered_blocked_ell_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("owns_stream", C.c_int), ("private_workspace", c_void_p),
])

sharded_blocked_ell_repack_config = _aligned16("sharded_blocked_ell_repack_config", [
    ("output_rows", c_u32), ("output_cols", c_u32), ("block_size", c_u32), ("requested_bucket_count", c_u32),
])

sharded_blocked_ell_repack_stats = _aligned16("sharded_blocked_ell_repack_stats", [
    ("kept_rows", c_u32), ("output_rows", c_u32), ("output_cols", c_u32), ("live_nnz", c_u32),
    ("kept_row_slots", C.c_uint64), ("dead_slots", C.c_uint64), ("live_fill_ratio", C.c_double),
    ("dead_value_bytes", c_size_t), ("ell_width", c_u32),
])

sharded_blocked_ell_repack_result = _aligned16("sharded_blocked_ell_repack_result", [
    ("filtered", compressed_view), ("bucketed", compressed_major_bucket_result), ("rebuilt", blocked_ell_view),
    ("bucket_plan", major_nnz_bucket_plan_view), ("stats", sharded_blocked_ell_repack_stats), ("shard_id", C.c_ulong),
])

sharded_blocked_ell_repack_workspace = _aligned16("sharded_blocked_ell_repack_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("owns_stream", C.c_int), ("private_workspace", c_void_p),
])

synth_params = _aligned16("synth_params", [
    ("rows", c_u32), ("cols", c_u32), ("value_type", c_u32), ("seed", c_u32),
    ("row_nnz_median", C.c_float), ("row_nnz_sigma", C.c_float), ("nb_alpha", C.c_float), ("reserved", c_u32),
    ("d_col_inv_weight", c_void_p), ("d_col_mean", c_void_p),
])

host_compressed = _aligned16("host_compressed", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("axis", c_u32), ("flags", c_u32),
    ("storage", c_void_p), ("majorPtr", c_void_p), ("minorIdx", c_void_p), ("val", c_void_p),
])

bucketed_blocked_ell = _aligned16("bucketed_blocked_ell", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("segment_count", c_u32),
    ("segments", C.POINTER(blocked_ell_view)), ("segment_row_offsets", C.POINTER(c_u32)),
    ("exec_to_canonical_rows", c_void_p), ("canonical_to_exec_rows", c_void_p),
    ("exec_to_canonical_cols", c_void_p), ("canonical_to_exec_cols", c_void_p),
    ("bytes", C.c_uint64), ("storage", c_void_p),
])

bucketed_sliced_ell = _aligned16("bucketed_sliced_ell", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("segment_count", c_u32),
    ("segments", C.POINTER(sliced_ell_view)), ("segment_row_offsets", C.POINTER(c_u32)),
    ("exec_to_canonical_rows", c_void_p), ("canonical_to_exec_rows", c_void_p),
    ("canonical_slice_count", c_u32), ("reserved", c_u32),
    ("canonical_slice_row_offsets", c_void_p), ("canonical_slice_widths", c_void_p),
    ("bytes", C.c_uint64), ("storage", c_void_p),
])

host_coo = _aligned16("host_coo", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("storage", c_void_p), ("rowIdx", c_void_p), ("colIdx", c_void_p),
    ("val", c_void_p),
])

host_blocked_ell = _aligned16("host_blocked_ell", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("block_size", c_u32), ("ell_cols", c_u32), ("flags", c_u32),
    ("storage", c_void_p), ("blockColIdx", c_void_p), ("val", c_void_p),
])

host_sliced_ell = _aligned16("host_sliced_ell", [
    ("rows", c_u32), ("cols", c_u32), ("nnz", c_u32), ("slice_count", c_u32), ("flags", c_u32),
    ("storage", c_void_p), ("slice_row_offsets", c_void_p), ("slice_widths", c_void_p), ("col_idx", c_void_p),
    ("val", c_void_p),
])

partition_record = _aligned16("partition_record", [
    ("allocation", c_void_p), ("allocation_bytes", c_size_t), ("storage", c_void_p), ("view", c_void_p),
    ("a0", c_void_p), ("a1", c_void_p), ("a2", c_void_p), ("a3", c_void_p),
    ("group_begin", C.c_ulong), ("group_end", C.c_ulong), ("device_id", C.c_int),
])

cached_allocation = _aligned16("cached_allocation", [("ptr", c_void_p), ("bytes", c_size_t), ("device_id", C.c_int)])

FORMAT_COMPRESSED = 1
FORMAT_BLOCKED_ELL = 2
FORMAT_SLICED_ELL = 3

c_ulong = C.c_ulong
c_ulong_p = C.POINTER(C.c_ulong)

sharded = _aligned16("sharded", [
    ("rows", c_ulong), ("cols", c_ulong), ("nnz", c_ulong), ("num_partitions", c_ulong), ("partition_capacity", c_ulong),
    ("parts", C.POINTER(c_void_p)), ("partition_offsets", c_ulong_p), ("partition_rows", c_ulong_p),
    ("partition_nnz", c_ulong_p), ("partition_aux", c_ulong_p), ("num_shards", c_ulong), ("shard_capacity", c_ulong),
    ("shard_offsets", c_ulong_p), ("shard_parts", c_ulong_p),
])

sharded_device = type("sharded_device", (C.Structure,), {"_fields_": [
    ("capacity", c_ulong), ("parts", C.POINTER(partition_record)), ("cache_count", c_ulong), ("cache_capacity", c_ulong),
    ("cache", C.POINTER(cached_allocation)),
]})

compressed_major_bucket_workspace = _aligned16("compressed_major_bucket_workspace", [
    ("device", C.c_int), ("stream", c_void_p), ("major_capacity", c_u32), ("nnz_capacity", c_u32), ("bucket_capacity", c_u32),
    ("sort_capacity", c_size_t), ("scan_capacity", c_size_t), ("d_major_block", c_void_p), ("d_rebuilt_block", c_void_p),
    ("d_major_nnz", c_void_p), ("d_major_nnz_sorted", c_void_p), ("d_major_order_in", c_void_p), ("d_major_order_out", c_void_p),
    ("d_inverse_major_order", c_void_p), ("d_bucket_offsets", c_void_p), ("d_rebuilt_major_ptr", c_void_p),
    ("d_rebuilt_minor_idx", c_void_p), ("d_rebuilt_val", c_void_p), ("d_sort_tmp", c_void_p), ("d_scan_tmp", c_void_p),
    ("owns_stream", C.c_int), ("value_bytes_capacity", c_u32),
])

sharded_major_bucket_workspace = _aligned16("sharded_major_bucket_workspace", [
    ("base", compressed_major_bucket_workspace), ("part_capacity", c_ulong), ("h_part_block", c_void_p), ("d_part_block", c_void_p),
    ("h_part_major_ptr", c_void_p), ("h_part_minor_idx", c_void_p), ("h_part_val", c_void_p), ("h_part_row_offsets", c_void_p),
    ("d_part_major_ptr", c_void_p), ("d_part_minor_idx", c_void_p), ("d_part_val", c_void_p), ("d_part_row_offsets", c_void_p),
    ("tables_event", c_void_p),
])

sharded_major_bucket_result = _aligned16("sharded_major_bucket_result", [
    ("rebuilt", compressed_view), ("plan", major_nnz_bucket_plan_view), ("shard_id", c_ulong),
])

allocation_cache = _aligned16("allocation_cache", [
    ("cache_count", C.c_ulong), ("cache_capacity", C.c_ulong), ("cache", C.POINTER(cached_allocation)),
])

serve_query = _aligned16("serve_query", [
    ("shard", c_u32), ("op", c_u32), ("group_count", c_u32), ("feature_count", c_u32),
    ("features", c_void_p), ("row_labels", c_void_p), ("row_keep", c_void_p),
    ("label_seed", c_u32), ("mask_seed", c_u32), ("keep_thr24", c_u32), ("reserved", c_u32),
])
serve_result = _aligned16("serve_result", [
    ("status", C.c_int), ("op", c_u32), ("group_count", c_u32), ("feature_count", c_u32),
    ("out", c_void_p), ("latency_us", C.c_double),
])
serve_config = _aligned16("serve_config", [
    ("device", C.c_int), ("max_slots", c_u32), ("max_batch_queries", c_u32), ("max_batches_in_flight", c_u32),
    ("slot_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
])
serve_workspace = _aligned16("serve_workspace", [("device", C.c_int), ("state", c_void_p)])
SERVE_SUM, SERVE_SUMSQ, SERVE_NNZ = 0, 1, 2

label_reduce_partition = _aligned16("label_reduce_partition", [
    ("layout", c_u32), ("reserved", c_u32 * 3), ("compressed", compressed_view), ("sliced_ell", sliced_ell_view),
])

pipeline_partition = _aligned16("pipeline_partition", [
    ("layout", c_u32), ("slice_rows", c_u32), ("block_size", c_u32), ("ell_width", c_u32),
    ("staged_from", C.POINTER(host_compressed)), ("d_nnz", c_void_p), ("slots_capacity", C.c_uint64),
    ("_align16", C.c_uint64),  # the C compiler places the 16-byte-aligned view at offset 48
    ("compressed", compressed_view), ("sliced_ell", sliced_ell_view), ("blocked_ell", blocked_ell_view),
])
label_reduce_graph = _aligned16("label_reduce_graph", [("device", C.c_int), ("state", c_void_p)])
partition_tune_result = _aligned16("partition_tune_result", [
    ("layout", c_u32), ("bell_block_size", c_u32), ("bell_ell_width", c_u32), ("sell_slice_rows", c_u32),
    ("sell_slice_count", c_u32), ("reserved", c_u32), ("sell_padded_slots", C.c_uint64), ("csr_bytes", C.c_uint64),
    ("bell_bytes", C.c_uint64), ("sell_bytes", C.c_uint64), ("bell_fill_ratio", C.c_double), ("sell_score", C.c_double),
    ("row_length_log2_hist", c_u32 * 32),
])

CSPACK_ANY_SHARD = 0xFFFFFFFFFFFFFFFF

cspack_file = _aligned16("cspack_file", [
    ("shard_id", C.c_uint64), ("partition_count", C.c_uint64), ("file_bytes", C.c_uint64),
    ("partition_offsets", C.POINTER(C.c_uint64)), ("partition_bytes", C.POINTER(C.c_uint64)), ("impl", c_void_p),
])

cspack_workspace = _aligned16("cspack_workspace", [
    ("device", C.c_int), ("reader_threads", c_u32), ("stream", c_void_p), ("owns_stream", C.c_int),
    ("chunk_bytes", c_u32), ("private_workspace", c_void_p),
])

# Structs whose layout tests/test_abi.py compares against the C header.
ALL_STRUCTS = {
    "csamd_compressed_view": compressed_view,
    "csamd_blocked_ell_view": blocked_ell_view,
    "csamd_sliced_ell_view": sliced_ell_view,
    "csamd_group_mask_config_view": group_mask_config_view,
    "csamd_row_feature_mask_view": row_feature_mask_view,
    "csamd_sparse_group_filter_params": sparse_group_filter_params,
    "csamd_sparse_group_reduce_result": sparse_group_reduce_result,
    "csamd_sparse_group_reduce_workspace": sparse_group_reduce_workspace,
    "csamd_sparse_group_reduce_fleet_workspace": sparse_group_reduce_fleet_workspace,
    "csamd_sparse_group_reduce_fleet_result": sparse_group_reduce_fleet_result,
    "csamd_sparse_group_reduce_fleet_config": sparse_group_reduce_fleet_config,
    "csamd_masked_sparse_reoptimize_config": masked_sparse_reoptimize_config,
    "csamd_masked_sparse_reoptimize_metadata": masked_sparse_reoptimize_metadata,
    "csamd_masked_sparse_reoptimize_result": masked_sparse_reoptimize_result,
    "csamd_masked_sparse_reoptimize_workspace": masked_sparse_reoptimize_workspace,
    "csamd_auto_reoptimize_prediction_input": auto_reoptimize_prediction_input,
    "csamd_auto_reoptimize_prediction": auto_reoptimize_prediction,
    "csamd_label_reduce_query": label_reduce_query,
    "csamd_label_reduce_result": label_reduce_result,
    "csamd_label_reduce_workspace": label_reduce_workspace,
    "csamd_label_reduce_partition": label_reduce_partition,
    "csamd_label_reduce_graph": label_reduce_graph,
    "csamd_partition_tune_result": partition_tune_result,
    "csamd_host_coo": host_coo,
    "csamd_bucketed_blocked_ell": bucketed_blocked_ell,
    "csamd_bucketed_sliced_ell": bucketed_sliced_ell,
    "csamd_cspack_file": cspack_file,
    "csamd_cspack_workspace": cspack_workspace,
    "csamd_convert_workspace": convert_workspace,
    "csamd_blocked_ell_tune_result": blocked_ell_tune_result,
    "csamd_sliced_ell_tune_result": sliced_ell_tune_result,
    "csamd_major_nnz_bucket_plan_view": major_nnz_bucket_plan_view,
    "csamd_filtered_blocked_ell_stats": filtered_blocked_ell_stats,
    "csamd_compressed_major_bucket_result": compressed_major_bucket_result,
    "csamd_compressed_major_bucket_workspace": compressed_major_bucket_workspace,
    "csamd_pipeline_partition": pipeline_partition,
    "csamd_sharded_major_bucket_workspace": sharded_major_bucket_workspace,
    "csamd_sharded_major_bucket_result": sharded_major_bucket_result,
    "csamd_filtered_blocked_ell_result": filtered_blocked_ell_result,
    "csamd_filtered_blocked_ell_workspace": filtered_blocked_ell_workspace,
    "csamd_sharded_blocked_ell_repack_config": sharded_blocked_ell_repack_config,
    "csamd_sharded_blocked_ell_repack_stats": sharded_blocked_ell_repack_stats,
    "csamd_sharded_blocked_ell_repack_result": sharded_blocked_ell_repack_result,
    "csamd_sharded_blocked_ell_repack_workspace": sharded_blocked_ell_repack_workspace,
    "csamd_synth_params": synth_params,
    "csamd_host_compressed": host_compressed,
    "csamd_host_blocked_ell": host_blocked_ell,
    "csamd_host_sliced_ell": host_sliced_ell,
    "csamd_partition_record": partition_record,
    "csamd_cached_allocation": cached_allocation,
    "csamd_allocation_cache": allocation_cache,
    "csamd_sharded": sharded,
    "csamd_sharded_device": sharded_device,
    "csamd_serve_query": serve_query,
    "csamd_serve_result": serve_result,
    "csamd_serve_config": serve_config,
    "csamd_serve_workspace": serve_workspace,
}
