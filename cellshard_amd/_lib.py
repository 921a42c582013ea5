"""Loader for libcellshard_amd.so and the C-ABI prototypes.

The library is required: there is no CPU fallback for any entry point.  A
missing or unloadable library raises LibraryMissing at first use.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from . import abi as A

LIB_PATH = Path(__file__).resolve().parent / "libcellshard_amd.so"

P = C.POINTER
vp = C.c_void_p
u32 = C.c_uint32
u64 = C.c_uint64
i32 = C.c_int
f32 = C.c_float
sz = C.c_size_t

# name -> (restype, [argtypes])
PROTOTYPES = {
    "csamd_last_error": (C.c_char_p, []),
    "csamd_abi_version": (i32, []),
    "csamd_device_count": (i32, []),
    "csamd_set_device": (i32, [i32]),
    # mask_groups
    "csamd_group_reduce_init": (None, [P(A.sparse_group_reduce_workspace)]),
    "csamd_group_reduce_clear": (None, [P(A.sparse_group_reduce_workspace)]),
    "csamd_group_reduce_setup": (i32, [P(A.sparse_group_reduce_workspace), i32, vp]),
    "csamd_group_reduce_reserve": (i32, [P(A.sparse_group_reduce_workspace), u32, u32, u32, u32]),
    "csamd_upload_feature_group_masks": (i32, [P(A.sparse_group_reduce_workspace), u32, vp]),
    "csamd_compute_sparse_group_reduce_blocked_ell": (i32, [P(A.blocked_ell_view), P(A.sparse_group_reduce_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result)]),
    "csamd_compute_sparse_group_reduce_sliced_ell": (i32, [P(A.sliced_ell_view), P(A.sparse_group_reduce_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result)]),
    "csamd_compute_sparse_group_reduce_compressed": (i32, [P(A.compressed_view), P(A.sparse_group_reduce_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result)]),
    "csamd_compute_sparse_group_reduce_compressed_typed": (i32, [P(A.compressed_view), u32, P(A.sparse_group_reduce_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result)]),
    "csamd_compute_sparse_group_reduce_compressed_noalloc": (i32, [P(A.compressed_view), u32, P(A.group_mask_config_view),
        P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result), vp]),
    "csamd_compute_sparse_group_reduce_blocked_ell_typed": (i32, [P(A.blocked_ell_view), u32, P(A.sparse_group_reduce_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result)]),
    "csamd_compute_sparse_group_reduce_sliced_ell_typed": (i32, [P(A.sliced_ell_view), u32, P(A.sparse_group_reduce_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_result)]),
    "csamd_label_reduce_pipeline_capture": (i32, [P(A.pipeline_partition), u32, P(A.label_reduce_query), i32,
        P(A.label_reduce_graph)]),
    "csamd_major_bucket_init": (None, [P(A.compressed_major_bucket_workspace)]),
    "csamd_major_bucket_clear": (None, [P(A.compressed_major_bucket_workspace)]),
    "csamd_major_bucket_setup": (i32, [P(A.compressed_major_bucket_workspace), i32, vp]),
    "csamd_major_bucket_reserve": (i32, [P(A.compressed_major_bucket_workspace), u32, u32, u32]),
    "csamd_major_bucket_reserve_typed": (i32, [P(A.compressed_major_bucket_workspace), u32, u32, u32, u32]),
    "csamd_major_bucket_build_plan": (i32, [P(A.compressed_view), u32, P(A.compressed_major_bucket_workspace),
        P(A.major_nnz_bucket_plan_view)]),
    "csamd_major_bucket_rebuild": (i32, [P(A.compressed_view), P(A.major_nnz_bucket_plan_view), P(A.compressed_major_bucket_workspace),
        P(A.compressed_major_bucket_result)]),
    "csamd_major_bucket_rebuild_typed": (i32, [P(A.compressed_view), u32, P(A.major_nnz_bucket_plan_view),
        P(A.compressed_major_bucket_workspace), P(A.compressed_major_bucket_result)]),
    "csamd_build_bucketed_major_view": (i32, [P(A.compressed_view), u32, P(A.compressed_major_bucket_workspace),
        P(A.compressed_major_bucket_result)]),
    "csamd_build_bucketed_major_view_typed": (i32, [P(A.compressed_view), u32, u32, P(A.compressed_major_bucket_workspace),
        P(A.compressed_major_bucket_result)]),
    "csamd_repack_sharded_compressed_to_blocked_ell": (i32, [P(A.sharded), u32, C.c_ulong, P(A.sharded)]),
    "csamd_repack_sharded_compressed_to_blocked_ell_auto": (i32, [P(A.sharded), P(u32), u32, C.c_ulong, P(A.sharded),
        P(A.blocked_ell_tune_result)]),
    "csamd_sharded_major_bucket_init": (None, [P(A.sharded_major_bucket_workspace)]),
    "csamd_sharded_major_bucket_clear": (None, [P(A.sharded_major_bucket_workspace)]),
    "csamd_sharded_major_bucket_setup": (i32, [P(A.sharded_major_bucket_workspace), i32, vp]),
    "csamd_sharded_major_bucket_reserve_parts": (i32, [P(A.sharded_major_bucket_workspace), C.c_ulong]),
    "csamd_build_bucketed_shard_major_view": (i32, [P(A.sharded_device), P(A.sharded), C.c_ulong, u32,
        P(A.sharded_major_bucket_workspace), P(A.sharded_major_bucket_result)]),
    "csamd_build_bucketed_shard_major_view_typed": (i32, [P(A.sharded_device), P(A.sharded), C.c_ulong, u32, u32,
        P(A.sharded_major_bucket_workspace), P(A.sharded_major_bucket_result)]),
    "csamd_fleet_init": (None, [P(A.sparse_group_reduce_fleet_workspace)]),
    "csamd_fleet_clear": (None, [P(A.sparse_group_reduce_fleet_workspace)]),
    "csamd_setup_fleet": (i32, [P(A.sparse_group_reduce_fleet_workspace), P(A.sparse_group_reduce_fleet_config)]),
    "csamd_compute_sparse_group_reduce_fleet_blocked_ell": (i32, [P(A.blocked_ell_view), P(A.sparse_group_reduce_fleet_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_fleet_result)]),
    "csamd_compute_sparse_group_reduce_fleet_sliced_ell": (i32, [P(A.sliced_ell_view), P(A.sparse_group_reduce_fleet_workspace),
        P(A.group_mask_config_view), P(A.row_feature_mask_view), P(A.sparse_group_filter_params), P(A.sparse_group_reduce_fleet_result)]),
    "csamd_predict_masked_sparse_reoptimization": (i32, [P(A.auto_reoptimize_prediction_input), P(A.auto_reoptimize_prediction)]),
    "csamd_reoptimize_init": (None, [P(A.masked_sparse_reoptimize_workspace)]),
    "csamd_reoptimize_clear": (None, [P(A.masked_sparse_reoptimize_workspace)]),
    "csamd_reoptimize_setup": (i32, [P(A.masked_sparse_reoptimize_workspace), i32, vp]),
    "csamd_manual_reoptimize_masked_sparse": (i32, [P(A.blocked_ell_view), P(A.row_feature_mask_view),
        P(A.masked_sparse_reoptimize_config), P(A.masked_sparse_reoptimize_workspace), P(A.masked_sparse_reoptimize_result)]),
    "csamd_filtered_blocked_ell_init": (None, [P(A.filtered_blocked_ell_workspace)]),
    "csamd_filtered_blocked_ell_clear": (None, [P(A.filtered_blocked_ell_workspace)]),
    "csamd_filtered_blocked_ell_setup": (i32, [P(A.filtered_blocked_ell_workspace), i32, vp]),
    "csamd_filter_blocked_ell_to_compressed": (i32, [P(A.blocked_ell_view), u32, u32, u32, vp, vp, vp, vp,
        P(A.filtered_blocked_ell_workspace), P(A.filtered_blocked_ell_result)]),
    "csamd_build_bucketed_filtered_blocked_ell_major_view": (i32, [P(A.blocked_ell_view), u32, u32, u32, vp, vp, vp, vp, u32,
        P(A.filtered_blocked_ell_workspace), P(A.filtered_blocked_ell_result)]),
    "csamd_build_shard_column_maps": (i32, [P(A.blocked_ell_view), u32, u32, u32, vp, vp, P(A.convert_workspace)]),
    "csamd_build_column_order_from_sampled_coo": (i32, [P(A.host_coo), vp, i32, vp, vp]),
    "csamd_build_row_order_from_coo": (i32, [P(A.host_coo), vp, i32, vp, vp]),
    "csamd_repack_init": (None, [P(A.sharded_blocked_ell_repack_workspace)]),
    "csamd_repack_clear": (None, [P(A.sharded_blocked_ell_repack_workspace)]),
    "csamd_repack_setup": (i32, [P(A.sharded_blocked_ell_repack_workspace), i32, vp]),
    "csamd_build_repacked_blocked_ell": (i32, [P(A.blocked_ell_view), u32, u32, P(A.sharded_blocked_ell_repack_config), vp, vp, vp,
        vp, C.c_ulong, P(A.sharded_blocked_ell_repack_workspace), P(A.sharded_blocked_ell_repack_result)]),
    # label reduce
    "csamd_label_reduce_init": (None, [P(A.label_reduce_workspace)]),
    "csamd_label_reduce_clear": (None, [P(A.label_reduce_workspace)]),
    "csamd_label_reduce_setup": (i32, [P(A.label_reduce_workspace), i32, vp]),
    "csamd_label_reduce_compressed": (i32, [P(A.compressed_view), P(A.label_reduce_query), P(A.label_reduce_workspace), P(A.label_reduce_result)]),
    "csamd_label_reduce_sliced_ell": (i32, [P(A.sliced_ell_view), P(A.label_reduce_query), P(A.label_reduce_workspace), P(A.label_reduce_result)]),
    "csamd_label_reduce_blocked_ell": (i32, [P(A.blocked_ell_view), P(A.label_reduce_query), P(A.label_reduce_workspace), P(A.label_reduce_result)]),
    "csamd_label_reduce_last_kernel_ms": (f32, [P(A.label_reduce_workspace)]),
    "csamd_label_reduce_batched": (i32, [P(A.label_reduce_partition), u32, P(A.label_reduce_query), P(A.label_reduce_workspace), P(A.label_reduce_result)]),
    "csamd_label_reduce_graph_capture": (i32, [P(A.label_reduce_partition), u32, P(A.label_reduce_query), i32, P(A.label_reduce_graph)]),
    "csamd_label_reduce_graph_launch": (i32, [P(A.label_reduce_graph), vp]),
    "csamd_label_reduce_graph_result": (i32, [P(A.label_reduce_graph), P(A.label_reduce_result)]),
    "csamd_label_reduce_graph_destroy": (None, [P(A.label_reduce_graph)]),
    "csamd_tune_partitions": (i32, [P(A.compressed_view), u32, u32, P(A.convert_workspace), P(A.partition_tune_result)]),
    "csamd_choose_blocked_ell_block_size_from_coo": (i32, [P(A.host_coo), u32, vp, vp, u32, i32, P(A.blocked_ell_tune_result)]),
    "csamd_blocked_ell_from_coo": (i32, [P(A.host_coo), u32, vp, u32, P(A.host_blocked_ell), i32]),
    "csamd_blocked_ell_from_coo_auto": (i32, [P(A.host_coo), u32, vp, vp, u32, P(A.host_blocked_ell), i32, P(A.blocked_ell_tune_result)]),
    "csamd_compressed_from_coo_scratch_bytes": (sz, [u32]),
    "csamd_build_bucketed_blocked_ell": (i32, [P(A.blocked_ell_view), u32, u32, P(A.convert_workspace), P(A.bucketed_blocked_ell)]),
    "csamd_choose_bucket_count_blocked_ell": (i32, [P(A.blocked_ell_view), P(A.convert_workspace), P(u32), P(C.c_uint64)]),
    "csamd_bucketed_blocked_ell_clear": (None, [P(A.bucketed_blocked_ell)]),
    "csamd_build_bucketed_sliced_ell": (i32, [P(A.sliced_ell_view), u32, u32, P(A.convert_workspace), P(A.bucketed_sliced_ell)]),
    "csamd_choose_bucket_count_sliced_ell": (i32, [P(A.sliced_ell_view), P(A.convert_workspace), P(u32), P(C.c_uint64)]),
    "csamd_bucketed_sliced_ell_clear": (None, [P(A.bucketed_sliced_ell)]),
    # CSPACK01 shard packs
    "csamd_cspack_init": (None, [P(A.cspack_workspace)]),
    "csamd_cspack_clear": (None, [P(A.cspack_workspace)]),
    "csamd_cspack_setup": (i32, [P(A.cspack_workspace), i32, vp]),
    "csamd_cspack_write_blocked_ell": (i32, [C.c_char_p, u64, P(A.bucketed_blocked_ell), u32, u32, P(A.cspack_workspace)]),
    "csamd_cspack_write_sliced_ell": (i32, [C.c_char_p, u64, P(A.bucketed_sliced_ell), u32, u32, P(A.cspack_workspace)]),
    "csamd_cspack_open": (i32, [C.c_char_p, u64, P(A.cspack_file)]),
    "csamd_cspack_close": (None, [P(A.cspack_file)]),
    "csamd_cspack_stage_blocked_ell": (i32, [P(A.cspack_file), u64, u32, P(A.cspack_workspace), P(A.bucketed_blocked_ell)]),
    "csamd_cspack_stage_sliced_ell": (i32, [P(A.cspack_file), u64, u32, P(A.cspack_workspace), P(A.bucketed_sliced_ell)]),
    "csamd_cspack_stage_shard_blocked_ell": (i32, [P(A.cspack_file), u32, P(A.cspack_workspace), P(A.bucketed_blocked_ell)]),
    "csamd_canonical_compressed_from_bucketed_blocked_ell": (i32, [P(A.bucketed_blocked_ell), u32, vp, i32, vp, vp, u64, P(u64),
                                                                    P(A.convert_workspace)]),
    "csamd_canonical_sliced_ell_from_bucketed_sliced_ell": (i32, [P(A.bucketed_sliced_ell), u32, vp, vp, vp, u64, P(u64),
                                                                   P(A.convert_workspace)]),
    "csamd_compressed_from_coo_raw": (i32, [u32, u32, u32, vp, vp, vp, u32, vp, u32, vp, vp, vp, vp, sz, vp]),
    "csamd_label_reduce_set_timing": (i32, [P(A.label_reduce_workspace), i32]),
    "csamd_serve_init": (None, [P(A.serve_workspace)]),
    "csamd_serve_clear": (None, [P(A.serve_workspace)]),
    "csamd_serve_setup": (i32, [P(A.serve_workspace), P(A.serve_config)]),
    "csamd_serve_add_shard": (i32, [P(A.serve_workspace), P(A.compressed_view), u32, P(u32)]),
    "csamd_serve_submit": (i32, [P(A.serve_workspace), P(A.serve_query), P(u64)]),
    "csamd_serve_pump": (i32, [P(A.serve_workspace), P(u32)]),
    "csamd_serve_drain": (i32, [P(A.serve_workspace)]),
    "csamd_serve_get_result": (i32, [P(A.serve_workspace), u64, P(A.serve_result)]),
    "csamd_serve_release": (i32, [P(A.serve_workspace), u64]),
    "csamd_serve_replay": (i32, [P(A.serve_workspace), P(A.serve_query), P(C.c_double), u32, P(u64), P(C.c_double), P(C.c_double)]),
    "csamd_serve_materialize": (i32, [u32, u32, u32, u32, u32, vp, vp, vp]),
    "csamd_decode_values": (i32, [u32, vp, u64, vp, vp]),
    # convert
    "csamd_convert_init": (None, [P(A.convert_workspace)]),
    "csamd_convert_clear": (None, [P(A.convert_workspace)]),
    "csamd_convert_setup": (i32, [P(A.convert_workspace), i32, vp]),
    "csamd_blocked_ell_shape": (i32, [P(A.compressed_view), u32, P(A.convert_workspace), P(u32), P(u64), P(C.c_double)]),
    "csamd_choose_blocked_ell_block_size": (i32, [P(A.compressed_view), P(u32), u32, u32, P(A.convert_workspace), P(A.blocked_ell_tune_result)]),
    "csamd_blocked_ell_from_compressed": (i32, [P(A.compressed_view), u32, P(A.blocked_ell_view), P(A.convert_workspace)]),
    "csamd_sliced_ell_layout": (i32, [P(A.compressed_view), u32, vp, vp, vp, P(A.convert_workspace), P(u32), P(u64)]),
    "csamd_sliced_ell_layout_from_blocked_ell": (i32, [P(A.blocked_ell_view), u32, u32, vp, vp, vp, P(A.convert_workspace), P(u32), P(u64)]),
    "csamd_choose_sliced_ell_slice_rows": (i32, [P(A.compressed_view), P(u32), u32, P(A.convert_workspace), P(A.sliced_ell_tune_result)]),
    "csamd_sliced_ell_from_compressed": (i32, [P(A.compressed_view), u32, P(A.sliced_ell_view), P(A.convert_workspace)]),
    "csamd_sliced_ell_from_blocked_ell": (i32, [P(A.blocked_ell_view), u32, P(A.sliced_ell_view), P(A.convert_workspace)]),
    "csamd_compressed_transpose_scratch_bytes": (sz, [u32, u32, u32]),
    "csamd_build_compressed_transpose_raw": (i32, [u32, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
    "csamd_build_compressed_transpose_raw_typed": (i32, [u32, u32, u32, vp, vp, vp, u32, vp, vp, vp, vp, sz, vp]),
    # bucket
    "csamd_clamp_bucket_count": (u32, [u32, u32]),
    "csamd_major_nnz_bucket_sort_scratch_bytes": (i32, [u32, P(sz)]),
    "csamd_major_nnz_bucket_scan_scratch_bytes": (i32, [u32, P(sz)]),
    "csamd_build_major_nnz_bucket_plan_raw": (i32, [vp, u32, vp, vp, vp, vp, vp, u32, vp, sz, vp]),
    "csamd_fill_log2_bucket_offsets_raw": (i32, [vp, u32, u32, u32, vp, vp]),
    "csamd_rebuild_compressed_major_order_raw": (i32, [vp, vp, vp, u32, u32, vp, vp, vp, vp, vp, vp, sz, vp]),
    "csamd_blocked_ell_bucket_plan_raw": (i32, [vp, u32, u32, vp, vp, vp, P(A.convert_workspace)]),
    # host-struct converters
    "csamd_host_compressed_clear": (None, [P(A.host_compressed)]),
    "csamd_host_blocked_ell_clear": (None, [P(A.host_blocked_ell)]),
    "csamd_host_sliced_ell_clear": (None, [P(A.host_sliced_ell)]),
    "csamd_host_choose_blocked_ell_block_size": (i32, [P(A.host_compressed), P(u32), u32, P(A.blocked_ell_tune_result)]),
    "csamd_host_blocked_ell_from_compressed": (i32, [P(A.host_compressed), u32, P(A.host_blocked_ell)]),
    "csamd_host_blocked_ell_from_compressed_auto": (i32, [P(A.host_compressed), P(u32), u32, P(A.host_blocked_ell),
                                                          P(A.blocked_ell_tune_result)]),
    "csamd_host_choose_sliced_ell_slice_rows": (i32, [P(A.host_compressed), P(u32), u32, P(A.sliced_ell_tune_result)]),
    "csamd_host_sliced_ell_from_compressed": (i32, [P(A.host_compressed), u32, P(A.host_sliced_ell)]),
    "csamd_host_sliced_ell_from_compressed_auto": (i32, [P(A.host_compressed), P(u32), u32, P(A.host_sliced_ell),
                                                         P(A.sliced_ell_tune_result)]),
    "csamd_host_sliced_ell_from_blocked_ell": (i32, [P(A.host_blocked_ell), u32, P(A.host_sliced_ell)]),
    # staging
    "csamd_allocation_cache_init": (None, [P(A.allocation_cache)]),
    "csamd_allocation_cache_clear": (None, [P(A.allocation_cache)]),
    "csamd_take_cached_allocation": (vp, [P(A.allocation_cache), i32, sz, P(sz)]),
    "csamd_cache_allocation": (i32, [P(A.allocation_cache), i32, vp, sz]),
    "csamd_upload_compressed": (i32, [P(A.host_compressed), P(A.partition_record), P(A.compressed_view), P(A.allocation_cache), vp]),
    "csamd_upload_blocked_ell": (i32, [P(A.host_blocked_ell), P(A.partition_record), P(A.blocked_ell_view), P(A.allocation_cache), vp]),
    "csamd_upload_sliced_ell": (i32, [P(A.host_sliced_ell), P(A.partition_record), P(A.sliced_ell_view), P(A.allocation_cache), vp]),
    "csamd_release_record": (i32, [P(A.partition_record), P(A.allocation_cache)]),
    "csamd_upload_blocked_ell_shard": (i32, [P(P(A.host_blocked_ell)), u32, P(A.partition_record), P(A.blocked_ell_view),
                                             P(A.allocation_cache), vp]),
    "csamd_upload_compressed_shard": (i32, [P(P(A.host_compressed)), u32, P(A.partition_record), P(A.compressed_view),
                                            P(A.allocation_cache), vp]),
    # sharded<T> metadata and sharded_device<T> residency
    "csamd_sharded_init": (None, [P(A.sharded)]),
    "csamd_sharded_clear": (None, [P(A.sharded), u32]),
    "csamd_host_part_alloc": (vp, [u32]),
    "csamd_host_compressed_allocate": (i32, [P(A.host_compressed)]),
    "csamd_sharded_reserve_partitions": (i32, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_reserve_shards": (i32, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_rebuild_shard_parts": (None, [P(A.sharded)]),
    "csamd_sharded_rebuild_partition_offsets": (None, [P(A.sharded)]),
    "csamd_sharded_define_partitions": (i32, [P(A.sharded), u32, C.c_ulong, C.c_ulong, vp, vp, vp]),
    "csamd_sharded_set_shards_to_partitions": (i32, [P(A.sharded)]),
    "csamd_sharded_append_partition": (i32, [P(A.sharded), u32, vp]),
    "csamd_sharded_concatenate": (i32, [P(A.sharded), P(A.sharded)]),
    "csamd_sharded_set_equal_shards": (i32, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_build_shard_offsets_by_rows": (i32, [P(A.sharded), C.c_ulong, P(P(C.c_ulong)), P(C.c_ulong)]),
    "csamd_sharded_reshard": (i32, [P(A.sharded), C.c_ulong, vp]),
    "csamd_sharded_set_shards_by_nnz": (i32, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_set_shards_by_part_bytes": (i32, [P(A.sharded), u32, sz]),
    "csamd_sharded_drop_partition": (i32, [P(A.sharded), u32, C.c_ulong]),
    "csamd_sharded_drop_shard": (i32, [P(A.sharded), u32, C.c_ulong]),
    "csamd_partition_aux": (C.c_ulong, [u32, vp]),
    "csamd_sharded_find_partition": (C.c_ulong, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_find_shard": (C.c_ulong, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_first_partition_in_shard": (C.c_ulong, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_last_partition_in_shard": (C.c_ulong, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_row_is_partition_boundary": (i32, [P(A.sharded), C.c_ulong]),
    "csamd_sharded_partition_bytes": (sz, [P(A.sharded), u32, C.c_ulong]),
    "csamd_sharded_shard_bytes": (sz, [P(A.sharded), u32, C.c_ulong]),
    "csamd_sharded_bytes": (sz, [P(A.sharded), u32]),
    "csamd_sharded_device_init": (None, [P(A.sharded_device)]),
    "csamd_sharded_device_clear": (i32, [P(A.sharded_device)]),
    "csamd_sharded_device_reserve": (i32, [P(A.sharded_device), C.c_ulong]),
    "csamd_sharded_device_partition_bytes": (sz, [P(A.sharded), u32, C.c_ulong]),
    "csamd_sharded_device_shard_bytes": (sz, [P(A.sharded), u32, C.c_ulong]),
    "csamd_sharded_set_shards_by_device_bytes": (i32, [P(A.sharded), u32, sz]),
    "csamd_sharded_device_upload_partition": (i32, [P(A.sharded_device), P(A.sharded), u32, C.c_ulong, i32, vp]),
    "csamd_sharded_device_release_partition": (i32, [P(A.sharded_device), C.c_ulong]),
    "csamd_sharded_device_upload_shard": (i32, [P(A.sharded_device), P(A.sharded), u32, C.c_ulong, i32, vp]),
    "csamd_sharded_device_release_shard": (i32, [P(A.sharded_device), P(A.sharded), C.c_ulong]),
    "csamd_sharded_device_stage_shard": (i32, [P(A.sharded_device), P(A.sharded), u32, P(A.cspack_file), C.c_ulong, i32, vp, i32]),
    "csamd_sharded_device_stage_partition": (i32, [P(A.sharded_device), P(A.sharded), u32, P(A.cspack_file), C.c_ulong, i32, vp, i32]),
    "csamd_sharded_device_swap_shard": (i32, [P(A.sharded_device), P(A.sharded), u32, P(A.cspack_file), C.c_ulong, C.c_ulong, i32,
                                              vp, i32, i32]),
    # synth / scan
    "csamd_synth_row_nnz": (i32, [P(A.synth_params), vp, vp]),
    "csamd_exclusive_scan_u32": (i32, [vp, u32, vp, P(A.convert_workspace), P(u64)]),
    "csamd_synth_fill": (i32, [P(A.synth_params), vp, vp, vp, vp]),
    "csamd_synth_log1p_normalize": (i32, [u32, u32, vp, vp, vp, f32, u32, vp, vp, P(A.convert_workspace)]),
}


class LibraryMissing(RuntimeError):
    pass


class CellShardError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises LibraryMissing if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().csamd_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc != 1:
        raise CellShardError(f"{what} failed: {last_error()}")
