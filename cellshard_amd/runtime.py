"""cellshard::runtime (mask_groups) in Python, over the C ABI.

Same names, argument meaning and 1/0 return convention as
include/CellShard/runtime/mask_groups.cuh:195-260; the structs are the
reference layouts (cellshard_amd.abi).  Results alias workspace memory until
the next call, and calls return before the GPU finishes (synchronize the
workspace stream before reading), as in the reference.
"""
from __future__ import annotations

import ctypes as C

from . import abi as A
from ._lib import load

CELLSHARD_MAX_MASK_GROUPS = A.MAX_MASK_GROUPS

sparse_group_reduce_workspace = A.sparse_group_reduce_workspace
sparse_group_reduce_result = A.sparse_group_reduce_result
group_mask_config_view = A.group_mask_config_view
row_feature_mask_view = A.row_feature_mask_view
sparse_group_filter_params = A.sparse_group_filter_params
sparse_group_reduce_fleet_workspace = A.sparse_group_reduce_fleet_workspace
sparse_group_reduce_fleet_result = A.sparse_group_reduce_fleet_result
sparse_group_reduce_fleet_config = A.sparse_group_reduce_fleet_config
masked_sparse_reoptimize_workspace = A.masked_sparse_reoptimize_workspace
masked_sparse_reoptimize_config = A.masked_sparse_reoptimize_config
masked_sparse_reoptimize_result = A.masked_sparse_reoptimize_result
auto_reoptimize_prediction_input = A.auto_reoptimize_prediction_input
auto_reoptimize_prediction = A.auto_reoptimize_prediction


def _ref(x):
    return C.byref(x) if x is not None else None


def init(workspace) -> None:
    lib = load()
    if isinstance(workspace, A.sparse_group_reduce_workspace):
        lib.csamd_group_reduce_init(C.byref(workspace))
    elif isinstance(workspace, A.sparse_group_reduce_fleet_workspace):
        lib.csamd_fleet_init(C.byref(workspace))
    elif isinstance(workspace, A.masked_sparse_reoptimize_workspace):
        lib.csamd_reoptimize_init(C.byref(workspace))
    else:
        raise TypeError(type(workspace))


def clear(workspace) -> None:
    lib = load()
    if isinstance(workspace, A.sparse_group_reduce_workspace):
        lib.csamd_group_reduce_clear(C.byref(workspace))
    elif isinstance(workspace, A.sparse_group_reduce_fleet_workspace):
        lib.csamd_fleet_clear(C.byref(workspace))
    elif isinstance(workspace, A.masked_sparse_reoptimize_workspace):
        lib.csamd_reoptimize_clear(C.byref(workspace))
    else:
        raise TypeError(type(workspace))


def setup(workspace, device: int, stream: int | None = None) -> int:
    lib = load()
    if isinstance(workspace, A.sparse_group_reduce_workspace):
        return lib.csamd_group_reduce_setup(C.byref(workspace), device, stream)
    if isinstance(workspace, A.masked_sparse_reoptimize_workspace):
        return lib.csamd_reoptimize_setup(C.byref(workspace), device, stream)
    raise TypeError(type(workspace))


def setup_fleet(fleet, config=None) -> int:
    return load().csamd_setup_fleet(C.byref(fleet), _ref(config))


def reserve(workspace, rows: int, cols: int, values: int, group_count: int) -> int:
    return load().csamd_group_reduce_reserve(C.byref(workspace), rows, cols, values, group_count)


def upload_feature_group_masks(workspace, cols: int, host_masks) -> int:
    """host_masks: a sequence of u32 (or None to zero the workspace masks)."""
    if host_masks is None:
        return load().csamd_upload_feature_group_masks(C.byref(workspace), cols, None)
    arr = (C.c_uint32 * cols)(*[int(m) for m in host_masks])
    return load().csamd_upload_feature_group_masks(C.byref(workspace), cols, C.cast(arr, C.c_void_p))


def compute_sparse_group_reduce(src, workspace, groups, masks, filter_, out, value_type: int | None = None) -> int:
    """BELL / SELL overloads (mask_groups.cuh:221-233); value_type selects the
    typed entry point (extension) when given."""
    lib = load()
    args = (C.byref(workspace), _ref(groups), _ref(masks), _ref(filter_), _ref(out))
    if isinstance(src, A.blocked_ell_view):
        if value_type is None:
            return lib.csamd_compute_sparse_group_reduce_blocked_ell(C.byref(src), *args)
        return lib.csamd_compute_sparse_group_reduce_blocked_ell_typed(C.byref(src), value_type, *args)
    if isinstance(src, A.sliced_ell_view):
        if value_type is None:
            return lib.csamd_compute_sparse_group_reduce_sliced_ell(C.byref(src), *args)
        return lib.csamd_compute_sparse_group_reduce_sliced_ell_typed(C.byref(src), value_type, *args)
    raise TypeError(type(src))


def compute_sparse_group_reduce_compressed_fallback(src, workspace, groups, masks, filter_, out,
                                                   value_type: int | None = None) -> int:
    lib = load()
    args = (C.byref(workspace), _ref(groups), _ref(masks), _ref(filter_), _ref(out))
    if value_type is None:
        return lib.csamd_compute_sparse_group_reduce_compressed(C.byref(src), *args)
    return lib.csamd_compute_sparse_group_reduce_compressed_typed(C.byref(src), value_type, *args)


def compute_sparse_group_reduce_compressed_noalloc(src, value_type: int, groups, masks, filter_, out, stream) -> int:
    """compute_sparse_group_reduce over compressed rows into caller-owned
    device outputs (`out` describes them), allocating nothing: the scratch
    is a fixed per-GPU device array.  Returns 1, or 0 with the error set
    (more than 122,303 columns, unaligned arrays, mismatched outputs)."""
    lib = load()
    return lib.csamd_compute_sparse_group_reduce_compressed_noalloc(C.byref(src), value_type, _ref(groups), _ref(masks),
                                                                    _ref(filter_), C.byref(out), stream)


def compute_sparse_group_reduce_fleet(src_by_slot, fleet, groups, masks, filter_, out) -> int:
    lib = load()
    first = src_by_slot[0]
    arr = (type(first) * len(src_by_slot))(*src_by_slot)
    args = (C.byref(fleet), _ref(groups), _ref(masks), _ref(filter_), _ref(out))
    if isinstance(first, A.blocked_ell_view):
        return lib.csamd_compute_sparse_group_reduce_fleet_blocked_ell(arr, *args)
    if isinstance(first, A.sliced_ell_view):
        return lib.csamd_compute_sparse_group_reduce_fleet_sliced_ell(arr, *args)
    raise TypeError(type(first))


def predict_masked_sparse_reoptimization(inp, out) -> int:
    return load().csamd_predict_masked_sparse_reoptimization(_ref(inp), _ref(out))


def manual_reoptimize_masked_sparse(src, masks, config, workspace, out) -> int:
    return load().csamd_manual_reoptimize_masked_sparse(C.byref(src), _ref(masks), _ref(config), C.byref(workspace),
                                                        C.byref(out))
