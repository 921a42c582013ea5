"""Device operations over the C ABI: conversion, transpose, bucketing, the
row-label grouped reduction and the synthetic generator.

Each function calls exactly one family of C-ABI entry points; nothing here
computes results on the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi as A
from ._lib import check, load
from .device import (DeviceBlockedELL, DeviceCSR, DeviceSlicedELL, default_device, ptr, to_dev_f32, to_dev_u32, torch_bits_dtype,
                     value_bytes)


def current_stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def _empty_u32(n: int, device=None) -> torch.Tensor:
    return torch.empty(max(int(n), 0), dtype=torch.int32, device=device or default_device())


class ConvertWorkspace:
    """Owns a csamd_convert_workspace bound to torch's current stream."""

    def __init__(self, device: int | None = None, stream: int | None = None):
        lib = load()
        self.ws = A.convert_workspace()
        lib.csamd_convert_init(C.byref(self.ws))
        dev = torch.cuda.current_device() if device is None else device
        check(lib.csamd_convert_setup(C.byref(self.ws), dev, stream if stream is not None else current_stream_handle()),
              "convert setup")

    def close(self):
        if self.ws is not None:
            load().csamd_convert_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MajorBucketWorkspace:
    """A csamd_compressed_major_bucket_workspace (bucket/operators/major_nnz.cuh:11-213)
    on torch's current stream: build_plan / rebuild / build_bucketed_major_view.
    Results alias the workspace until its next call."""

    def __init__(self, device: int | None = None, stream: int | None = None):
        lib = load()
        self.ws = A.compressed_major_bucket_workspace()
        lib.csamd_major_bucket_init(C.byref(self.ws))
        dev = torch.cuda.current_device() if device is None else device
        check(lib.csamd_major_bucket_setup(C.byref(self.ws), dev, stream if stream is not None else current_stream_handle()),
              "major bucket setup")

    def build_plan(self, src: DeviceCSR, requested_bucket_count: int) -> A.major_nnz_bucket_plan_view:
        plan = A.major_nnz_bucket_plan_view()
        check(load().csamd_major_bucket_build_plan(C.byref(src.view()), requested_bucket_count, C.byref(self.ws), C.byref(plan)),
              "build_plan")
        return plan

    def rebuild(self, src: DeviceCSR, plan) -> A.compressed_major_bucket_result:
        out = A.compressed_major_bucket_result()
        check(load().csamd_major_bucket_rebuild_typed(C.byref(src.view()), src.value_type, C.byref(plan), C.byref(self.ws),
                                                      C.byref(out)), "rebuild")
        return out

    def build_bucketed_major_view(self, src: DeviceCSR, requested_bucket_count: int) -> A.compressed_major_bucket_result:
        out = A.compressed_major_bucket_result()
        check(load().csamd_build_bucketed_major_view_typed(C.byref(src.view()), src.value_type, requested_bucket_count,
                                                           C.byref(self.ws), C.byref(out)), "build_bucketed_major_view")
        return out

    def close(self):
        if self.ws is not None:
            load().csamd_major_bucket_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Blocked-ELL
# ---------------------------------------------------------------------------
def blocked_ell_shape(csr: DeviceCSR, block_size: int, ws: ConvertWorkspace):
    lib = load()
    w = C.c_uint32(0)
    total = C.c_uint64(0)
    fill = C.c_double(0.0)
    v = csr.view()
    check(lib.csamd_blocked_ell_shape(C.byref(v), block_size, C.byref(ws.ws), C.byref(w), C.byref(total), C.byref(fill)),
          "blocked_ell_shape")
    return int(w.value), int(total.value), float(fill.value)


def choose_blocked_ell_block_size(csr: DeviceCSR, candidates, ws: ConvertWorkspace, value_bytes_: int | None = None):
    lib = load()
    cands = (C.c_uint32 * len(candidates))(*candidates)
    out = A.blocked_ell_tune_result()
    v = csr.view()
    vb = value_bytes_ if value_bytes_ is not None else value_bytes(csr.value_type)
    check(lib.csamd_choose_blocked_ell_block_size(C.byref(v), cands, len(candidates), vb, C.byref(ws.ws), C.byref(out)),
          "choose_blocked_ell_block_size")
    return int(out.block_size), float(out.fill_ratio), int(out.padded_bytes)


def blocked_ell_from_compressed(csr: DeviceCSR, block_size: int, ws: ConvertWorkspace) -> DeviceBlockedELL:
    width, _, _ = blocked_ell_shape(csr, block_size, ws)
    rb = (csr.rows + block_size - 1) // block_size
    dev = csr.minor_idx.device
    bci = torch.empty(rb * width, dtype=torch.int32, device=dev)
    val = torch.empty(csr.rows * width * block_size, dtype=torch_bits_dtype(csr.value_type), device=dev)
    out = DeviceBlockedELL(csr.rows, csr.cols, 0, block_size, width * block_size, bci, val, csr.value_type)
    v = out.view()
    src = csr.view()
    check(load().csamd_blocked_ell_from_compressed(C.byref(src), csr.value_type, C.byref(v), C.byref(ws.ws)),
          "blocked_ell_from_compressed")
    out.nnz = int(v.nnz)
    return out


class FilteredBlockedEllWorkspace:
    """filtered_blocked_ell_workspace: the filtered CSR (and its bucketed
    rebuild) alias this workspace's memory until the next call."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.ws = A.filtered_blocked_ell_workspace()
        lib = load()
        lib.csamd_filtered_blocked_ell_init(C.byref(self.ws))
        check(lib.csamd_filtered_blocked_ell_setup(C.byref(self.ws), device,
                                                   stream if stream is not None else current_stream_handle()),
              "filtered_blocked_ell_setup")

    def close(self):
        if self.ws is not None:
            load().csamd_filtered_blocked_ell_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def filter_blocked_ell_to_compressed(bell: DeviceBlockedELL, output_rows: int, output_cols: int, keep_rows=None, keep_cols=None,
                                     row_remap=None, col_remap=None, ws: FilteredBlockedEllWorkspace | None = None,
                                     bucket_count: int | None = None) -> A.filtered_blocked_ell_result:
    """Masked Blocked-ELL -> CSR with row / column remaps
    (convert/filtered_blocked_ell_to_compressed.cuh:264-373); with
    bucket_count, also the major-nnz bucket plan and rebuild (:375-412).
    keep_* are device uint8 tensors, *_remap device uint32/int32 tensors (or
    None).  The result's views point into `ws`."""
    lib = load()
    if ws is None:
        raise ValueError("filter_blocked_ell_to_compressed: pass a FilteredBlockedEllWorkspace (the result aliases it)")
    out = A.filtered_blocked_ell_result()
    v = bell.view()
    args = [C.byref(v), bell.value_type, output_rows, output_cols] + [ptr(t) if t is not None else None
                                                                        for t in (keep_rows, keep_cols, row_remap, col_remap)]
    if bucket_count is None:
        check(lib.csamd_filter_blocked_ell_to_compressed(*args, C.byref(ws.ws), C.byref(out)), "filter_blocked_ell_to_compressed")
    else:
        check(lib.csamd_build_bucketed_filtered_blocked_ell_major_view(*args, bucket_count, C.byref(ws.ws), C.byref(out)),
              "build_bucketed_filtered_blocked_ell_major_view")
    return out


class RepackWorkspace:
    """sharded_blocked_ell_repack_workspace: results alias its memory until
    the next call."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.ws = A.sharded_blocked_ell_repack_workspace()
        lib = load()
        lib.csamd_repack_init(C.byref(self.ws))
        check(lib.csamd_repack_setup(C.byref(self.ws), device, stream if stream is not None else current_stream_handle()),
              "repack_setup")

    def close(self):
        if self.ws is not None:
            load().csamd_repack_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_repacked_shard_blocked_ell(parts: list, output_rows: int, output_cols: int, block_size: int, bucket_count: int,
                                     keep_rows=None, keep_cols=None, row_remap=None, col_remap=None, shard_id: int = 0,
                                     ws: RepackWorkspace | None = None) -> A.sharded_blocked_ell_repack_result:
    """One shard's Blocked-ELL partitions (in shard order) -> masked CSR ->
    bucketed CSR -> Blocked-ELL of `block_size`
    (repack/operators/sharded_blocked_ell.cuh:393-723).  keep_rows /
    row_remap index the shard's rows."""
    if ws is None:
        raise ValueError("build_repacked_shard_blocked_ell: pass a RepackWorkspace (the result aliases it)")
    vts = {b.value_type for b in parts}
    if len(vts) > 1:
        raise ValueError("build_repacked_shard_blocked_ell: partitions of different value types")
    views = (A.blocked_ell_view * max(len(parts), 1))(*[b.view() for b in parts])
    cfg = A.sharded_blocked_ell_repack_config(output_rows, output_cols, block_size, bucket_count)
    out = A.sharded_blocked_ell_repack_result()
    check(load().csamd_build_repacked_blocked_ell(views, len(parts), vts.pop() if vts else A.VALUE_F16, C.byref(cfg),
                                                 *[ptr(t) if t is not None else None for t in (keep_rows, keep_cols, row_remap, col_remap)],
                                                 shard_id, C.byref(ws.ws), C.byref(out)), "build_repacked_shard_blocked_ell")
    return out


def build_shard_column_maps(parts: list, cols: int, ws: ConvertWorkspace):
    """Column order by row-block support signature over a shard's Blocked-ELL
    partitions (convert_layout_support_part.hh:20-86).  Returns host
    (exec_to_canonical, canonical_to_exec) uint32 arrays."""
    vts = {b.value_type for b in parts}
    if len(vts) > 1:
        raise ValueError("build_shard_column_maps: partitions of different value types")
    views = (A.blocked_ell_view * max(len(parts), 1))(*[b.view() for b in parts])
    e2c = np.zeros(max(cols, 1), np.uint32)
    c2e = np.zeros(max(cols, 1), np.uint32)
    check(load().csamd_build_shard_column_maps(views, len(parts), vts.pop() if vts else A.VALUE_F16, cols,
                                              e2c.ctypes.data_as(C.c_void_p), c2e.ctypes.data_as(C.c_void_p), C.byref(ws.ws)),
          "build_shard_column_maps")
    return e2c[:cols], c2e[:cols]


def _coo_view(rows: int, cols: int, row_idx: np.ndarray, col_idx: np.ndarray):
    r = np.ascontiguousarray(row_idx, np.uint32)
    c = np.ascontiguousarray(col_idx, np.uint32)
    coo = A.host_coo(rows=rows, cols=cols, nnz=len(r), storage=None, rowIdx=r.ctypes.data_as(C.c_void_p),
                     colIdx=c.ctypes.data_as(C.c_void_p), val=None)
    return coo, (r, c)


def build_column_order_from_sampled_coo(rows: int, cols: int, row_idx, col_idx, sample_row_rank, device: int = 0):
    """Columns by mean rank of their sampled rows (blocked_ell_bipartite_optimize.cuh:361-434)."""
    coo, keep = _coo_view(rows, cols, row_idx, col_idx)
    rk = np.ascontiguousarray(sample_row_rank, np.uint32)
    e2c, c2e = np.zeros(max(cols, 1), np.uint32), np.zeros(max(cols, 1), np.uint32)
    check(load().csamd_build_column_order_from_sampled_coo(C.byref(coo), rk.ctypes.data_as(C.c_void_p), device,
                                                          e2c.ctypes.data_as(C.c_void_p), c2e.ctypes.data_as(C.c_void_p)),
          "build_column_order_from_sampled_coo")
    return e2c[:cols], c2e[:cols]


def build_row_order_from_coo(rows: int, cols: int, row_idx, col_idx, canonical_to_exec_cols, device: int = 0):
    """Rows by mean exec rank of their columns (blocked_ell_bipartite_optimize.cuh:436-500)."""
    coo, keep = _coo_view(rows, cols, row_idx, col_idx)
    ck = np.ascontiguousarray(canonical_to_exec_cols, np.uint32)
    e2c, c2e = np.zeros(max(rows, 1), np.uint32), np.zeros(max(rows, 1), np.uint32)
    check(load().csamd_build_row_order_from_coo(C.byref(coo), ck.ctypes.data_as(C.c_void_p), device, e2c.ctypes.data_as(C.c_void_p),
                                               c2e.ctypes.data_as(C.c_void_p)),
          "build_row_order_from_coo")
    return e2c[:rows], c2e[:rows]


# ---------------------------------------------------------------------------
# Sliced-ELL
# ---------------------------------------------------------------------------
def _sell_alloc(rows: int, slice_rows: int, dev):
    s = 0 if rows == 0 else min(max(slice_rows, 1), rows)
    count = 0 if rows == 0 else (rows + s - 1) // s
    return s, count, _empty_u32(count + 1, dev), _empty_u32(count, dev), _empty_u32(count, dev)


def sliced_ell_layout(csr: DeviceCSR, slice_rows: int, ws: ConvertWorkspace):
    dev = csr.minor_idx.device
    s, count, offs, widths, slots = _sell_alloc(csr.rows, slice_rows, dev)
    c_count = C.c_uint32(0)
    total = C.c_uint64(0)
    v = csr.view()
    check(load().csamd_sliced_ell_layout(C.byref(v), slice_rows, ptr(offs), ptr(widths), ptr(slots), C.byref(ws.ws),
                                         C.byref(c_count), C.byref(total)), "sliced_ell_layout")
    return s, int(c_count.value), offs, widths, slots, int(total.value)


def sliced_ell_from_compressed(csr: DeviceCSR, slice_rows: int, ws: ConvertWorkspace) -> DeviceSlicedELL:
    s, count, offs, widths, slots, total = sliced_ell_layout(csr, slice_rows, ws)
    dev = csr.minor_idx.device
    out = DeviceSlicedELL(csr.rows, csr.cols, csr.nnz, count, s, offs, widths, slots,
                          torch.empty(total, dtype=torch.int32, device=dev),
                          torch.empty(total, dtype=torch_bits_dtype(csr.value_type), device=dev), csr.value_type)
    v = out.view()
    src = csr.view()
    check(load().csamd_sliced_ell_from_compressed(C.byref(src), csr.value_type, C.byref(v), C.byref(ws.ws)),
          "sliced_ell_from_compressed")
    return out


def choose_sliced_ell_slice_rows(csr: DeviceCSR, candidates, ws: ConvertWorkspace):
    cands = (C.c_uint32 * len(candidates))(*candidates)
    out = A.sliced_ell_tune_result()
    v = csr.view()
    check(load().csamd_choose_sliced_ell_slice_rows(C.byref(v), cands, len(candidates), C.byref(ws.ws), C.byref(out)),
          "choose_sliced_ell_slice_rows")
    return {k: getattr(out, k) for k in ("slice_rows", "slice_count", "padded_slots", "padded_bytes", "storage_ratio",
                                         "slice_penalty", "score")}


def sliced_ell_from_blocked_ell(bell: DeviceBlockedELL, slice_rows: int, ws: ConvertWorkspace) -> DeviceSlicedELL:
    dev = bell.val.device
    s, count, offs, widths, slots = _sell_alloc(bell.rows, slice_rows, dev)
    c_count = C.c_uint32(0)
    total = C.c_uint64(0)
    bv = bell.view()
    check(load().csamd_sliced_ell_layout_from_blocked_ell(C.byref(bv), bell.value_type, slice_rows, ptr(offs), ptr(widths),
                                                          ptr(slots), C.byref(ws.ws), C.byref(c_count), C.byref(total)),
          "sliced_ell_layout_from_blocked_ell")
    t = int(total.value)
    out = DeviceSlicedELL(bell.rows, bell.cols, bell.nnz, count, s, offs, widths, slots,
                          torch.empty(t, dtype=torch.int32, device=dev),
                          torch.empty(t, dtype=torch_bits_dtype(bell.value_type), device=dev), bell.value_type)
    v = out.view()
    check(load().csamd_sliced_ell_from_blocked_ell(C.byref(bv), bell.value_type, C.byref(v), C.byref(ws.ws)),
          "sliced_ell_from_blocked_ell")
    return out


# ---------------------------------------------------------------------------
# Transpose
# ---------------------------------------------------------------------------
LAYOUT_NAMES = {A.LAYOUT_COMPRESSED: "compressed", A.LAYOUT_SLICED_ELL: "sliced_ell", A.LAYOUT_BLOCKED_ELL: "blocked_ell"}


def tune_partitions(parts, ws: ConvertWorkspace) -> list[dict]:
    """Per-partition layout choice for CSR partitions (one launch for all):
    the reference's Blocked-ELL {8,16,32} and Sliced-ELL {32,64} choices,
    the bytes of each format, the smallest layout and the log2 row-length
    histogram."""
    if not parts:
        return []
    vt = parts[0].value_type
    views = (A.compressed_view * len(parts))(*[p.view() for p in parts])
    out = (A.partition_tune_result * len(parts))()
    check(load().csamd_tune_partitions(views, len(parts), vt, C.byref(ws.ws), out), "tune_partitions")
    res = []
    for r in out:
        res.append(dict(layout=LAYOUT_NAMES.get(int(r.layout), "unknown"), bell_block_size=int(r.bell_block_size),
                        bell_ell_width=int(r.bell_ell_width), bell_fill_ratio=float(r.bell_fill_ratio),
                        sell_slice_rows=int(r.sell_slice_rows), sell_slice_count=int(r.sell_slice_count),
                        sell_padded_slots=int(r.sell_padded_slots), sell_score=float(r.sell_score),
                        csr_bytes=int(r.csr_bytes), bell_bytes=int(r.bell_bytes), sell_bytes=int(r.sell_bytes),
                        row_length_log2_hist=[int(x) for x in r.row_length_log2_hist]))
    return res


def _host_coo(rows, cols_src, row_idx, col_idx, val16, keep):
    arrs = [np.ascontiguousarray(row_idx, np.uint32), np.ascontiguousarray(col_idx, np.uint32),
            np.ascontiguousarray(val16).view(np.uint16)]
    keep += arrs
    h = A.host_coo()
    h.rows, h.cols, h.nnz = rows, cols_src, len(arrs[0])
    h.rowIdx, h.colIdx, h.val = (a.ctypes.data if a.size else None for a in arrs)
    return h


def blocked_ell_from_coo(rows, cols_src, row_idx, col_idx, val16, cols, block_size, feature_to_global=None, device=0):
    """Host COO (fp16 bits) -> host Blocked-ELL through the GPU: returns
    (block_col_idx u32, val u16 bits, ell_cols, nnz)."""
    keep = []
    h = _host_coo(rows, cols_src, row_idx, col_idx, val16, keep)
    fmap = None if feature_to_global is None else np.ascontiguousarray(feature_to_global, np.uint32)
    dst = A.host_blocked_ell()
    lib = load()
    check(lib.csamd_blocked_ell_from_coo(C.byref(h), cols, fmap.ctypes.data if fmap is not None else None, block_size,
                                         C.byref(dst), device), "blocked_ell_from_coo")
    rb = (rows + block_size - 1) // block_size
    w = dst.ell_cols // block_size
    bci = np.ctypeslib.as_array(C.cast(dst.blockColIdx, C.POINTER(C.c_uint32)), (rb * w,)).copy() if rb * w else np.zeros(0, np.uint32)
    val = (np.ctypeslib.as_array(C.cast(dst.val, C.POINTER(C.c_uint16)), (rows * dst.ell_cols,)).copy()
           if rows * dst.ell_cols else np.zeros(0, np.uint16))
    ell_cols, nnz = int(dst.ell_cols), int(dst.nnz)
    lib.csamd_host_blocked_ell_clear(C.byref(dst))
    return bci, val, ell_cols, nnz


def choose_blocked_ell_block_size_from_coo(rows, cols_src, row_idx, col_idx, val16, cols, candidates, feature_to_global=None,
                                           device=0):
    keep = []
    h = _host_coo(rows, cols_src, row_idx, col_idx, val16, keep)
    fmap = None if feature_to_global is None else np.ascontiguousarray(feature_to_global, np.uint32)
    cands = (C.c_uint32 * len(candidates))(*candidates)
    out = A.blocked_ell_tune_result()
    check(load().csamd_choose_blocked_ell_block_size_from_coo(C.byref(h), cols, fmap.ctypes.data if fmap is not None else None,
                                                              cands, len(candidates), device, C.byref(out)),
          "choose_blocked_ell_block_size_from_coo")
    return int(out.block_size), float(out.fill_ratio), int(out.padded_bytes)


def compressed_from_coo(rows, cols, row_idx: torch.Tensor, col_idx: torch.Tensor, val: torch.Tensor, value_type: int,
                        feature_to_global: torch.Tensor | None = None, stream: int | None = None) -> DeviceCSR:
    """Device COO -> device CSR (stable by row, then column)."""
    lib = load()
    nnz = int(row_idx.numel())
    dev = row_idx.device
    scratch = torch.empty(int(lib.csamd_compressed_from_coo_scratch_bytes(nnz)), dtype=torch.uint8, device=dev)
    out_ptr = _empty_u32(rows + 1, dev)
    out_idx = _empty_u32(nnz, dev)
    out_val = torch.empty(nnz, dtype=val.dtype, device=dev)
    check(lib.csamd_compressed_from_coo_raw(rows, cols, nnz, ptr(row_idx), ptr(col_idx), ptr(val), value_type, ptr(feature_to_global),
                                            int(feature_to_global.numel()) if feature_to_global is not None else 0, ptr(out_ptr),
                                            ptr(out_idx), ptr(out_val), ptr(scratch), scratch.numel(),
                                            stream if stream is not None else current_stream_handle()), "compressed_from_coo")
    torch.cuda.synchronize()
    return DeviceCSR(rows, cols, out_ptr, out_idx, out_val, value_type, A.COMPRESSED_BY_ROW)


def _bits_dtype(value_type):
    return {1: np.uint8, 2: np.uint16, 4: np.uint32}[value_bytes(value_type)]


def build_bucketed_blocked_ell(bell: DeviceBlockedELL, requested_bucket_count: int, ws: ConvertWorkspace) -> dict:
    """Bucketed execution partition of a device Blocked-ELL, copied to host
    numpy: segments [(rows, ell_cols, block_col_idx, val bits)], segment row
    offsets, exec <-> canonical row maps, the packed-byte estimate."""
    lib = load()
    out = A.bucketed_blocked_ell()
    v = bell.view()
    check(lib.csamd_build_bucketed_blocked_ell(C.byref(v), bell.value_type, requested_bucket_count, C.byref(ws.ws), C.byref(out)),
          "build_bucketed_blocked_ell")
    try:
        segs = []
        dt = _bits_dtype(bell.value_type)
        for k in range(out.segment_count):
            sv = out.segments[k]
            rb = (sv.rows + sv.block_size - 1) // sv.block_size
            w = sv.ell_cols // sv.block_size
            segs.append(dict(rows=int(sv.rows), ell_cols=int(sv.ell_cols), nnz=int(sv.nnz),
                             block_col_idx=device_to_numpy(sv.blockColIdx, rb * w, np.uint32),
                             val=device_to_numpy(sv.val, sv.rows * sv.ell_cols, dt)))
        return dict(segments=segs, segment_row_offsets=[int(out.segment_row_offsets[k]) for k in range(out.segment_count + 1)],
                    exec_to_canonical=device_to_numpy(out.exec_to_canonical_rows, out.rows, np.uint32),
                    canonical_to_exec=device_to_numpy(out.canonical_to_exec_rows, out.rows, np.uint32),
                    col_maps=(device_to_numpy(out.exec_to_canonical_cols, out.cols, np.uint32),
                              device_to_numpy(out.canonical_to_exec_cols, out.cols, np.uint32)),
                    bytes=int(out.bytes))
    finally:
        lib.csamd_bucketed_blocked_ell_clear(C.byref(out))


def choose_bucket_count_blocked_ell(bell: DeviceBlockedELL, ws: ConvertWorkspace):
    n, b = C.c_uint32(0), C.c_uint64(0)
    v = bell.view()
    check(load().csamd_choose_bucket_count_blocked_ell(C.byref(v), C.byref(ws.ws), C.byref(n), C.byref(b)), "choose_bucket_count")
    return int(n.value), int(b.value)


def build_bucketed_sliced_ell(sell: DeviceSlicedELL, requested_bucket_count: int, ws: ConvertWorkspace) -> dict:
    lib = load()
    out = A.bucketed_sliced_ell()
    v = sell.view()
    check(lib.csamd_build_bucketed_sliced_ell(C.byref(v), sell.value_type, requested_bucket_count, C.byref(ws.ws), C.byref(out)),
          "build_bucketed_sliced_ell")
    try:
        segs = []
        dt = _bits_dtype(sell.value_type)
        for k in range(out.segment_count):
            sv = out.segments[k]
            width = int(device_to_numpy(sv.slice_widths, 1, np.uint32)[0]) if sv.slice_count else 0
            segs.append(dict(rows=int(sv.rows), width=width, nnz=int(sv.nnz), slice_count=int(sv.slice_count),
                             col_idx=device_to_numpy(sv.col_idx, sv.rows * width, np.uint32),
                             val=device_to_numpy(sv.val, sv.rows * width, dt)))
        return dict(segments=segs, segment_row_offsets=[int(out.segment_row_offsets[k]) for k in range(out.segment_count + 1)],
                    exec_to_canonical=device_to_numpy(out.exec_to_canonical_rows, out.rows, np.uint32),
                    canonical_to_exec=device_to_numpy(out.canonical_to_exec_rows, out.rows, np.uint32),
                    bytes=int(out.bytes))
    finally:
        lib.csamd_bucketed_sliced_ell_clear(C.byref(out))


def choose_bucket_count_sliced_ell(sell: DeviceSlicedELL, ws: ConvertWorkspace):
    n, b = C.c_uint32(0), C.c_uint64(0)
    v = sell.view()
    check(load().csamd_choose_bucket_count_sliced_ell(C.byref(v), C.byref(ws.ws), C.byref(n), C.byref(b)), "choose_bucket_count_sliced")
    return int(n.value), int(b.value)


# ---------------------------------------------------------------------------
# Device-resident bucketed execution partitions and CSPACK01 shard packs
# ---------------------------------------------------------------------------
def _sliced_segment_to_host(sv, dt) -> dict:
    n = int(sv.slice_count)
    offs = device_to_numpy(sv.slice_row_offsets, n + 1, np.uint32) if sv.slice_row_offsets else np.zeros(1, np.uint32)
    wid = device_to_numpy(sv.slice_widths, n, np.uint32)
    slot = device_to_numpy(sv.slice_slot_offsets, n, np.uint32)
    slots = int(((offs[1:].astype(np.int64) - offs[:-1]) * wid).sum()) if n else 0
    return dict(rows=int(sv.rows), cols=int(sv.cols), nnz=int(sv.nnz), slice_count=n, slice_rows=int(sv.slice_rows),
                slice_row_offsets=offs, slice_widths=wid, slice_slot_offsets=slot, width=int(wid[0]) if n else 0,
                col_idx=device_to_numpy(sv.col_idx, slots, np.uint32), val=device_to_numpy(sv.val, slots, dt))


class BucketedBlockedEll:
    """An owned, device-resident bucketed Blocked-ELL execution partition
    (csamd_bucketed_blocked_ell); freed by close()."""

    def __init__(self, st, value_type: int):
        self.st = st
        self.value_type = value_type

    def to_host(self) -> dict:
        out, dt = self.st, _bits_dtype(self.value_type)
        segs = []
        for k in range(out.segment_count):
            sv = out.segments[k]
            rb = (sv.rows + sv.block_size - 1) // sv.block_size if sv.block_size else 0
            w = sv.ell_cols // sv.block_size if sv.block_size else 0
            segs.append(dict(rows=int(sv.rows), cols=int(sv.cols), nnz=int(sv.nnz), block_size=int(sv.block_size),
                             ell_cols=int(sv.ell_cols), block_col_idx=device_to_numpy(sv.blockColIdx, rb * w, np.uint32),
                             val=device_to_numpy(sv.val, sv.rows * sv.ell_cols, dt)))
        return dict(rows=int(out.rows), cols=int(out.cols), nnz=int(out.nnz), segments=segs,
                    segment_row_offsets=np.array([out.segment_row_offsets[k] for k in range(out.segment_count + 1)], np.uint32),
                    exec_to_canonical=device_to_numpy(out.exec_to_canonical_rows, out.rows, np.uint32),
                    canonical_to_exec=device_to_numpy(out.canonical_to_exec_rows, out.rows, np.uint32),
                    col_maps=(device_to_numpy(out.exec_to_canonical_cols, out.cols, np.uint32),
                              device_to_numpy(out.canonical_to_exec_cols, out.cols, np.uint32)),
                    bytes=int(out.bytes))

    def close(self):
        if self.st is not None:
            load().csamd_bucketed_blocked_ell_clear(C.byref(self.st))
            self.st = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BucketedSlicedEll:
    """An owned, device-resident bucketed Sliced-ELL execution partition."""

    def __init__(self, st, value_type: int):
        self.st = st
        self.value_type = value_type

    def to_host(self) -> dict:
        out, dt = self.st, _bits_dtype(self.value_type)
        segs = [_sliced_segment_to_host(out.segments[k], dt) for k in range(out.segment_count)]
        return dict(rows=int(out.rows), cols=int(out.cols), nnz=int(out.nnz), segments=segs,
                    segment_row_offsets=np.array([out.segment_row_offsets[k] for k in range(out.segment_count + 1)], np.uint32),
                    exec_to_canonical=device_to_numpy(out.exec_to_canonical_rows, out.rows, np.uint32),
                    canonical_to_exec=device_to_numpy(out.canonical_to_exec_rows, out.rows, np.uint32),
                    canonical_slice_row_offsets=device_to_numpy(out.canonical_slice_row_offsets, out.canonical_slice_count + 1,
                                                                np.uint32),
                    canonical_slice_widths=device_to_numpy(out.canonical_slice_widths, out.canonical_slice_count, np.uint32),
                    bytes=int(out.bytes))

    def close(self):
        if self.st is not None:
            load().csamd_bucketed_sliced_ell_clear(C.byref(self.st))
            self.st = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_bucketed_blocked_ell_device(bell: DeviceBlockedELL, requested_bucket_count: int,
                                      ws: ConvertWorkspace) -> BucketedBlockedEll:
    out = A.bucketed_blocked_ell()
    v = bell.view()
    check(load().csamd_build_bucketed_blocked_ell(C.byref(v), bell.value_type, requested_bucket_count, C.byref(ws.ws), C.byref(out)),
          "build_bucketed_blocked_ell")
    return BucketedBlockedEll(out, bell.value_type)


def build_bucketed_sliced_ell_device(sell: DeviceSlicedELL, requested_bucket_count: int, ws: ConvertWorkspace) -> BucketedSlicedEll:
    out = A.bucketed_sliced_ell()
    v = sell.view()
    check(load().csamd_build_bucketed_sliced_ell(C.byref(v), sell.value_type, requested_bucket_count, C.byref(ws.ws), C.byref(out)),
          "build_bucketed_sliced_ell")
    return BucketedSlicedEll(out, sell.value_type)


class CspackWorkspace:
    """Owns a csamd_cspack_workspace (pinned and device staging buffers kept
    between calls).  reader_threads / chunk_bytes of 0 take the library's
    defaults (8 readers, 8 MiB pieces)."""

    def __init__(self, device: int | None = None, reader_threads: int = 0, chunk_bytes: int = 0, stream: int | None = None):
        lib = load()
        self.ws = A.cspack_workspace()
        lib.csamd_cspack_init(C.byref(self.ws))
        self.ws.reader_threads = reader_threads
        self.ws.chunk_bytes = chunk_bytes
        dev = torch.cuda.current_device() if device is None else device
        check(lib.csamd_cspack_setup(C.byref(self.ws), dev, stream), "cspack setup")

    def close(self):
        if self.ws is not None:
            load().csamd_cspack_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CspackFile:
    """An open CSPACK01 file (csamd_cspack_open): header checked, offsets
    read; `shard_id` None accepts any shard."""

    def __init__(self, path, shard_id: int | None = None):
        self.f = A.cspack_file()
        sid = A.CSPACK_ANY_SHARD if shard_id is None else int(shard_id)
        check(load().csamd_cspack_open(str(path).encode(), sid, C.byref(self.f)), "cspack_open")

    @property
    def shard_id(self) -> int:
        return int(self.f.shard_id)

    @property
    def partition_count(self) -> int:
        return int(self.f.partition_count)

    @property
    def partition_offsets(self) -> list:
        return [int(self.f.partition_offsets[i]) for i in range(self.partition_count)]

    @property
    def partition_bytes(self) -> list:
        return [int(self.f.partition_bytes[i]) for i in range(self.partition_count)]

    def close(self):
        if self.f is not None:
            load().csamd_cspack_close(C.byref(self.f))
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_cspack_blocked_ell(path, shard_id: int, parts: list, value_type: int, ws: CspackWorkspace) -> None:
    arr = (A.bucketed_blocked_ell * max(len(parts), 1))()
    for i, p in enumerate(parts):
        arr[i] = p.st
    check(load().csamd_cspack_write_blocked_ell(str(path).encode(), shard_id, arr, len(parts), value_type, C.byref(ws.ws)),
          "cspack_write_blocked_ell")


def write_cspack_sliced_ell(path, shard_id: int, parts: list, value_type: int, ws: CspackWorkspace) -> None:
    arr = (A.bucketed_sliced_ell * max(len(parts), 1))()
    for i, p in enumerate(parts):
        arr[i] = p.st
    check(load().csamd_cspack_write_sliced_ell(str(path).encode(), shard_id, arr, len(parts), value_type, C.byref(ws.ws)),
          "cspack_write_sliced_ell")


def stage_cspack_blocked_ell(f: CspackFile, partition: int, value_type: int, ws: CspackWorkspace) -> BucketedBlockedEll:
    out = A.bucketed_blocked_ell()
    check(load().csamd_cspack_stage_blocked_ell(C.byref(f.f), partition, value_type, C.byref(ws.ws), C.byref(out)),
          "cspack_stage_blocked_ell")
    return BucketedBlockedEll(out, value_type)


def stage_cspack_sliced_ell(f: CspackFile, partition: int, value_type: int, ws: CspackWorkspace) -> BucketedSlicedEll:
    out = A.bucketed_sliced_ell()
    check(load().csamd_cspack_stage_sliced_ell(C.byref(f.f), partition, value_type, C.byref(ws.ws), C.byref(out)),
          "cspack_stage_sliced_ell")
    return BucketedSlicedEll(out, value_type)


def stage_cspack_shard_blocked_ell(f: CspackFile, value_type: int, ws: CspackWorkspace) -> list:
    n = f.partition_count
    arr = (A.bucketed_blocked_ell * max(n, 1))()
    check(load().csamd_cspack_stage_shard_blocked_ell(C.byref(f.f), value_type, C.byref(ws.ws), arr), "cspack_stage_shard_blocked_ell")
    return [BucketedBlockedEll(A.bucketed_blocked_ell.from_buffer_copy(arr[i]), value_type) for i in range(n)]


def canonical_compressed_from_bucketed_blocked_ell(part: BucketedBlockedEll, ws: ConvertWorkspace) -> DeviceCSR:
    """The partition's canonical-order CSR (the reference's canonical COO,
    rows in canonical order); Blocked-ELL of it via
    blocked_ell_from_compressed is the reference's reconstructed partition."""
    lib = load()
    rows = int(part.st.rows)
    dev = default_device()
    d_ptr = _empty_u32(rows + 1, dev)
    n = C.c_uint64(0)
    check(lib.csamd_canonical_compressed_from_bucketed_blocked_ell(C.byref(part.st), part.value_type, ptr(d_ptr), 0, None, None, 0,
                                                                   C.byref(n), C.byref(ws.ws)), "canonical_compressed (size)")
    nnz = int(n.value)
    idx = _empty_u32(nnz, dev)
    val = torch.empty(nnz, dtype=torch_bits_dtype(part.value_type), device=dev)
    if nnz:
        check(lib.csamd_canonical_compressed_from_bucketed_blocked_ell(C.byref(part.st), part.value_type, ptr(d_ptr), 1, ptr(idx),
                                                                       ptr(val), nnz, C.byref(n), C.byref(ws.ws)), "canonical_compressed")
    return DeviceCSR(rows, int(part.st.cols), d_ptr, idx, val, part.value_type, A.COMPRESSED_BY_ROW)


def reconstruct_canonical_blocked_ell(part: BucketedBlockedEll, block_size: int, ws: ConvertWorkspace) -> DeviceBlockedELL:
    """reconstruct_canonical_blocked_ell_part (execution_builders_part.hh
    :1631-1652) on the device."""
    return blocked_ell_from_compressed(canonical_compressed_from_bucketed_blocked_ell(part, ws), block_size, ws)


def rebuild_canonical_sliced_ell(part: BucketedSlicedEll, ws: ConvertWorkspace) -> DeviceSlicedELL:
    """rebuild_canonical_sliced_partition (execution_runtime_part.hh
    :275-318) on the device: the canonical Sliced-ELL in the partition's
    canonical slice layout."""
    lib = load()
    st = part.st
    dev = default_device()
    ncs = int(st.canonical_slice_count)
    slot = _empty_u32(max(ncs, 1), dev)
    n = C.c_uint64(0)
    check(lib.csamd_canonical_sliced_ell_from_bucketed_sliced_ell(C.byref(st), part.value_type, ptr(slot), None, None, 0, C.byref(n),
                                                                  C.byref(ws.ws)), "canonical_sliced_ell (size)")
    total = int(n.value)
    col = _empty_u32(total, dev)
    val = torch.empty(total, dtype=torch_bits_dtype(part.value_type), device=dev)
    if total:
        check(lib.csamd_canonical_sliced_ell_from_bucketed_sliced_ell(C.byref(st), part.value_type, ptr(slot), ptr(col), ptr(val), total,
                                                                      C.byref(n), C.byref(ws.ws)), "canonical_sliced_ell")
    off = device_to_numpy(st.canonical_slice_row_offsets, ncs + 1, np.uint32)
    wid = device_to_numpy(st.canonical_slice_widths, ncs, np.uint32)
    h = np.diff(off.astype(np.int64))
    uniform = ncs > 0 and bool(np.all(h[:-1] == h[0])) and h[-1] <= h[0]
    return DeviceSlicedELL(int(st.rows), int(st.cols), int(st.nnz), ncs, int(h[0]) if uniform else 0, to_dev_u32(off, dev),
                           to_dev_u32(wid, dev), slot[:ncs], col, val, part.value_type)


class TransposeScratch:
    def __init__(self):
        self.buf = None

    def get(self, nbytes: int, dev) -> torch.Tensor:
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
        return self.buf


def compressed_transpose(csr: DeviceCSR, scratch: TransposeScratch | None = None, stream: int | None = None) -> DeviceCSR:
    """CSR -> CSC (or CSC -> CSR); ascending source index inside each segment."""
    lib = load()
    c_dim = csr.major_dim
    u_dim = csr.rows if csr.axis == A.COMPRESSED_BY_COL else csr.cols
    dev = csr.minor_idx.device
    nbytes = int(lib.csamd_compressed_transpose_scratch_bytes(c_dim, u_dim, csr.nnz))
    scratch = scratch or TransposeScratch()
    buf = scratch.get(nbytes, dev)
    out_ptr = _empty_u32(u_dim + 1, dev)
    out_idx = _empty_u32(csr.nnz, dev)
    out_val = torch.empty(csr.nnz, dtype=csr.val.dtype, device=dev)
    check(lib.csamd_build_compressed_transpose_raw_typed(c_dim, u_dim, csr.nnz, ptr(csr.major_ptr), ptr(csr.minor_idx),
                                                         ptr(csr.val), csr.value_type, ptr(out_ptr), ptr(out_idx), ptr(out_val),
                                                         ptr(buf), buf.numel(),
                                                         stream if stream is not None else current_stream_handle()),
          "compressed_transpose")
    return DeviceCSR(csr.rows, csr.cols, out_ptr, out_idx, out_val, csr.value_type, 1 - csr.axis)


# ---------------------------------------------------------------------------
# Bucketing
# ---------------------------------------------------------------------------
def build_major_nnz_bucket_plan(csr: DeviceCSR, requested_bucket_count: int, stream: int | None = None):
    lib = load()
    n = csr.major_dim
    dev = csr.minor_idx.device
    sort_bytes = C.c_size_t(0)
    check(lib.csamd_major_nnz_bucket_sort_scratch_bytes(n, C.byref(sort_bytes)), "sort scratch bytes")
    sort_tmp = torch.empty(max(int(sort_bytes.value), 256), dtype=torch.uint8, device=dev)
    nnz = _empty_u32(n, dev)
    nnz_sorted = _empty_u32(n, dev)
    order_in = _empty_u32(n, dev)
    order = _empty_u32(n, dev)
    b = int(lib.csamd_clamp_bucket_count(n, requested_bucket_count))
    offsets = _empty_u32(b + 1, dev)
    check(lib.csamd_build_major_nnz_bucket_plan_raw(ptr(csr.major_ptr), n, ptr(nnz), ptr(nnz_sorted), ptr(order_in),
                                                    ptr(order), ptr(offsets), requested_bucket_count, ptr(sort_tmp),
                                                    sort_tmp.numel(), stream if stream is not None else current_stream_handle()),
          "build_major_nnz_bucket_plan")
    return {"major_nnz": nnz, "major_nnz_sorted": nnz_sorted, "major_order": order, "bucket_offsets": offsets,
            "bucket_count": b}


def log2_bucket_offsets(major_nnz_sorted: torch.Tensor, bucket_count: int = 12, min_log2: int = 5,
                        stream: int | None = None) -> torch.Tensor:
    offsets = _empty_u32(bucket_count + 1, major_nnz_sorted.device)
    check(load().csamd_fill_log2_bucket_offsets_raw(ptr(major_nnz_sorted), major_nnz_sorted.numel(), bucket_count, min_log2,
                                                    ptr(offsets), stream if stream is not None else current_stream_handle()),
          "fill_log2_bucket_offsets")
    return offsets


def rebuild_compressed_major_order(csr: DeviceCSR, order: torch.Tensor, stream: int | None = None):
    lib = load()
    n = csr.major_dim
    dev = csr.minor_idx.device
    scan_bytes = C.c_size_t(0)
    check(lib.csamd_major_nnz_bucket_scan_scratch_bytes(n, C.byref(scan_bytes)), "scan scratch bytes")
    scan_tmp = torch.empty(max(int(scan_bytes.value), 256), dtype=torch.uint8, device=dev)
    dptr = _empty_u32(n + 1, dev)
    didx = _empty_u32(csr.nnz, dev)
    dval = torch.empty(csr.nnz, dtype=csr.val.dtype, device=dev)
    inv = _empty_u32(n, dev)
    check(lib.csamd_rebuild_compressed_major_order_raw(ptr(csr.major_ptr), ptr(csr.minor_idx), ptr(csr.val), csr.value_type, n,
                                                       ptr(order), ptr(dptr), ptr(didx), ptr(dval), ptr(inv), ptr(scan_tmp),
                                                       scan_tmp.numel(), stream if stream is not None else current_stream_handle()),
          "rebuild_compressed_major_order")
    return DeviceCSR(csr.rows, csr.cols, dptr, didx, dval, csr.value_type, csr.axis), inv


def blocked_ell_bucket_plan(bell: DeviceBlockedELL, ws: ConvertWorkspace):
    dev = bell.val.device
    n = bell.row_blocks
    order = _empty_u32(n, dev)
    widths = _empty_u32(n, dev)
    keys = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    check(load().csamd_blocked_ell_bucket_plan_raw(ptr(bell.block_col_idx), n, bell.width, ptr(order), ptr(widths), ptr(keys),
                                                   C.byref(ws.ws)), "blocked_ell_bucket_plan")
    return order, widths, keys[:n]


# ---------------------------------------------------------------------------
# Row-label grouped reduction
# ---------------------------------------------------------------------------
class LabelReduce:
    """Workspace-owning runner of csamd_label_reduce_* (results alias the
    workspace until the next call, like the reference's reduce results)."""

    def __init__(self, device: int | None = None, stream: int | None = None):
        lib = load()
        self.ws = A.label_reduce_workspace()
        lib.csamd_label_reduce_init(C.byref(self.ws))
        dev = torch.cuda.current_device() if device is None else device
        check(lib.csamd_label_reduce_setup(C.byref(self.ws), dev, stream if stream is not None else current_stream_handle()),
              "label reduce setup")

    def close(self):
        if self.ws is not None:
            load().csamd_label_reduce_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _query(value_type, labels, group_count, row_keep, feature_keep, feature_scale, moments):
        q = A.label_reduce_query()
        q.group_count = group_count
        q.value_type = value_type
        q.row_labels = ptr(labels)
        q.row_keep = ptr(row_keep)
        q.feature_keep = ptr(feature_keep)
        q.feature_scale = ptr(feature_scale)
        q.compute_moments = 1 if moments else 0
        return q

    def run(self, src, labels: torch.Tensor, group_count: int, row_keep: torch.Tensor | None = None,
            feature_keep: torch.Tensor | None = None, feature_scale: torch.Tensor | None = None, moments: bool = True):
        q = self._query(src.value_type, labels, group_count, row_keep, feature_keep, feature_scale, moments)
        res = A.label_reduce_result()
        v = src.view()
        lib = load()
        if isinstance(src, DeviceCSR):
            rc = lib.csamd_label_reduce_compressed(C.byref(v), C.byref(q), C.byref(self.ws), C.byref(res))
        elif isinstance(src, DeviceSlicedELL):
            rc = lib.csamd_label_reduce_sliced_ell(C.byref(v), C.byref(q), C.byref(self.ws), C.byref(res))
        elif isinstance(src, DeviceBlockedELL):
            rc = lib.csamd_label_reduce_blocked_ell(C.byref(v), C.byref(q), C.byref(self.ws), C.byref(res))
        else:
            raise TypeError(type(src))
        check(rc, "label_reduce")
        return res

    @staticmethod
    def partition_table(parts) -> "C.Array":
        """csamd_label_reduce_partition array for CSR / Sliced-ELL partitions."""
        arr = (A.label_reduce_partition * len(parts))()
        for i, p in enumerate(parts):
            if isinstance(p, DeviceCSR):
                arr[i].layout = A.LAYOUT_COMPRESSED
                arr[i].compressed = p.view()
            elif isinstance(p, DeviceSlicedELL):
                arr[i].layout = A.LAYOUT_SLICED_ELL
                arr[i].sliced_ell = p.view()
            else:
                raise TypeError(type(p))
        return arr

    def run_batched(self, parts, labels: torch.Tensor, group_count: int, row_keep: torch.Tensor | None = None,
                    feature_keep: torch.Tensor | None = None, feature_scale: torch.Tensor | None = None,
                    moments: bool = True, table=None):
        """One reduce over partitions (same columns, same value type); the
        row arrays cover the partitions' rows in order."""
        if not parts:
            raise ValueError("no partitions")
        vt = parts[0].value_type
        if any(p.value_type != vt for p in parts):
            raise ValueError("partitions must share the value type")
        q = self._query(vt, labels, group_count, row_keep, feature_keep, feature_scale, moments)
        arr = table if table is not None else self.partition_table(parts)
        res = A.label_reduce_result()
        check(load().csamd_label_reduce_batched(arr, len(parts), C.byref(q), C.byref(self.ws), C.byref(res)),
              "label_reduce_batched")
        return res

    def set_timing(self, enable: bool) -> None:
        """Record HIP events around the main reduce kernel on later calls."""
        check(load().csamd_label_reduce_set_timing(C.byref(self.ws), 1 if enable else 0), "label reduce timing")

    def last_kernel_ms(self) -> float:
        return float(load().csamd_label_reduce_last_kernel_ms(C.byref(self.ws)))

    @staticmethod
    def fetch(res, with_moments: bool = True) -> dict:
        """Copy a result to host numpy arrays (synchronizes)."""
        G, F = int(res.group_count), int(res.kept_features)
        dl = device_to_numpy
        out = {
            "group_count": G,
            "kept_features": F,
            "kept_rows": int(res.kept_rows),
            "sum": dl(res.sum, G * F, np.float64).reshape(G, F),
            "sumsq": dl(res.sumsq, G * F, np.float64).reshape(G, F),
            "nnz": dl(res.nnz, G * F, np.uint32).reshape(G, F),
            "group_rows": dl(res.group_rows, G, np.uint32),
            "kept_feature_ids": dl(res.kept_feature_ids, F, np.uint32),
        }
        if with_moments and res.mean:
            out["mean"] = dl(res.mean, G * F, np.float64).reshape(G, F)
            out["var"] = dl(res.var, G * F, np.float64).reshape(G, F)
        return out


class LabelReduceGraph:
    """A batched label reduce captured into a HIP graph (own buffers); each
    launch() replays it on a stream, result() waits and returns the
    result struct.  The partitions' and the query's device arrays are bound
    at capture and must stay alive; replays read their current contents."""

    def __init__(self, parts, labels: torch.Tensor, group_count: int, row_keep: torch.Tensor | None = None,
                 feature_keep: torch.Tensor | None = None, feature_scale: torch.Tensor | None = None,
                 moments: bool = True, device: int | None = None):
        if not parts:
            raise ValueError("no partitions")
        vt = parts[0].value_type
        if any(p.value_type != vt for p in parts):
            raise ValueError("partitions must share the value type")
        self._keep = (list(parts), labels, row_keep, feature_keep, feature_scale)
        q = LabelReduce._query(vt, labels, group_count, row_keep, feature_keep, feature_scale, moments)
        table = LabelReduce.partition_table(parts)
        self.g = A.label_reduce_graph()
        dev = torch.cuda.current_device() if device is None else device
        check(load().csamd_label_reduce_graph_capture(table, len(parts), C.byref(q), dev, C.byref(self.g)),
              "label_reduce_graph_capture")

    def launch(self, stream: int | None = None) -> None:
        check(load().csamd_label_reduce_graph_launch(C.byref(self.g), stream if stream is not None else current_stream_handle()),
              "label_reduce_graph_launch")

    def result(self):
        res = A.label_reduce_result()
        check(load().csamd_label_reduce_graph_result(C.byref(self.g), C.byref(res)), "label_reduce_graph_result")
        return res

    def close(self):
        if self.g is not None:
            load().csamd_label_reduce_graph_destroy(C.byref(self.g))
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LabelReducePipeline:
    """Staging -> conversion -> reduce of batched partitions captured into one
    HIP graph (csamd_label_reduce_pipeline_capture).

    parts: the partitions' device CSRs (conversion sources; with `staged`,
    also the staging destinations).  layouts: per partition a dict from
    tune_partitions (or the same keys): layout "compressed", "sliced_ell"
    (sell_slice_rows, sell_slice_count, sell_padded_slots) or "blocked_ell"
    (bell_block_size, bell_ell_width).  staged: True (every partition) or a
    list of flags, to stage those CSRs from pinned host copies inside the
    graph (made here from the device CSRs).  The
    Sliced-ELL / Blocked-ELL outputs are allocated here and filled by every
    launch; sell(i) / bell(i) give them as device layouts."""

    def __init__(self, parts, layouts, labels: torch.Tensor, group_count: int, row_keep: torch.Tensor | None = None,
                 feature_keep: torch.Tensor | None = None, staged=False, moments: bool = True, device: int | None = None):
        if not parts or len(parts) != len(layouts):
            raise ValueError("one layout per partition required")
        staged = [bool(staged)] * len(parts) if isinstance(staged, bool) else [bool(x) for x in staged]
        vt = parts[0].value_type
        if any(p.value_type != vt for p in parts):
            raise ValueError("partitions must share the value type")
        dev = default_device() if device is None else torch.device("cuda", device)
        bits = torch_bits_dtype(vt)
        self.parts, self.layouts = list(parts), list(layouts)
        self._sell, self._bell, self._host = {}, {}, {}
        arr = (A.pipeline_partition * len(parts))()
        for i, (p, lay) in enumerate(zip(parts, layouts)):
            e = arr[i]
            e.compressed = p.view()
            name = lay["layout"]
            if name == "sliced_ell":
                S = min(int(lay["sell_slice_rows"]), p.rows)
                ns = -(-p.rows // S)
                slots = int(lay["sell_padded_slots"])
                u32 = lambda n: torch.empty(max(n, 1), dtype=torch.int32, device=dev)
                sell = DeviceSlicedELL(p.rows, p.cols, p.nnz, ns, S, u32(ns + 1), u32(ns), u32(ns), u32(slots),
                                       torch.empty(max(slots, 1), dtype=bits, device=dev), vt)
                self._sell[i] = (sell, slots)
                e.layout, e.slice_rows, e.slots_capacity, e.sliced_ell = A.LAYOUT_SLICED_ELL, S, slots, sell.view()
            elif name == "blocked_ell":
                bs, W = int(lay["bell_block_size"]), int(lay["bell_ell_width"])
                rb = -(-p.rows // bs)
                bell = DeviceBlockedELL(p.rows, p.cols, 0, bs, W * bs, torch.empty(max(rb * W, 1), dtype=torch.int32, device=dev),
                                        torch.empty(max(p.rows * W * bs, 1), dtype=bits, device=dev), vt)
                nnz = torch.zeros(1, dtype=torch.int32, device=dev)
                self._bell[i] = (bell, nnz)
                e.layout, e.block_size, e.ell_width, e.blocked_ell, e.d_nnz = A.LAYOUT_BLOCKED_ELL, bs, W, bell.view(), ptr(nnz)
            elif name == "compressed":
                e.layout = A.LAYOUT_COMPRESSED
            else:
                raise ValueError(f"partition {i}: layout {name}")
            if staged[i]:
                hp = p.major_ptr.cpu().pin_memory()
                hi = p.minor_idx.cpu().pin_memory()
                hv = p.val.cpu().pin_memory()
                h = A.host_compressed(rows=p.rows, cols=p.cols, nnz=p.nnz, axis=A.COMPRESSED_BY_ROW)
                h.majorPtr, h.minorIdx, h.val = hp.data_ptr(), hi.data_ptr() if p.nnz else None, hv.data_ptr() if p.nnz else None
                self._host[i] = (hp, hi, hv, h)
                e.staged_from = C.pointer(h)
        self._table = arr
        self._keep = (labels, row_keep, feature_keep)
        q = LabelReduce._query(vt, labels, group_count, row_keep, feature_keep, None, moments)
        self.g = A.label_reduce_graph()
        check(load().csamd_label_reduce_pipeline_capture(arr, len(parts), C.byref(q), dev.index, C.byref(self.g)),
              "label_reduce_pipeline_capture")

    def launch(self, stream: int | None = None) -> None:
        check(load().csamd_label_reduce_graph_launch(C.byref(self.g), stream if stream is not None else current_stream_handle()),
              "label_reduce_graph_launch")

    def result(self):
        res = A.label_reduce_result()
        check(load().csamd_label_reduce_graph_result(C.byref(self.g), C.byref(res)), "label_reduce_graph_result")
        return res

    def sell(self, i) -> DeviceSlicedELL:
        return self._sell[i][0]

    def bell(self, i) -> DeviceBlockedELL:
        """The converted Blocked-ELL of partition i with its stored nonzero count (synchronizes)."""
        bell, nnz = self._bell[i]
        torch.cuda.synchronize()
        bell.nnz = int(nnz.cpu().numpy().view(np.uint32)[0])
        return bell

    def close(self):
        if self.g is not None:
            load().csamd_label_reduce_graph_destroy(C.byref(self.g))
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy_d2h(host_u8: torch.Tensor, dev_ptr: int, nbytes: int) -> None:
    """Copy nbytes from a raw device pointer into a host uint8 tensor."""
    hip = _hip()
    rc = hip.hipMemcpy(C.c_void_p(host_u8.data_ptr()), C.c_void_p(dev_ptr), C.c_size_t(nbytes), 2)  # D2H
    if rc != 0:
        raise RuntimeError(f"hipMemcpy D2H failed ({rc})")


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.restype = C.c_int
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _HIP


def device_to_numpy(dev_ptr: int, count: int, dtype) -> np.ndarray:
    """Synchronous copy of `count` elements from a raw device pointer."""
    nbytes = int(count) * np.dtype(dtype).itemsize
    if nbytes == 0 or not dev_ptr:
        return np.zeros(int(count), dtype=dtype)
    torch.cuda.synchronize()
    t = torch.empty(nbytes, dtype=torch.uint8)
    _copy_d2h(t, dev_ptr, nbytes)
    return t.numpy().view(dtype).copy()


# ---------------------------------------------------------------------------
# Synthetic generator
# ---------------------------------------------------------------------------
def column_tables(cols: int, seed: int, zipf_s: float = 1.1, mu_sigma: float = 1.5):
    """Per-column Zipf weights over a seeded permutation and NB means.
    Returns (inverse weights, means) as float32 numpy arrays."""
    rng = np.random.default_rng(seed)
    rank = rng.permutation(cols).astype(np.float64)   # rank of column c
    w = (rank + 1.0) ** (-zipf_s)
    w = w / w.sum()
    mu = rng.lognormal(mean=0.0, sigma=mu_sigma, size=cols)
    return (1.0 / w).astype(np.float32), mu.astype(np.float32)


def synth_csr(rows: int, cols: int, median: float, sigma: float, seed: int, value_type: int = A.VALUE_F16,
              nb_alpha: float = 0.5, zipf_s: float = 1.1, mu_sigma: float = 1.5, ws: ConvertWorkspace | None = None,
              device=None) -> DeviceCSR:
    """Generate a count CSR on the device (see csamd_synth_params)."""
    lib = load()
    dev = device or default_device()
    inv_w, mu = column_tables(cols, seed, zipf_s, mu_sigma)
    d_inv_w = to_dev_f32(inv_w, dev)
    d_mu = to_dev_f32(mu, dev)
    p = A.synth_params()
    p.rows, p.cols, p.value_type, p.seed = rows, cols, value_type, seed
    p.row_nnz_median, p.row_nnz_sigma, p.nb_alpha = median, sigma, nb_alpha
    p.d_col_inv_weight, p.d_col_mean = ptr(d_inv_w), ptr(d_mu)
    own_ws = ws is None
    ws = ws or ConvertWorkspace()
    stream = ws.ws.stream
    row_nnz = _empty_u32(rows, dev)
    check(lib.csamd_synth_row_nnz(C.byref(p), ptr(row_nnz), stream), "synth_row_nnz")
    d_ptr = _empty_u32(rows + 1, dev)
    total = C.c_uint64(0)
    check(lib.csamd_exclusive_scan_u32(ptr(row_nnz), rows, ptr(d_ptr), C.byref(ws.ws), C.byref(total)), "synth scan")
    nnz = int(total.value)
    idx = _empty_u32(nnz, dev)
    val = torch.empty(nnz, dtype=torch_bits_dtype(value_type), device=dev)
    check(lib.csamd_synth_fill(C.byref(p), ptr(d_ptr), ptr(idx), ptr(val), stream), "synth_fill")
    torch.cuda.synchronize()
    if own_ws:
        ws.close()
    del d_inv_w, d_mu, row_nnz
    return DeviceCSR(rows, cols, d_ptr, idx, val, value_type, A.COMPRESSED_BY_ROW)


def log1p_normalize(counts: DeviceCSR, value_type: int, scale: float = 1e4, ws: ConvertWorkspace | None = None):
    """log1p(count / row_total * scale) stored as value_type; for FP8 e4m3 also
    returns the per-feature absmax scale (decoded value * scale[c])."""
    if counts.value_type != A.VALUE_F32:
        raise ValueError("log1p_normalize expects f32 counts")
    lib = load()
    dev = counts.minor_idx.device
    own = ws is None
    ws = ws or ConvertWorkspace()
    out = torch.empty(counts.nnz, dtype=torch_bits_dtype(value_type), device=dev)
    fscale = torch.empty(counts.cols, dtype=torch.float32, device=dev) if value_type == A.VALUE_FP8_E4M3 else None
    check(lib.csamd_synth_log1p_normalize(counts.rows, counts.cols, ptr(counts.major_ptr), ptr(counts.minor_idx), ptr(counts.val),
                                          scale, value_type, ptr(out), ptr(fscale), C.byref(ws.ws)), "log1p_normalize")
    if own:
        ws.close()
    return DeviceCSR(counts.rows, counts.cols, counts.major_ptr, counts.minor_idx, out, value_type, counts.axis), fscale
