"""csamd_compute_sparse_group_reduce_compressed_noalloc: the allocation-free
variant of compute_sparse_group_reduce over compressed rows.

Two calls in a row through the C ABI (the second on a smaller matrix of
another value type, with fewer groups) write caller-owned outputs that must
equal, bit for bit, what csamd_compute_sparse_group_reduce_compressed_typed
writes into its workspace for the same inputs: the two share their kernels,
and every row is reduced by one wave in a fixed order.  The outputs start
filled with a poison pattern, so an output the variant leaves unwritten
fails.  Oversized and inconsistent requests are refused before any launch.
"""
import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd import runtime as rt
from cellshard_amd._lib import last_error
from cellshard_amd.device import DeviceCSR, ptr, to_dev_u32, to_dev_u8
from cellshard_amd.ops import device_to_numpy
from helpers import BF16, F16, encode

pytestmark = pytest.mark.gpu

POISON = 0x7FC00001  # a NaN bit pattern no reduction writes


def _csr(rows, cols, mean_nnz, seed, vt):
    rng = np.random.default_rng(seed)
    n = rng.poisson(mean_nnz, rows).astype(np.int64)
    n[rng.random(rows) < 0.05] = 0
    ptr_ = np.concatenate([[0], np.cumsum(n)]).astype(np.uint32)
    idx = rng.integers(0, cols, int(n.sum())).astype(np.uint32)
    vals = rng.choice([0.0, 1.0, 2.0, 3.0, 5.0, 7.5], size=len(idx), p=[0.05, 0.4, 0.2, 0.15, 0.1, 0.1])
    return DeviceCSR.from_numpy(rows, cols, ptr_, idx, encode(vals, vt), value_type=vt)


def _masks(rows, cols, G, seed):
    rng = np.random.default_rng(seed)
    rk = (rng.random(rows) < 0.85).astype(np.uint8)
    fk = (rng.random(cols) < 0.3).astype(np.uint8)
    fgm = rng.integers(0, 2 ** 32, size=cols, dtype=np.uint64).astype(np.uint32)
    fgm[rng.random(cols) < 0.5] = 0  # features in no group
    return to_dev_u8(rk), to_dev_u8(fk), to_dev_u32(fgm)


class _Outputs:
    """Caller-owned device outputs for `rows` x `G`, poisoned."""

    def __init__(self, rows, G):
        dev = torch.device("cuda", 0)
        full = lambda n: torch.full((max(n, 1),), POISON, dtype=torch.int32, device=dev)
        self.totals, self.maxv, self.det = full(rows), full(rows), full(rows)
        self.keep = torch.full((max(rows, 1),), 0xA5, dtype=torch.uint8, device=dev)
        self.gc, self.gp = full(rows * G), full(rows * G)
        self.res = A.sparse_group_reduce_result(rows=rows, group_count=G)
        self.res.row_totals, self.res.max_values = ptr(self.totals), ptr(self.maxv)
        self.res.detected_features, self.res.row_keep = ptr(self.det), ptr(self.keep)
        if G:
            self.res.group_counts, self.res.group_percentages = ptr(self.gc), ptr(self.gp)


def _fetch(res, rows, G):
    torch.cuda.synchronize()
    out = [device_to_numpy(res.row_totals, rows, np.uint32), device_to_numpy(res.max_values, rows, np.uint32),
           device_to_numpy(res.detected_features, rows, np.uint32), device_to_numpy(res.row_keep, rows, np.uint8)]
    if G:
        out += [device_to_numpy(res.group_counts, rows * G, np.uint32), device_to_numpy(res.group_percentages, rows * G, np.uint32)]
    return out


@pytest.fixture()
def ws(gpu):
    w = A.sparse_group_reduce_workspace()
    rt.init(w)
    assert rt.setup(w, 0, torch.cuda.current_stream().cuda_stream) == 1
    yield w
    rt.clear(w)


def test_noalloc_matches_workspace_entry_point(ws):
    stream = torch.cuda.current_stream().cuda_stream
    filt = A.sparse_group_filter_params(2.0, 1, 0.8, 1)
    cases = [(6000, 9000, 120.0, 11, F16, 9), (900, 2100, 40.0, 12, BF16, 3)]  # the second call is smaller
    for rows, cols, mean, seed, vt, G in cases:
        print(f"noalloc case rows={rows} cols={cols} vt={vt} G={G}", flush=True)
        src = _csr(rows, cols, mean, seed, vt)
        d_rk, d_fk, d_fgm = _masks(rows, cols, G, seed + 100)
        groups = A.group_mask_config_view(group_count=G, feature_group_masks=ptr(d_fgm))
        masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk))
        ref = A.sparse_group_reduce_result()
        assert rt.compute_sparse_group_reduce_compressed_fallback(src.view(), ws, groups, masks, filt, ref, value_type=vt) == 1
        want = _fetch(ref, rows, G)
        out = _Outputs(rows, G)
        assert rt.compute_sparse_group_reduce_compressed_noalloc(src.view(), vt, groups, masks, filt, out.res, stream) == 1, last_error()
        got = _fetch(out.res, rows, G)
        for name, a, b in zip(["totals", "max", "detected", "keep", "group_counts", "group_pct"], got, want):
            assert np.array_equal(a, b), name
        assert int(want[3].sum()) not in (0, rows)  # the filter keeps some rows and drops others


def test_noalloc_refuses_before_launch(gpu):
    stream = torch.cuda.current_stream().cuda_stream
    filt = A.sparse_group_filter_params(0.0, 0, 1.0, 0)
    src = _csr(64, 500, 20.0, 3, F16)
    d_rk, d_fk, d_fgm = _masks(64, 500, 4, 4)
    groups = A.group_mask_config_view(group_count=4, feature_group_masks=ptr(d_fgm))
    out = _Outputs(64, 4)
    # more columns than the fixed scratch covers (the stored indices stay < 500)
    wide = src.view()
    wide.cols = 122_304
    assert rt.compute_sparse_group_reduce_compressed_noalloc(wide, F16, groups, None, filt, out.res, stream) == 0
    assert "scratch" in last_error()
    # outputs sized for another group count / too few rows
    assert rt.compute_sparse_group_reduce_compressed_noalloc(src.view(), F16, A.group_mask_config_view(group_count=5,
                                                             feature_group_masks=ptr(d_fgm)), None, filt, out.res, stream) == 0
    short = _Outputs(63, 4)
    assert rt.compute_sparse_group_reduce_compressed_noalloc(src.view(), F16, groups, None, filt, short.res, stream) == 0
    # groups without device masks: there is no workspace to hold uploaded ones
    assert rt.compute_sparse_group_reduce_compressed_noalloc(src.view(), F16, A.group_mask_config_view(group_count=4), None, filt,
                                                             out.res, stream) == 0
    torch.cuda.synchronize()
    # nothing ran: the poison is intact
    assert (device_to_numpy(out.res.row_totals, 64, np.uint32) == POISON).all()
    # the largest width it takes runs (group masks cover every column)
    edge = src.view()
    edge.cols = 122_303
    wide_fgm = to_dev_u32(np.random.default_rng(5).integers(0, 16, size=edge.cols).astype(np.uint32))
    wide_groups = A.group_mask_config_view(group_count=4, feature_group_masks=ptr(wide_fgm))
    assert rt.compute_sparse_group_reduce_compressed_noalloc(edge, F16, wide_groups, None, filt, out.res, stream) == 1, last_error()
    torch.cuda.synchronize()
    assert not (device_to_numpy(out.res.row_totals, 64, np.uint32) == POISON).any()
