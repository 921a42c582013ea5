"""mask_groups (reference ABI) on the GPU: the reference test's fixtures
(tests/mask_groups_runtime_test.cu, F3-corrected) through the same call
sequence, then random parity against the oracle."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd import runtime as csr_rt
from cellshard_amd.device import DeviceBlockedELL, DeviceCSR, DeviceSlicedELL, ptr, to_dev_u32, to_dev_u8
from cellshard_amd.ops import device_to_numpy
from helpers import BF16, F16, F32, FP8, encode, f16_bits, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def fetch(ws, res):
    torch.cuda.synchronize()
    rows, G = res.rows, res.group_count
    out = dict(
        row_totals=device_to_numpy(res.row_totals, rows, np.float32),
        detected=device_to_numpy(res.detected_features, rows, np.uint32),
        max_values=device_to_numpy(res.max_values, rows, np.float32),
        keep=device_to_numpy(res.row_keep, rows, np.uint8),
    )
    if G:
        out["group_counts"] = device_to_numpy(res.group_counts, rows * G, np.float32).reshape(rows, G)
        out["group_pct"] = device_to_numpy(res.group_percentages, rows * G, np.float32).reshape(rows, G)
    return out


@pytest.fixture()
def ws(gpu):
    w = A.sparse_group_reduce_workspace()
    csr_rt.init(w)
    assert csr_rt.setup(w, 0, torch.cuda.current_stream().cuda_stream) == 1
    yield w
    csr_rt.clear(w)


def test_reference_fixture_sequence(ws):
    """Port of tests/mask_groups_runtime_test.cu:52-198."""
    mg = json.loads((GOLD / "mask_groups_fixture.json").read_text())
    e = mg["expected"]
    tol = e["tolerance_abs"]
    rows, cols = mg["rows"], mg["cols"]
    row_keep = to_dev_u8(mg["row_keep"])
    feature_keep = to_dev_u8(mg["feature_keep"])
    assert csr_rt.upload_feature_group_masks(ws, cols, mg["group_masks"]) == 1
    groups = A.group_mask_config_view(group_count=mg["group_count"], feature_group_masks=ws.feature_group_masks)
    masks = A.row_feature_mask_view(row_keep=ptr(row_keep), feature_keep=ptr(feature_keep))
    filt = A.sparse_group_filter_params(*mg["filter"])
    res = A.sparse_group_reduce_result()

    b = mg["blocked_ell"]
    bell = DeviceBlockedELL.from_numpy(rows, cols, b["nnz"], b["block_size"], b["ell_cols"], b["block_col_idx"],
                                       f16_bits(b["values"]))
    assert csr_rt.compute_sparse_group_reduce(bell.view(), ws, groups, masks, filt, res) == 1
    o = fetch(ws, res)
    be = e["blocked_ell_masked"]
    for k, v in be["row_totals"].items():
        assert abs(o["row_totals"][int(k)] - v) <= tol
    for k, v in be["detected"].items():
        assert o["detected"][int(k)] == v
    for k, v in be["keep"].items():
        assert o["keep"][int(k)] == v
    for k, v in be["group_counts"].items():
        assert abs(o["group_counts"].ravel()[int(k)] - v) <= tol
    for k, v in be["group_pct"].items():
        assert abs(o["group_pct"].ravel()[int(k)] - v) <= tol

    no_groups = A.group_mask_config_view(group_count=0)
    assert csr_rt.compute_sparse_group_reduce(bell.view(), ws, no_groups, None, filt, res) == 1
    o = fetch(ws, res)
    assert res.group_count == 0 and abs(o["row_totals"][0] - 15.0) <= tol and o["detected"][0] == 3

    c = mg["compressed"]
    csr = DeviceCSR.from_numpy(rows, cols, c["ptr"], c["idx"], f16_bits(c["values"]))
    assert csr_rt.compute_sparse_group_reduce_compressed_fallback(csr.view(), ws, groups, masks, filt, res) == 1
    o = fetch(ws, res)
    assert abs(o["row_totals"][2] - 7.0) <= tol and abs(o["group_counts"].ravel()[8] - 2.0) <= tol

    s = mg["sliced_ell"]
    sell = DeviceSlicedELL.from_numpy(rows, cols, s["nnz"], s["slice_rows"], s["slice_row_offsets"], s["slice_widths"],
                                      s["slice_slot_offsets"], s["col_idx"], f16_bits(s["values"]))
    assert csr_rt.compute_sparse_group_reduce(sell.view(), ws, groups, masks, filt, res) == 1
    o = fetch(ws, res)
    assert abs(o["row_totals"][0] - 8.0) <= tol and abs(o["row_totals"][3] - 18.0) <= tol

    r = mg["reoptimize"]
    rws = A.masked_sparse_reoptimize_workspace()
    csr_rt.init(rws)
    assert csr_rt.setup(rws, 0) == 1
    row_remap = to_dev_u32(r["row_remap"])
    feature_remap = to_dev_u32(r["feature_remap"])
    rmasks = A.row_feature_mask_view(row_keep=ptr(row_keep), feature_keep=ptr(feature_keep), row_remap=ptr(row_remap),
                                     feature_remap=ptr(feature_remap))
    cfg = A.masked_sparse_reoptimize_config(**r["config"])
    out = A.masked_sparse_reoptimize_result()
    assert csr_rt.manual_reoptimize_masked_sparse(bell.view(), rmasks, cfg, rws, out) == 1
    md = out.metadata
    assert (md.kept_rows, md.kept_features, md.live_nnz, md.output_layout) == (3, 4, 7, A.LAYOUT_COMPRESSED)
    # the compacted CSR equals the oracle's
    mp, mi, mv, nnz = O.reoptimize(rows, cols, b["block_size"], b["ell_cols"], b["block_col_idx"],
                                   f16_bits(b["values"]).view(np.uint16), mg["row_keep"], mg["feature_keep"], r["row_remap"],
                                   r["feature_remap"], 3, 4)
    assert np.array_equal(device_to_numpy(out.compressed.majorPtr, 4, np.uint32), mp)
    assert np.array_equal(device_to_numpy(out.compressed.minorIdx, nnz, np.uint32), mi)
    assert np.array_equal(device_to_numpy(out.compressed.val, nnz, np.uint16), mv)
    csr_rt.clear(rws)


def test_error_paths(ws):
    filt = A.sparse_group_filter_params(0.0, 0, 1.0, 0)
    res = A.sparse_group_reduce_result()
    csr = DeviceCSR.from_numpy(2, 3, [0, 1, 2], [0, 2], f16_bits([1, 2]))
    big = A.group_mask_config_view(group_count=33)
    assert csr_rt.compute_sparse_group_reduce_compressed_fallback(csr.view(), ws, big, None, filt, res) == 0
    csc = DeviceCSR.from_numpy(2, 3, [0, 1, 1, 2], [0, 1], f16_bits([1, 2]), axis=A.COMPRESSED_BY_COL)
    assert csr_rt.compute_sparse_group_reduce_compressed_fallback(csc.view(), ws, None, None, filt, res) == 0
    assert csr_rt.compute_sparse_group_reduce_compressed_fallback(csr.view(), ws, None, None, None, res) == 0


def _compare(o, ref, G, vt):
    # exact for integer-valued inputs well below 2^24
    assert np.array_equal(o["detected"], ref["detected"])
    assert np.array_equal(o["keep"], ref["keep"])
    np.testing.assert_allclose(o["row_totals"], ref["row_totals"], rtol=1e-6, atol=0)
    np.testing.assert_array_equal(o["max_values"], ref["max_values"])
    if G:
        np.testing.assert_allclose(o["group_counts"], ref["group_counts"], rtol=1e-6, atol=0)
        np.testing.assert_allclose(o["group_pct"], ref["group_pct"], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("vt", [F16, F32, BF16, FP8])
@pytest.mark.parametrize("layout", ["csr", "sell", "bell"])
def test_random_parity(ws, vt, layout):
    rows, cols, G = 300, 211, 32 if vt == F16 else 9
    ptr_, idx, val = random_csr(rows, cols, 0.08, seed=31 + vt, zeros=0.03, empty_rows=0.05, fp8_safe=(vt == FP8))
    bits = encode(val, vt)
    rng = np.random.default_rng(vt)
    rk = (rng.random(rows) < 0.9).astype(np.uint8)
    fk = (rng.random(cols) < 0.8).astype(np.uint8)
    fgm = rng.integers(0, 2 ** 32, size=cols, dtype=np.uint64).astype(np.uint32)
    filt_t = [10.0, 3, 0.4, 2]
    d_rk, d_fk, d_fgm = to_dev_u8(rk), to_dev_u8(fk), to_dev_u32(fgm)
    groups = A.group_mask_config_view(group_count=G, feature_group_masks=ptr(d_fgm))
    masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk))
    filt = A.sparse_group_filter_params(*filt_t)
    res = A.sparse_group_reduce_result()
    if layout == "csr":
        src = DeviceCSR.from_numpy(rows, cols, ptr_, idx, bits, vt)
        assert csr_rt.compute_sparse_group_reduce_compressed_fallback(src.view(), ws, groups, masks, filt, res, vt) == 1
        ref = O.group_reduce_csr(rows, cols, ptr_, idx, bits, vt, rk, fk, fgm, G, filt_t)
    elif layout == "sell":
        s = O.sliced_ell_from_compressed(rows, cols, 0, ptr_, idx, bits, 32)
        src = DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                         s["col_idx"], s["val"], vt)
        assert csr_rt.compute_sparse_group_reduce(src.view(), ws, groups, masks, filt, res, vt) == 1
        ref = O.group_reduce_sell(rows, cols, s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"], s["col_idx"],
                                  s["val"], vt, rk, fk, fgm, G, filt_t)
    else:
        bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr_, idx, bits, 8)
        src = DeviceBlockedELL.from_numpy(rows, cols, nnz, 8, ell, bci, bv, vt)
        assert csr_rt.compute_sparse_group_reduce(src.view(), ws, groups, masks, filt, res, vt) == 1
        ref = O.group_reduce_bell(rows, cols, 8, ell, bci, bv, vt, rk, fk, fgm, G, filt_t)
    _compare(fetch(ws, res), ref, G, vt)


def test_fleet_single_gpu(gpu):
    fleet = A.sparse_group_reduce_fleet_workspace()
    csr_rt.init(fleet)
    cfg = A.sparse_group_reduce_fleet_config()
    ids = (C.c_int * 1)(0)
    cfg.device_ids = C.cast(ids, C.c_void_p)
    cfg.device_count = 1
    assert csr_rt.setup_fleet(fleet, cfg) == 1
    rows, cols = 64, 40
    ptr_, idx, val = random_csr(rows, cols, 0.2, seed=3)
    bits = f16_bits(val)
    bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr_, idx, bits, 4)
    src = DeviceBlockedELL.from_numpy(rows, cols, nnz, 4, ell, bci, bv)
    out = A.sparse_group_reduce_fleet_result()
    filt = A.sparse_group_filter_params(0.0, 0, 1.0, 0)
    assert csr_rt.compute_sparse_group_reduce_fleet([src.view()], fleet, None, None, filt, out) == 1
    assert out.slot_count == 1
    torch.cuda.synchronize()
    r = out.slot_results[0]
    got = device_to_numpy(r.row_totals, rows, np.float32)
    ref = O.group_reduce_bell(rows, cols, 4, ell, bci, bv, F16, None, None, None, 0, [0, 0, 1, 0])
    np.testing.assert_allclose(got, ref["row_totals"], rtol=1e-6)
    csr_rt.clear(fleet)


@pytest.mark.parametrize("shift_idx,shift_val", [(0, 0), (1, 1), (3, 3), (1, 2)])
@pytest.mark.parametrize("vt", [F16, F32, FP8])
def test_compressed_views_off_unit_boundaries(ws, vt, shift_idx, shift_val):
    """CSR views into the middle of allocations (indices and values a few
    entries past a 16-byte boundary, equal or different shifts) and long
    rows: the unit-load path, its head / tail entries and the per-entry
    fallback all give the oracle's sums."""
    rows, cols, G = 64, 3000, 32
    ptr_, idx, val = random_csr(rows, cols, 0.45, seed=5 + vt, zeros=0.03, empty_rows=0.05, fp8_safe=(vt == FP8))
    bits = encode(val, vt)
    rng = np.random.default_rng(vt)
    rk = (rng.random(rows) < 0.9).astype(np.uint8)
    fk = (rng.random(cols) < 0.3).astype(np.uint8)
    fgm = rng.integers(0, 2 ** 32, size=cols, dtype=np.uint64).astype(np.uint32)
    filt_t = [10.0, 3, 0.4, 2]
    src = DeviceCSR.from_numpy(rows, cols, ptr_, idx, bits, vt)
    big_i = to_dev_u32(np.concatenate([np.zeros(shift_idx, np.uint32), idx]))
    big_v = torch.cat([torch.zeros(shift_val, dtype=src.val.dtype, device=src.val.device), src.val])
    src.minor_idx = big_i[shift_idx:]
    src.val = big_v[shift_val:]
    d_rk, d_fk, d_fgm = to_dev_u8(rk), to_dev_u8(fk), to_dev_u32(fgm)
    groups = A.group_mask_config_view(group_count=G, feature_group_masks=ptr(d_fgm))
    masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk))
    filt = A.sparse_group_filter_params(*filt_t)
    res = A.sparse_group_reduce_result()
    assert csr_rt.compute_sparse_group_reduce_compressed_fallback(src.view(), ws, groups, masks, filt, res, vt) == 1
    ref = O.group_reduce_csr(rows, cols, ptr_, idx, bits, vt, rk, fk, fgm, G, filt_t)
    _compare(fetch(ws, res), ref, G, vt)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr_, idx, bits, 32)
    sell = DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"], s["col_idx"],
                                      s["val"], vt)
    assert csr_rt.compute_sparse_group_reduce(sell.view(), ws, groups, masks, filt, res, vt) == 1
    _compare(fetch(ws, res), O.group_reduce_sell(rows, cols, s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                                 s["col_idx"], s["val"], vt, rk, fk, fgm, G, filt_t), G, vt)
    for bs in (4, 6, 16):
        bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr_, idx, bits, bs)
        bell = DeviceBlockedELL.from_numpy(rows, cols, nnz, bs, ell, bci, bv, vt)
        assert csr_rt.compute_sparse_group_reduce(bell.view(), ws, groups, masks, filt, res, vt) == 1
        _compare(fetch(ws, res), O.group_reduce_bell(rows, cols, bs, ell, bci, bv, vt, rk, fk, fgm, G, filt_t), G, vt)


@pytest.mark.parametrize("bs", [4, 8, 16, 24, 32, 64])
def test_reoptimize_random_parity(gpu, bs):
    """manual_reoptimize_masked_sparse on random Blocked-ELL (16-byte chunk
    reads when the block size is a multiple of 8, per value otherwise) with
    row / feature masks and compacting remaps, against the oracle."""
    rows, cols = 250, 403
    ptr_, idx, val = random_csr(rows, cols, 0.2, seed=bs, zeros=0.1, empty_rows=0.05)
    bits = f16_bits(val)
    bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr_, idx, bits, bs)
    bell = DeviceBlockedELL.from_numpy(rows, cols, nnz, bs, ell, bci, bv)
    rng = np.random.default_rng(bs)
    rk = (rng.random(rows) < 0.8).astype(np.uint8)
    fk = (rng.random(cols) < 0.5).astype(np.uint8)
    out_rows, out_cols = int(rk.sum()), int(fk.sum())
    rr = np.full(rows, 0xFFFFFFFF, np.uint32)
    rr[rk == 1] = rng.permutation(out_rows)
    fr = np.full(cols, 0xFFFFFFFF, np.uint32)
    fr[fk == 1] = rng.permutation(out_cols)
    d_rk, d_fk, d_rr, d_fr = to_dev_u8(rk), to_dev_u8(fk), to_dev_u32(rr), to_dev_u32(fr)
    rws = A.masked_sparse_reoptimize_workspace()
    csr_rt.init(rws)
    assert csr_rt.setup(rws, 0) == 1
    masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk), row_remap=ptr(d_rr), feature_remap=ptr(d_fr))
    cfg = A.masked_sparse_reoptimize_config(out_rows, out_cols, 0, bs, A.LAYOUT_COMPRESSED)
    out = A.masked_sparse_reoptimize_result()
    assert csr_rt.manual_reoptimize_masked_sparse(bell.view(), masks, cfg, rws, out) == 1
    mp, mi, mv, n = O.reoptimize(rows, cols, bs, ell, bci, np.asarray(bv).view(np.uint16), rk, fk, rr, fr, out_rows, out_cols)
    assert out.metadata.live_nnz == n
    assert np.array_equal(device_to_numpy(out.compressed.majorPtr, out_rows + 1, np.uint32), mp)
    assert np.array_equal(device_to_numpy(out.compressed.minorIdx, n, np.uint32), mi)
    assert np.array_equal(device_to_numpy(out.compressed.val, n, np.uint16), mv)
    csr_rt.clear(rws)


@pytest.mark.parametrize("many_to_one", [False, True])
def test_reoptimize_gather_and_fallback(gpu, many_to_one):
    """The one-read gather (16-byte chunks) against the oracle, and the
    count + scatter fallback when a row keeps more entries than its region
    holds (a feature remap onto 5 output columns), on one workspace."""
    rows, cols, bs = 300, 520, 16
    ptr_, idx, val = random_csr(rows, cols, 0.15, seed=7, zeros=0.1, empty_rows=0.05)
    bits = f16_bits(val)
    bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr_, idx, bits, bs)
    bell = DeviceBlockedELL.from_numpy(rows, cols, nnz, bs, ell, bci, bv)
    rk = np.ones(rows, np.uint8)
    fk = np.ones(cols, np.uint8)
    rr = np.arange(rows, dtype=np.uint32)
    out_cols = 5 if many_to_one else cols
    fr = (np.arange(cols) % out_cols).astype(np.uint32)
    d_rk, d_fk, d_rr, d_fr = to_dev_u8(rk), to_dev_u8(fk), to_dev_u32(rr), to_dev_u32(fr)
    rws = A.masked_sparse_reoptimize_workspace()
    csr_rt.init(rws)
    assert csr_rt.setup(rws, 0) == 1
    masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk), row_remap=ptr(d_rr), feature_remap=ptr(d_fr))
    for _ in range(2):
        cfg = A.masked_sparse_reoptimize_config(rows, out_cols, 0, bs, A.LAYOUT_COMPRESSED)
        out = A.masked_sparse_reoptimize_result()
        assert csr_rt.manual_reoptimize_masked_sparse(bell.view(), masks, cfg, rws, out) == 1
        mp, mi, mv, n = O.reoptimize(rows, cols, bs, ell, bci, np.asarray(bv).view(np.uint16), rk, fk, rr, fr, rows, out_cols)
        if many_to_one:
            assert int(np.diff(mp).max()) > out_cols
        assert out.metadata.live_nnz == n
        assert np.array_equal(device_to_numpy(out.compressed.majorPtr, rows + 1, np.uint32), mp)
        assert np.array_equal(device_to_numpy(out.compressed.minorIdx, n, np.uint32), mi)
        assert np.array_equal(device_to_numpy(out.compressed.val, n, np.uint16), mv)
    csr_rt.clear(rws)


def _fast_csr(rows, cols, mean_nnz, seed, vt, long_rows=0, empty=0.05):
    """Vectorised random CSR (columns may repeat inside a row, as stored
    entries may): integer counts, some explicit zeros and empty rows."""
    rng = np.random.default_rng(seed)
    n = rng.poisson(mean_nnz, rows).astype(np.int64)
    n[rng.random(rows) < empty] = 0
    if long_rows:
        n[rng.choice(rows, long_rows, replace=False)] = rng.integers(3000, 9000, long_rows)
    ptr_ = np.concatenate([[0], np.cumsum(n)]).astype(np.uint32)
    idx = rng.integers(0, cols, int(n.sum())).astype(np.uint32)
    vals = rng.choice([0.0, 1.0, 2.0, 3.0, 4.0, 6.0, 8.0], size=len(idx), p=[0.03, 0.37, 0.2, 0.15, 0.1, 0.1, 0.05])
    return ptr_, idx, encode(vals, vt)


def _sell_ragged(rows, cols, ptr_, idx, bits, seed):
    """Sliced-ELL with slices of random heights (no uniform height)."""
    rng = np.random.default_rng(seed)
    cuts = np.unique(np.concatenate([[0], rng.integers(1, rows, max(rows // 7, 1)), [rows]])).astype(np.uint32)
    rn = np.diff(ptr_.astype(np.int64))
    widths = np.array([rn[a:b].max(initial=0) for a, b in zip(cuts[:-1], cuts[1:])], np.uint32)
    heights = np.diff(cuts.astype(np.int64))
    slots = np.concatenate([[0], np.cumsum(heights * widths)]).astype(np.int64)
    col = np.full(int(slots[-1]), 0xFFFFFFFF, np.uint32)
    val = np.zeros(int(slots[-1]), bits.dtype)
    for s_, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
        for r in range(a, b):
            base = slots[s_] + (r - a) * widths[s_]
            k = int(rn[r])
            col[base:base + k] = idx[ptr_[r]:ptr_[r + 1]]
            val[base:base + k] = bits[ptr_[r]:ptr_[r + 1]]
    return cuts, widths, slots[:-1].astype(np.uint32), col, val


STREAM_CASES = {
    # name: rows, cols, mean nnz, long rows, G, vt, feature keep p (None: no mask), row keep p (None: no mask)
    "windows": (600_000, 500, 9, 0, 32, F16, 0.6, 0.8),
    "long_rows": (300, 20_000, 40, 40, 7, BF16, 0.3, 0.9),
    "tail_fp8": (1001, 300, 5, 0, 1, FP8, 0.7, 0.7),
    "global_masks": (3000, 40_000, 60, 3, 5, F32, None, 0.9),
    "no_groups": (5000, 1000, 30, 2, 0, F16, 0.5, None),
    "all_masked": (700, 100, 10, 0, 3, F16, 0.5, 0.0),
}


@pytest.mark.parametrize("case", sorted(STREAM_CASES))
def test_streaming_group_reduce_parity(ws, case):
    """The streaming path (LDS row windows, two-buffer chunk stream, group
    queue): many windows per workgroup, rows over several chunks, rows in
    the array's last partial unit, masks of every group count (1 to 32
    lanes per queued entry), kept-feature tables too large for LDS, no
    groups, no masks, every row masked; compressed rows and Sliced-ELL with
    uniform and ragged slices, against the oracle."""
    rows, cols, mean, longr, G, vt, fkp, rkp = STREAM_CASES[case]
    ptr_, idx, bits = _fast_csr(rows, cols, mean, seed=hash(case) % 1000, vt=vt, long_rows=longr)
    if len(idx) % 4 == 0:  # make the array end inside a unit (rows read entry by entry)
        ptr_[-1] += 1
        idx = np.append(idx, np.uint32(1)).astype(np.uint32)
        bits = np.append(bits, encode(np.ones(1), vt)).astype(bits.dtype)
    rows = len(ptr_) - 1
    rng = np.random.default_rng(7)
    rk = None if rkp is None else (rng.random(rows) < rkp).astype(np.uint8)
    fk = None if fkp is None else (rng.random(cols) < fkp).astype(np.uint8)
    fgm = rng.integers(0, 2 ** 32, size=cols, dtype=np.uint64).astype(np.uint32)
    if case == "long_rows":
        fgm &= np.uint32(0x41)  # sparse masks: at most two groups per feature
    filt_t = [5.0, 2, 0.5, 0]
    d_rk = to_dev_u8(rk) if rk is not None else None
    d_fk = to_dev_u8(fk) if fk is not None else None
    d_fgm = to_dev_u32(fgm)
    groups = A.group_mask_config_view(group_count=G, feature_group_masks=ptr(d_fgm) if G else None)
    masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk))
    filt = A.sparse_group_filter_params(*filt_t)
    res = A.sparse_group_reduce_result()
    src = DeviceCSR.from_numpy(rows, cols, ptr_, idx, bits, vt)
    assert csr_rt.compute_sparse_group_reduce_compressed_fallback(src.view(), ws, groups, masks, filt, res, vt) == 1
    ref = O.group_reduce_csr(rows, cols, ptr_, idx, bits, vt, rk, fk, fgm if G else None, G, filt_t)
    _compare(fetch(ws, res), ref, G, vt)
    if rows > 100_000:
        return
    for ragged in (False, True):
        if ragged:
            offs, wid, slot, col, sval = _sell_ragged(rows, cols, ptr_, idx, bits, seed=rows)
            sell = DeviceSlicedELL.from_numpy(rows, cols, len(idx), 0, offs, wid, slot, col, sval, vt)
            sref = O.group_reduce_sell(rows, cols, 0, offs, wid, slot, col, sval, vt, rk, fk, fgm if G else None, G, filt_t)
        else:
            s = O.sliced_ell_from_compressed(rows, cols, 0, ptr_, idx, bits, 32)
            sell = DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                              s["col_idx"], s["val"], vt)
            sref = O.group_reduce_sell(rows, cols, s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"], s["col_idx"],
                                       s["val"], vt, rk, fk, fgm if G else None, G, filt_t)
        assert csr_rt.compute_sparse_group_reduce(sell.view(), ws, groups, masks, filt, res, vt) == 1
        _compare(fetch(ws, res), sref, G, vt)
