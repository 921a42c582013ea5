"""Masked Blocked-ELL -> CSR (filter_blocked_ell_to_compressed) and its
bucketed variant on the GPU against the oracle restatement."""
import numpy as np
import pytest
import torch

from cellshard_amd import CellShardError
from cellshard_amd.device import DeviceCSR
from cellshard_amd.ops import (ConvertWorkspace, FilteredBlockedEllWorkspace, blocked_ell_from_compressed, device_to_numpy,
                               filter_blocked_ell_to_compressed)
from helpers import F16, F32, FP8, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
VB = {F16: 2, F32: 4, FP8: 1}
NP_BITS = {F16: np.uint16, F32: np.uint32, FP8: np.uint8}


@pytest.fixture(scope="module")
def ws():
    return ConvertWorkspace()


@pytest.fixture(scope="module")
def fws():
    w = FilteredBlockedEllWorkspace()
    yield w
    w.close()


def dev_u8(a):
    return torch.as_tensor(np.asarray(a, np.uint8), device="cuda")


def dev_u32(a):
    return torch.as_tensor(np.asarray(a, np.uint32).view(np.int32), device="cuda")


def make_bell(rows, cols, bs, vt, seed, ws, zeros=0.1):
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=seed, zeros=zeros, empty_rows=0.05, fp8_safe=(vt == FP8))
    bits = encode(val, vt)
    bell = blocked_ell_from_compressed(DeviceCSR.from_numpy(rows, cols, ptr, idx, bits, vt), bs, ws)
    gb, gv = bell.to_numpy()
    return bell, gb, np.ascontiguousarray(gv).view(NP_BITS[vt])


def fetch_csr(view, vt):
    torch.cuda.synchronize()
    p = device_to_numpy(view.majorPtr, view.rows + 1, np.uint32)
    n = int(view.nnz)
    return p, device_to_numpy(view.minorIdx, n, np.uint32), device_to_numpy(view.val, n, NP_BITS[vt])


def masks(rows, cols, seed, drop_rows=3, drop_cols=2):
    rng = np.random.default_rng(seed)
    rk = (rng.random(rows) < 0.8).astype(np.uint8)
    fk = (rng.random(cols) < 0.6).astype(np.uint8)
    rr = np.full(rows, 0xFFFFFFFF, np.uint32)
    rr[rk != 0] = rng.permutation(int(rk.sum()))
    cr = np.full(cols, 0xFFFFFFFF, np.uint32)
    cr[fk != 0] = np.arange(int(fk.sum()))
    return rk, fk, rr, cr, int(rk.sum()) - drop_rows, int(fk.sum()) - drop_cols


@pytest.mark.parametrize("vt", [F16, F32, FP8])
@pytest.mark.parametrize("bs", [4, 16, 32])
def test_filter_matches_oracle(ws, fws, vt, bs):
    rows, cols = 701, 1300
    bell, gb, gv = make_bell(rows, cols, bs, vt, 30 + bs + vt, ws)
    rk, fk, rr, cr, orows, ocols = masks(rows, cols, bs + vt)
    out = filter_blocked_ell_to_compressed(bell, orows, ocols, dev_u8(rk), dev_u8(fk), dev_u32(rr), dev_u32(cr), ws=fws)
    p, i, v = fetch_csr(out.filtered, vt)
    rp, ri, rv, st = O.filter_blocked_ell(rows, cols, bs, bell.ell_cols, gb, gv, vt, orows, ocols, rk, fk, rr, cr)
    assert np.array_equal(p, rp) and np.array_equal(i, ri) and np.array_equal(v, rv.view(NP_BITS[vt]))
    s = out.stats
    assert (s.kept_rows, s.output_rows, s.output_cols, s.live_nnz) == (orows, orows, ocols, st["live_nnz"])
    assert (s.kept_row_slots, s.dead_slots) == (st["kept_row_slots"], st["dead_slots"])
    assert abs(s.live_fill_ratio - st["live_fill_ratio"]) < 1e-12 and s.dead_value_bytes == st["dead_slots"] * VB[vt]
    assert out.has_bucketed == 0


def test_filter_identity_and_empty_outputs(ws, fws):
    rows, cols, bs = 300, 500, 8
    bell, gb, gv = make_bell(rows, cols, bs, F16, 77, ws, zeros=0.3)
    out = filter_blocked_ell_to_compressed(bell, rows, cols, ws=fws)  # no masks, no remaps
    p, i, v = fetch_csr(out.filtered, F16)
    rp, ri, rv, st = O.filter_blocked_ell(rows, cols, bs, bell.ell_cols, gb, gv, F16, rows, cols)
    assert np.array_equal(p, rp) and np.array_equal(i, ri) and np.array_equal(v, rv)
    assert int(out.filtered.nnz) == bell.nnz  # explicit zeros dropped, every nonzero kept
    for orows, ocols in ((0, cols), (rows, 0)):
        out = filter_blocked_ell_to_compressed(bell, orows, ocols, ws=fws)
        assert int(out.filtered.nnz) == 0 and out.stats.live_fill_ratio == 1.0
        if orows:
            assert np.all(fetch_csr(out.filtered, F16)[0] == 0)


@pytest.mark.parametrize("vt", [F16, F32])
def test_filter_many_to_one_column_remap(ws, fws, vt):
    # every column onto one of 7 output columns: a row keeps more entries than
    # there are output columns, so the one-read gather overflows its per-row
    # region and the call counts and emits instead; the result is the same
    rows, cols, bs = 300, 512, 8
    bell, gb, gv = make_bell(rows, cols, bs, vt, 12 + vt, ws)
    cr = (np.arange(cols) % 7).astype(np.uint32)
    out = filter_blocked_ell_to_compressed(bell, rows, 7, col_remap=dev_u32(cr), ws=fws)
    p, i, v = fetch_csr(out.filtered, vt)
    rp, ri, rv, st = O.filter_blocked_ell(rows, cols, bs, bell.ell_cols, gb, gv, vt, rows, 7, None, None, None, cr)
    assert int(np.diff(rp).max()) > 7
    assert np.array_equal(p, rp) and np.array_equal(i, ri) and np.array_equal(v, rv.view(NP_BITS[vt]))
    assert out.stats.live_nnz == st["live_nnz"]
    # then a call that fits its regions again, on the same workspace
    rk, fk, rr, cr2, orows, ocols = masks(rows, cols, 4)
    out = filter_blocked_ell_to_compressed(bell, orows, ocols, dev_u8(rk), dev_u8(fk), dev_u32(rr), dev_u32(cr2), ws=fws)
    p, i, v = fetch_csr(out.filtered, vt)
    rp, ri, rv, st = O.filter_blocked_ell(rows, cols, bs, bell.ell_cols, gb, gv, vt, orows, ocols, rk, fk, rr, cr2)
    assert np.array_equal(p, rp) and np.array_equal(i, ri) and np.array_equal(v, rv.view(NP_BITS[vt]))
    assert (out.stats.kept_rows, out.stats.live_nnz) == (orows, st["live_nnz"])


def test_filter_rejects_two_rows_on_one_output_row(ws, fws):
    rows, cols, bs = 40, 60, 4
    bell, gb, gv = make_bell(rows, cols, bs, F16, 5, ws)
    rr = np.arange(rows, dtype=np.uint32)
    rr[7] = 3
    with pytest.raises(CellShardError, match="output row 3"):
        filter_blocked_ell_to_compressed(bell, rows, cols, row_remap=dev_u32(rr), ws=fws)


@pytest.mark.parametrize("buckets", [1, 5, 64])
def test_bucketed_filtered_view(ws, fws, buckets):
    rows, cols, bs = 512, 900, 16
    bell, gb, gv = make_bell(rows, cols, bs, F16, 91, ws)
    rk, fk, rr, cr, orows, ocols = masks(rows, cols, 3, drop_rows=0, drop_cols=0)
    out = filter_blocked_ell_to_compressed(bell, orows, ocols, dev_u8(rk), dev_u8(fk), dev_u32(rr), dev_u32(cr), ws=fws,
                                           bucket_count=buckets)
    assert out.has_bucketed == 1
    rp, ri, rv, st = O.filter_blocked_ell(rows, cols, bs, bell.ell_cols, gb, gv, F16, orows, ocols, rk, fk, rr, cr)
    ref = O.bucket_plan(rp, orows, buckets)
    pl = out.bucketed.plan
    torch.cuda.synchronize()
    assert (pl.major_dim, pl.bucket_count) == (orows, ref["bucket_count"])
    order = device_to_numpy(pl.major_order, orows, np.uint32)
    assert np.array_equal(order, ref["major_order"])
    assert np.array_equal(device_to_numpy(pl.bucket_offsets, pl.bucket_count + 1, np.uint32), ref["bucket_offsets"])
    assert np.array_equal(device_to_numpy(pl.major_nnz_sorted, orows, np.uint32), ref["major_nnz_sorted"])
    inv = device_to_numpy(pl.inverse_major_order, orows, np.uint32)
    assert np.array_equal(inv[order], np.arange(orows))
    bp, bi, bv = fetch_csr(out.bucketed.rebuilt, F16)
    for k, r in enumerate(order):  # rebuilt row k = filtered row order[k]
        assert np.array_equal(bi[bp[k]:bp[k + 1]], ri[rp[r]:rp[r + 1]])
        assert np.array_equal(bv[bp[k]:bp[k + 1]], rv[rp[r]:rp[r + 1]])


@pytest.fixture(scope="module")
def rws():
    from cellshard_amd.ops import RepackWorkspace
    w = RepackWorkspace()
    yield w
    w.close()


def shard_parts(ws, vt, specs, seed):
    parts, oparts = [], []
    for k, (rows, bs) in enumerate(specs):
        bell, gb, gv = make_bell(rows, 1100, bs, vt, seed + k, ws)
        parts.append(bell)
        oparts.append({"rows": rows, "cols": 1100, "block_size": bs, "ell_cols": bell.ell_cols, "bci": gb, "val": gv})
    return parts, oparts


@pytest.mark.parametrize("vt", [F16, F32])
@pytest.mark.parametrize("remap", [True, False])
def test_repack_shard_matches_oracle(ws, rws, vt, remap):
    from cellshard_amd.ops import build_repacked_shard_blocked_ell
    parts, oparts = shard_parts(ws, vt, [(300, 8), (257, 16), (411, 32)], 60 + vt)
    rows = sum(p.rows for p in parts)
    if remap:
        rk, fk, rr, cr, orows, ocols = masks(rows, 1100, 9)
        args = (dev_u8(rk), dev_u8(fk), dev_u32(rr), dev_u32(cr))
        oargs = dict(keep_rows=rk, keep_cols=fk, row_remap=rr, col_remap=cr)
    else:
        orows, ocols, args, oargs = rows, 1100, (None, None, None, None), {}
    out = build_repacked_shard_blocked_ell(parts, orows, ocols, 16, 6, *args, shard_id=3, ws=rws)
    ref = O.repack_blocked_ell(oparts, vt, orows, ocols, 16, 6, **oargs)
    torch.cuda.synchronize()
    p, i, v = fetch_csr(out.filtered, vt)
    assert np.array_equal(p, ref["ptr"]) and np.array_equal(i, ref["idx"]) and np.array_equal(v, ref["val"].view(NP_BITS[vt]))
    s, rs = out.stats, ref["stats"]
    assert (s.kept_rows, s.output_rows, s.output_cols, s.live_nnz, s.kept_row_slots, s.dead_slots, s.ell_width) == (
        rs["kept_rows"], rs["output_rows"], rs["output_cols"], rs["live_nnz"], rs["kept_row_slots"], rs["dead_slots"], rs["ell_width"])
    assert abs(s.live_fill_ratio - rs["live_fill_ratio"]) < 1e-12 and out.shard_id == 3
    pl = out.bucket_plan
    assert pl.bucket_count == ref["plan"]["bucket_count"]
    assert np.array_equal(device_to_numpy(pl.major_order, orows, np.uint32), ref["plan"]["major_order"])
    bp, bi, bv = fetch_csr(out.bucketed.rebuilt, vt)
    assert np.array_equal(bp, ref["rptr"]) and np.array_equal(bi, ref["ridx"])
    b = out.rebuilt
    assert (b.rows, b.cols, b.block_size, b.ell_cols, b.nnz) == (orows, ocols, 16, ref["bell_ell_cols"], rs["live_nnz"])
    rb = (orows + 15) // 16
    w = b.ell_cols // 16
    assert np.array_equal(device_to_numpy(b.blockColIdx, rb * w, np.uint32), ref["bell_bci"])
    assert np.array_equal(device_to_numpy(b.val, orows * b.ell_cols, NP_BITS[vt]), ref["bell_val"].view(NP_BITS[vt]))


def test_repack_nothing_survives(ws, rws):
    from cellshard_amd.ops import build_repacked_shard_blocked_ell
    parts, oparts = shard_parts(ws, F16, [(100, 8), (60, 8)], 5)
    rk = np.zeros(160, np.uint8)
    out = build_repacked_shard_blocked_ell(parts, 50, 1100, 8, 4, keep_rows=dev_u8(rk), ws=rws)
    assert out.stats.kept_rows == 0 and out.stats.live_nnz == 0 and out.rebuilt.ell_cols == 0
    assert out.bucket_plan.bucket_count == 1 and out.bucketed.rebuilt.majorPtr == out.filtered.majorPtr
