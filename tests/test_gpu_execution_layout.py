"""Bucketed execution partitions on the GPU.

Blocked-ELL: the layout (stable width order with the partial last row block
in place, equal-count cuts, collapse, byte fallback, 1..8 choice) is a
closed form, so order, segments, widths, bytes and filled segments are
compared bit for bit with the numpy restatement.  Sliced-ELL: the split is
checked for optimality against a brute-force DP on its pre-merge cost
(exactly k segments), then the merge / fallback / byte rules, the maps and
the filled segments against closed forms on the split the library chose.
"""
import numpy as np
import pytest

from cellshard_amd.device import DeviceBlockedELL, DeviceSlicedELL
from cellshard_amd.ops import (ConvertWorkspace, build_bucketed_blocked_ell, build_bucketed_sliced_ell,
                               choose_bucket_count_blocked_ell, choose_bucket_count_sliced_ell)
from helpers import F16, encode
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture()
def ws(gpu):
    w = ConvertWorkspace()
    yield w
    w.close()


def _bell(rows, cols, bs, seed, skew=True):
    rng = np.random.default_rng(seed)
    ptr = [0]
    idx = []
    for r in range(rows):  # row lengths vary a lot between row blocks
        k = int(rng.integers(0, cols // (2 if (r // bs) % 3 == 0 or not skew else 12)))
        idx.append(np.sort(rng.choice(cols, k, replace=False)))
        ptr.append(ptr[-1] + k)
    ptr = np.asarray(ptr, np.uint32)
    idx = np.concatenate(idx).astype(np.uint32)
    bits = encode(rng.integers(0, 9, size=len(idx)).astype(np.float64), F16)
    bci, val, ell_cols, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    return DeviceBlockedELL.from_numpy(rows, cols, nnz, bs, ell_cols, bci, val), bci, val, ell_cols


@pytest.mark.parametrize("rows,bs,req", [(203, 8, 3), (256, 16, 4), (97, 16, 8), (40, 8, 1), (16, 16, 5)])
def test_bucketed_blocked_ell_matches_closed_form(ws, rows, bs, req):
    cols = 300
    dev, bci, val, ell_cols = _bell(rows, cols, bs, seed=rows + bs)
    got = build_bucketed_blocked_ell(dev, req, ws)
    lay = O.bell_execution_layout(rows, bs, ell_cols, bci, req)
    assert got["bytes"] == lay["bytes"]
    assert len(got["segments"]) == len(lay["width"])
    segs, e2c = O.bell_fill_reference(rows, bs, ell_cols, bci, val, lay)
    assert np.array_equal(got["exec_to_canonical"], e2c)
    c2e = np.empty(rows, np.uint32)
    c2e[e2c] = np.arange(rows, dtype=np.uint32)
    assert np.array_equal(got["canonical_to_exec"], c2e)
    assert np.array_equal(got["col_maps"][0], np.arange(cols)) and np.array_equal(got["col_maps"][1], np.arange(cols))
    for g, (seg_rows, sb, sv) in zip(got["segments"], segs):
        assert g["rows"] == seg_rows
        assert np.array_equal(g["block_col_idx"], sb)
        assert np.array_equal(g["val"], sv.view(np.uint16))
    n, b = choose_bucket_count_blocked_ell(dev, ws)
    assert (n, b) == O.bell_choose_bucket_count(rows, bs, ell_cols, bci)


def _sell(rows, cols, S, seed):
    rng = np.random.default_rng(seed)
    lens = np.where(rng.random(rows) < 0.7, rng.integers(0, 6, rows), rng.integers(20, 60, rows))
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    idx = np.concatenate([np.sort(rng.choice(cols, int(n), replace=False)) for n in lens]).astype(np.uint32)
    bits = encode(rng.integers(1, 9, size=len(idx)).astype(np.float64), F16)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, S)
    dev = DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                     s["col_idx"], s["val"])
    return dev, s, ptr, idx, bits


@pytest.mark.parametrize("rows,S,req", [(150, 32, 3), (64, 8, 8), (33, 32, 2), (1, 32, 4)])
def test_bucketed_sliced_ell_optimal_and_filled(ws, rows, S, req):
    cols = 200
    dev, s, ptr, idx, bits = _sell(rows, cols, S, seed=rows * 7 + S)
    got = build_bucketed_sliced_ell(dev, req, ws)
    rw = O.sell_row_widths(rows, s["offsets"], s["widths"], s["slot_offsets"], s["col_idx"])
    assert np.array_equal(rw, np.diff(ptr.astype(np.int64)))
    order = np.argsort(rw, kind="stable")
    assert np.array_equal(got["exec_to_canonical"], order)  # stable width order
    sw = rw[order]
    off = got["segment_row_offsets"]
    widths = [g["width"] for g in got["segments"]]
    K = max(1, min(req, rows))
    original = O.sell_packed_bytes(len(s["widths"]), int(np.sum(np.diff(s["offsets"].astype(np.int64)) * s["widths"])))
    for a, b_, w in zip(off[:-1], off[1:], widths):
        assert b_ > a and w == sw[b_ - 1]  # each segment's width is its widest (last) row
    assert all(widths[i] != widths[i + 1] for i in range(len(widths) - 1))  # equal neighbours merged
    split_bytes = sum(O.sell_packed_bytes(1, (b_ - a) * w) for a, b_, w in zip(off[:-1], off[1:], widths))
    if len(widths) == 1 and got["bytes"] == original and K > 1:
        pass  # fell back to one segment: its split did not shrink the bytes
    else:
        assert got["bytes"] == split_bytes
        # pre-merge cost = 29 per original segment + 6 x slots; optimal among exactly K segments
        pre = 29 * K + 6 * sum((b_ - a) * w for a, b_, w in zip(off[:-1], off[1:], widths))
        assert pre == O.sell_min_split_cost(sw, K)
    # filled rows: valid slots of the source row, in order, padded
    for g, a in zip(got["segments"], off[:-1]):
        w = g["width"]
        ci = g["col_idx"].reshape(-1, w) if w else np.zeros((g["rows"], 0), np.uint32)
        cv = g["val"].reshape(-1, w) if w else np.zeros((g["rows"], 0), np.uint16)
        for dr in range(g["rows"]):
            r = int(order[a + dr])
            n = int(rw[r])
            assert np.array_equal(ci[dr, :n], idx[ptr[r]:ptr[r + 1]])
            assert np.array_equal(cv[dr, :n], bits[ptr[r]:ptr[r + 1]].view(np.uint16))
            assert np.all(ci[dr, n:] == O.INVALID) and np.all(cv[dr, n:] == 0)
    c2e = np.empty(rows, np.uint32)
    c2e[order] = np.arange(rows, dtype=np.uint32)
    assert np.array_equal(got["canonical_to_exec"], c2e)
    # choice over 1..8 buckets: fewest bytes, ties to fewer buckets
    per = [build_bucketed_sliced_ell(dev, b, ws)["bytes"] for b in range(1, max(1, min(8, rows)) + 1)]
    n, b = choose_bucket_count_sliced_ell(dev, ws)
    assert b == min(per) and n == 1 + per.index(min(per))
