"""Conversions, transpose and bucketing on the GPU: bit-exact vs the oracle
(and the reference converter's golden vectors for Blocked-ELL)."""
from pathlib import Path

import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceBlockedELL, DeviceCSR, u32_np
from cellshard_amd.ops import (ConvertWorkspace, TransposeScratch, blocked_ell_bucket_plan, blocked_ell_from_compressed,
                               build_major_nnz_bucket_plan, choose_blocked_ell_block_size, choose_sliced_ell_slice_rows,
                               compressed_transpose, log2_bucket_offsets, rebuild_compressed_major_order,
                               sliced_ell_from_blocked_ell, sliced_ell_from_compressed, synth_csr)
from helpers import F16, F32, FP8, encode, f16_bits, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture()
def ws(gpu):
    w = ConvertWorkspace()
    yield w
    w.close()


def test_blocked_ell_matches_reference_golden(ws):
    g = np.load(GOLD / "bell_ref_vectors.npz")
    names = sorted({k.split("/")[0] for k in g.files})
    for name in names:
        rows, cols = (int(x) for x in g[f"{name}/meta"])
        src = DeviceCSR.from_numpy(rows, cols, g[f"{name}/ptr"], g[f"{name}/idx"], g[f"{name}/val"])
        for bs in (1, 2, 3, 4, 8, 16, 32):
            out = blocked_ell_from_compressed(src, bs, ws)
            ell, nnz = (int(x) for x in g[f"{name}/b{bs}/shape"])
            assert (out.ell_cols, out.nnz) == (ell, nnz), (name, bs)
            bci, val = out.to_numpy()
            assert np.array_equal(bci, g[f"{name}/b{bs}/bci"]), (name, bs)
            assert np.array_equal(val, g[f"{name}/b{bs}/val"]), (name, bs)
        for k in range(3):
            cands = [int(c) for c in g[f"{name}/choose{k}/cands"]]
            bs, fill, pad = choose_blocked_ell_block_size(src, cands, ws)
            rbs, rpad = (int(x) for x in g[f"{name}/choose{k}/result"])
            assert (bs, pad) == (rbs, rpad), (name, k)
            assert abs(fill - float(g[f"{name}/choose{k}/fill"][0])) < 1e-12


@pytest.mark.parametrize("vt", [F16, F32, FP8])
def test_blocked_ell_random_parity(ws, vt):
    rows, cols = 513, 1900
    ptr, idx, val = random_csr(rows, cols, 0.02, seed=5 + vt, zeros=0.03, empty_rows=0.1, fp8_safe=(vt == FP8))
    bits = encode(val, vt)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits, vt)
    for bs in (8, 16, 32):
        out = blocked_ell_from_compressed(src, bs, ws)
        bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
        gb, gv = out.to_numpy()
        assert (out.ell_cols, out.nnz) == (ell, nnz)
        assert np.array_equal(gb, bci) and np.array_equal(gv.view(v.dtype), v)


@pytest.mark.parametrize("extra", [-1, 0, 3])
def test_blocked_ell_given_width(ws, extra):
    # a caller-chosen ell width: wider than needed pads with invalid slots and
    # zero values; narrower fails with the width the rows need
    import ctypes as C
    from cellshard_amd import CellShardError, load
    from cellshard_amd.device import DeviceBlockedELL
    from cellshard_amd.ops import blocked_ell_shape
    rows, cols, bs = 300, 900, 8
    ptr, idx, val = random_csr(rows, cols, 0.03, seed=11, empty_rows=0.05)
    bits = f16_bits(val)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    need, _, _ = blocked_ell_shape(src, bs, ws)
    width = need + extra
    rb = (rows + bs - 1) // bs
    out = DeviceBlockedELL(rows, cols, 0, bs, width * bs, torch.empty(rb * width, dtype=torch.int32, device="cuda"),
                           torch.empty(rows * width * bs, dtype=torch.int16, device="cuda"), F16)
    v, sv = out.view(), src.view()
    ok = load().csamd_blocked_ell_from_compressed(C.byref(sv), F16, C.byref(v), C.byref(ws.ws))
    if extra < 0:
        assert ok == 0
        with pytest.raises(CellShardError, match="required"):
            from cellshard_amd._lib import check
            check(ok, "blocked_ell_from_compressed")
        return
    assert ok == 1
    bci, ref_v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    gb = out.block_col_idx.cpu().numpy().view(np.uint32).reshape(rb, width)
    gv = out.val.cpu().numpy().view(np.uint16).reshape(rows, width * bs)
    assert int(v.nnz) == nnz and ell == need * bs
    assert np.array_equal(gb[:, :need].ravel(), bci) and np.all(gb[:, need:] == 0xFFFFFFFF)
    assert np.array_equal(gv[:, :need * bs].ravel(), ref_v.view(np.uint16)) and np.all(gv[:, need * bs:] == 0)


def test_blocked_ell_unsorted_duplicates(ws):
    rows, cols = 90, 70
    ptr, idx, val = random_csr(rows, cols, 0.1, seed=2, sorted_rows=False, duplicates=0.2, zeros=0.2)
    bits = f16_bits(val)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    for bs in (2, 4, 16):
        out = blocked_ell_from_compressed(src, bs, ws)
        bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
        gb, gv = out.to_numpy()
        assert (out.ell_cols, out.nnz) == (ell, nnz)
        assert np.array_equal(gb, bci) and np.array_equal(gv, v)


@pytest.mark.parametrize("S", [1, 32, 64, 100])
def test_sliced_ell_parity(ws, S):
    rows, cols = 1000, 3000
    ptr, idx, val = random_csr(rows, cols, 0.01, seed=S, empty_rows=0.1)
    bits = f16_bits(val)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    out = sliced_ell_from_compressed(src, S, ws)
    ref = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, S)
    offs, widths, slots, col, v = out.to_numpy()
    assert out.slice_rows == ref["slice_rows"] and out.nnz == len(idx)
    assert np.array_equal(offs, ref["offsets"]) and np.array_equal(widths, ref["widths"])
    assert np.array_equal(slots, ref["slot_offsets"])
    assert np.array_equal(col, ref["col_idx"]) and np.array_equal(v, ref["val"])


def test_sliced_ell_tuner(ws):
    rows, cols = 800, 500
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=3, empty_rows=0.3)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, f16_bits(val))
    cands = [8, 16, 32, 64]
    got = choose_sliced_ell_slice_rows(src, cands, ws)
    ref = O.choose_sliced_ell_slice_rows(np.diff(ptr), rows, len(idx), cands)
    assert got["slice_rows"] == ref["slice_rows"] and got["padded_slots"] == ref["padded_slots"]
    assert abs(got["score"] - ref["score"]) < 1e-12


def test_sliced_from_blocked_parity(ws):
    rows, cols = 333, 400
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=4, zeros=0.1)
    bits = f16_bits(val)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    bell = blocked_ell_from_compressed(src, 8, ws)
    out = sliced_ell_from_blocked_ell(bell, 32, ws)
    gb, gv = bell.to_numpy()
    ref = O.sliced_ell_from_blocked_ell(rows, cols, 8, bell.ell_cols, gb, gv, 32)
    offs, widths, slots, col, v = out.to_numpy()
    assert np.array_equal(widths, ref["widths"]) and np.array_equal(slots, ref["slot_offsets"])
    assert np.array_equal(col, ref["col_idx"]) and np.array_equal(v, ref["val"])


@pytest.mark.parametrize("case", ["sorted", "unsorted", "empty_cols", "wide", "nnz0", "dense", "duplicates", "tall"])
def test_transpose_parity(gpu, case):
    if case == "dense":
        # more entries per (256-row tile, 256-column band) than one LDS batch
        # holds: chunks staged in several row batches; partial last tile and band
        rows, cols = 700, 1100
        ptr, idx, val = random_csr(rows, cols, 0.85, seed=3, empty_rows=0.05)
    elif case == "tall":
        # 36 tiles x 12 bands: several complete groups of tiles in the chunk
        # order plus a partial one, spread over all eight XCD ranges
        rows, cols = 9000, 3000
        ptr, idx, val = random_csr(rows, cols, 0.02, seed=5, empty_rows=0.05)
    elif case == "duplicates":  # repeated columns inside rows: the fallback scatter
        rows, cols = 400, 300
        ptr, idx, val = random_csr(rows, cols, 0.05, seed=9, duplicates=0.3)
        val = np.ones_like(val)  # equal values: the order of repeats inside a row is not fixed
    elif case == "nnz0":
        rows, cols = 10, 6
        ptr, idx, val = np.zeros(11, np.uint32), np.zeros(0, np.uint32), np.zeros(0)
    elif case == "wide":
        rows, cols = 3000, 50000  # > LDS histogram capacity: global-count path
        ptr, idx, val = random_csr(rows, cols, 0.001, seed=1)
    else:
        rows, cols = 2500, 9000 if case != "empty_cols" else 40000
        ptr, idx, val = random_csr(rows, cols, 0.004, seed=7, sorted_rows=(case != "unsorted"), empty_rows=0.1)
    bits = f16_bits(val)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    out = compressed_transpose(src, TransposeScratch())
    optr, oidx, oval = out.to_numpy()
    rptr, ridx, rval = O.transpose(rows, cols, ptr, idx, bits)
    assert np.array_equal(optr, rptr) and np.array_equal(oidx, ridx) and np.array_equal(oval, rval)
    # round trip back to CSR
    back = compressed_transpose(out, TransposeScratch())
    bptr, bidx, bval = back.to_numpy()
    assert np.array_equal(bptr, ptr)
    if case != "unsorted":
        assert np.array_equal(bidx, idx) and np.array_equal(bval, bits)


@pytest.mark.parametrize("tail", [0, 1, 2, 3])
def test_transpose_group_tails(gpu, tail):
    # the count pass reads indices in 16-byte groups: tile edges inside a
    # group, and an array whose last group is partial (nnz % 4 = tail)
    rows, cols = 777, 700
    ptr, idx, val = random_csr(rows, cols, 0.03, seed=40 + tail, empty_rows=0.05)
    cut = (len(idx) - tail) % 4
    keep = len(idx) - cut
    ptr = np.minimum(ptr, keep).astype(np.uint32)
    idx, val = idx[:keep], val[:keep]
    assert len(idx) % 4 == tail
    bits = f16_bits(val)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    optr, oidx, oval = compressed_transpose(src, TransposeScratch()).to_numpy()
    rptr, ridx, rval = O.transpose(rows, cols, ptr, idx, bits)
    assert np.array_equal(optr, rptr) and np.array_equal(oidx, ridx) and np.array_equal(oval, rval)


@pytest.mark.parametrize("vt", [F32, FP8])
def test_transpose_value_types(gpu, vt):
    rows, cols = 1300, 2100
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=21, fp8_safe=True)
    bits = encode(val, vt)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits, vt)
    optr, oidx, oval = compressed_transpose(src, TransposeScratch()).to_numpy()
    rptr, ridx, rval = O.transpose(rows, cols, ptr, idx, bits)
    assert np.array_equal(optr, rptr) and np.array_equal(oidx, ridx)
    assert np.array_equal(oval.view(np.uint8), np.ascontiguousarray(rval).view(np.uint8))


def test_bucket_plan_parity(gpu):
    rows, cols = 5000, 2000
    ptr, idx, val = random_csr(rows, cols, 0.01, seed=9, empty_rows=0.2)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, f16_bits(val))
    for req in (1, 7, 12, 6000):
        p = build_major_nnz_bucket_plan(src, req)
        ref = O.bucket_plan(ptr, rows, req)
        torch.cuda.synchronize()
        assert p["bucket_count"] == ref["bucket_count"]
        for k in ("major_nnz", "major_nnz_sorted", "major_order", "bucket_offsets"):
            assert np.array_equal(u32_np(p[k]), ref[k]), k
    l2 = log2_bucket_offsets(p["major_nnz_sorted"], 12, 5)
    assert np.array_equal(u32_np(l2), O.log2_offsets(ref["major_nnz_sorted"], 12, 5))
    reb, inv = rebuild_compressed_major_order(src, p["major_order"])
    dptr, didx, dval, rinv = O.rebuild(ptr, idx, f16_bits(val), ref["major_order"])
    gptr, gidx, gval = reb.to_numpy()
    assert np.array_equal(gptr, dptr) and np.array_equal(gidx, didx) and np.array_equal(gval, dval)
    assert np.array_equal(u32_np(inv), rinv)


def test_blocked_ell_bucket_plan(ws):
    rows, cols = 640, 900
    ptr, idx, val = random_csr(rows, cols, 0.01, seed=11, empty_rows=0.3)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, f16_bits(val))
    bell = blocked_ell_from_compressed(src, 16, ws)
    order, widths, keys = blocked_ell_bucket_plan(bell, ws)
    gb, _ = bell.to_numpy()
    rk = O.bell_bucket_keys(bell.row_blocks, bell.width, gb)
    ro = np.argsort(rk, kind="stable")
    assert np.array_equal(u32_np(order), ro)
    assert np.array_equal(keys.cpu().numpy().view(np.uint64), rk[ro])
    assert np.array_equal(u32_np(widths), (rk[ro] >> 32).astype(np.uint32))


def test_synth_generator_properties(ws):
    src = synth_csr(3000, 4000, median=200, sigma=0.6, seed=3, value_type=F16, ws=ws)
    ptr, idx, bits = src.to_numpy()
    rn = np.diff(ptr.astype(np.int64))
    assert rn.min() >= 0 and rn.max() <= 4000
    assert abs(np.median(rn) - 200) < 20
    for r in range(0, 3000, 97):
        cs = idx[ptr[r]:ptr[r + 1]]
        assert np.all(np.diff(cs.astype(np.int64)) > 0) and (len(cs) == 0 or cs[-1] < 4000)
    vals = bits.view(np.float16).astype(np.float64)
    assert vals.min() >= 1 and np.all(vals == np.round(vals))
    # deterministic in the seed
    again = synth_csr(3000, 4000, median=200, sigma=0.6, seed=3, value_type=F16, ws=ws)
    a2 = again.to_numpy()
    assert all(np.array_equal(x, y) for x, y in zip((ptr, idx, bits), a2))
    # Zipf popularity: the most popular column appears in far more rows than the median one
    counts = np.bincount(idx, minlength=4000)
    assert counts.max() > 20 * max(np.median(counts), 1)


@pytest.mark.parametrize("bs", [4, 8, 16])
@pytest.mark.parametrize("vt", [F16, F32])
def test_sliced_from_blocked_chunk_paths(ws, bs, vt):
    """16-byte chunk reads (block size a multiple of the chunk) and the
    per-value path give the oracle's Sliced-ELL, with explicit zeros and a
    column count that is not a multiple of the block size."""
    rows, cols = 300, 403
    ptr, idx, val = random_csr(rows, cols, 0.3, seed=bs + vt, zeros=0.1, empty_rows=0.05)
    bits = encode(val, vt)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits, vt)
    bell = blocked_ell_from_compressed(src, bs, ws)
    out = sliced_ell_from_blocked_ell(bell, 32, ws)
    gb, gv = bell.to_numpy()
    ref = O.sliced_ell_from_blocked_ell(rows, cols, bs, bell.ell_cols, gb, gv, 32)
    offs, widths, slots, col, v = out.to_numpy()
    assert np.array_equal(widths, ref["widths"]) and np.array_equal(slots, ref["slot_offsets"])
    assert np.array_equal(col, ref["col_idx"]) and np.array_equal(v.view(np.uint8), np.asarray(ref["val"]).view(np.uint8))
