"""The host-struct converters that back the drop-in C++ wrappers
(cellshard::convert::blocked_ell_from_compressed(_auto), choose_*,
sliced_ell_from_compressed(_auto), sliced_ell_from_blocked_ell), called
through the C ABI with host sparse::compressed / blocked_ell structs.

Blocked-ELL results are compared bit for bit with the golden vectors the
reference's own CPU converter produced (tests/golden/bell_ref_vectors.npz);
Sliced-ELL results with the oracle's restatement, from CSR and from CSC
input (the CSC path transposes on the device first).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
from pathlib import Path

from cellshard_amd import abi as A
from cellshard_amd._lib import load
from helpers import f16_bits, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).copy()


class HostCompressed:
    """A csamd_host_compressed over numpy arrays (storage NULL: the caller
    owns the arrays)."""

    def __init__(self, rows, cols, ptr, idx, bits, axis=A.COMPRESSED_BY_ROW):
        self.keep = [np.ascontiguousarray(ptr, np.uint32), np.ascontiguousarray(idx, np.uint32),
                     np.ascontiguousarray(np.asarray(bits).view(np.uint16))]
        m = A.host_compressed()
        m.rows, m.cols, m.nnz, m.axis, m.flags = rows, cols, len(idx), axis, 0
        m.storage = None
        m.majorPtr, m.minorIdx = self.keep[0].ctypes.data, self.keep[1].ctypes.data if len(idx) else None
        m.val = self.keep[2].ctypes.data if len(idx) else None
        self.m = m


def _bell_out(b):
    rb = (b.rows + b.block_size - 1) // b.block_size if b.block_size else 0
    w = b.ell_cols // b.block_size if b.block_size else 0
    return _arr(b.blockColIdx, rb * w, np.uint32), _arr(b.val, b.rows * b.ell_cols, np.int16)


def _sell_out(s):
    n = s.slice_count
    offs = _arr(s.slice_row_offsets, n + 1, np.uint32)
    wid = _arr(s.slice_widths, n, np.uint32)
    total = int(((offs[1:].astype(np.int64) - offs[:-1]) * wid).sum()) if n else 0
    return offs, wid, _arr(s.col_idx, total, np.uint32), _arr(s.val, total, np.int16)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD / "bell_ref_vectors.npz")


def _cases(g):
    return sorted({k.split("/")[0] for k in g.files})


def test_host_blocked_ell_matches_reference_vectors(gpu, gold):
    lib = load()
    for name in _cases(gold):
        rows, cols = (int(x) for x in gold[f"{name}/meta"])
        src = HostCompressed(rows, cols, gold[f"{name}/ptr"], gold[f"{name}/idx"], gold[f"{name}/val"])
        for key in [k for k in gold.files if k.startswith(f"{name}/b") and k.endswith("/bci")]:
            bs = int(key.split("/")[1][1:])
            dst = A.host_blocked_ell()
            assert lib.csamd_host_blocked_ell_from_compressed(C.byref(src.m), bs, C.byref(dst)) == 1, (name, bs)
            ell, nnz = (int(x) for x in gold[f"{name}/b{bs}/shape"])
            assert (dst.rows, dst.cols, dst.block_size, dst.ell_cols, dst.nnz) == (rows, cols, bs, ell, nnz), (name, bs)
            bci, val = _bell_out(dst)
            assert np.array_equal(bci, gold[f"{name}/b{bs}/bci"][:len(bci)]), (name, bs)
            assert np.array_equal(val, gold[f"{name}/b{bs}/val"][:len(val)]), (name, bs)
            lib.csamd_host_blocked_ell_clear(C.byref(dst))
            assert dst.storage is None and dst.rows == 0


def test_host_blocked_ell_tuner_and_auto_match_reference(gpu, gold):
    lib = load()
    for name in _cases(gold):
        rows, cols = (int(x) for x in gold[f"{name}/meta"])
        src = HostCompressed(rows, cols, gold[f"{name}/ptr"], gold[f"{name}/idx"], gold[f"{name}/val"])
        for k in range(3):
            if f"{name}/choose{k}/cands" not in gold.files:
                continue
            cands = gold[f"{name}/choose{k}/cands"].astype(np.uint32)
            carr = (C.c_uint32 * len(cands))(*cands.tolist())
            rbs, rpad = (int(x) for x in gold[f"{name}/choose{k}/result"])
            out = A.blocked_ell_tune_result()
            assert lib.csamd_host_choose_blocked_ell_block_size(C.byref(src.m), carr, len(cands), C.byref(out)) == 1
            assert (out.block_size, out.padded_bytes) == (rbs, rpad), (name, k)
            assert abs(out.fill_ratio - float(gold[f"{name}/choose{k}/fill"][0])) < 1e-12
            dst, picked = A.host_blocked_ell(), A.blocked_ell_tune_result()
            assert lib.csamd_host_blocked_ell_from_compressed_auto(C.byref(src.m), carr, len(cands), C.byref(dst),
                                                                   C.byref(picked)) == 1
            assert picked.block_size == rbs and dst.block_size == rbs
            bci, val = _bell_out(dst)
            one = A.host_blocked_ell()
            assert lib.csamd_host_blocked_ell_from_compressed(C.byref(src.m), rbs, C.byref(one)) == 1
            assert (np.array_equal(bci, _bell_out(one)[0]) and np.array_equal(val, _bell_out(one)[1])), (name, k)
            lib.csamd_host_blocked_ell_clear(C.byref(one))
            lib.csamd_host_blocked_ell_clear(C.byref(dst))


def _csc_of(rows, cols, ptr, idx, bits):
    m = sp.csr_matrix((np.arange(1, len(idx) + 1), idx, ptr), shape=(rows, cols)).tocsc()
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32), np.asarray(bits)[m.data - 1]


@pytest.mark.parametrize("axis", [A.COMPRESSED_BY_ROW, A.COMPRESSED_BY_COL])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_host_sliced_ell_from_compressed(gpu, axis, seed):
    lib = load()
    rows, cols = [(45, 30), (1, 9), (70, 200)][seed]
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=60 + seed, empty_rows=0.15)
    bits = f16_bits(val)
    if axis == A.COMPRESSED_BY_COL:
        p2, i2, b2 = _csc_of(rows, cols, ptr, idx, bits)
        src = HostCompressed(rows, cols, p2, i2, b2, axis)
    else:
        src = HostCompressed(rows, cols, ptr, idx, bits, axis)
    for S in (1, 4, 7, 32, 100):
        ref = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, S)
        dst = A.host_sliced_ell()
        assert lib.csamd_host_sliced_ell_from_compressed(C.byref(src.m), S, C.byref(dst)) == 1, S
        offs, wid, col, v = _sell_out(dst)
        assert (dst.rows, dst.cols, dst.nnz) == (rows, cols, len(idx))
        assert np.array_equal(offs, ref["offsets"]) and np.array_equal(wid, ref["widths"]), S
        assert np.array_equal(col, ref["col_idx"]) and np.array_equal(v, ref["val"].view(np.int16)), S
        lib.csamd_host_sliced_ell_clear(C.byref(dst))
    # tuner and auto
    cands = [2, 4, 8, 16]
    carr = (C.c_uint32 * 4)(*cands)
    ref = O.choose_sliced_ell_slice_rows(np.diff(ptr), rows, len(idx), cands)
    out = A.sliced_ell_tune_result()
    assert lib.csamd_host_choose_sliced_ell_slice_rows(C.byref(src.m), carr, 4, C.byref(out)) == 1
    assert out.slice_rows == ref["slice_rows"] and out.padded_slots == ref["padded_slots"]
    assert abs(out.score - ref["score"]) < 1e-12
    dst, picked = A.host_sliced_ell(), A.sliced_ell_tune_result()
    assert lib.csamd_host_sliced_ell_from_compressed_auto(C.byref(src.m), carr, 4, C.byref(dst), C.byref(picked)) == 1
    want = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, ref["slice_rows"])
    offs, wid, col, v = _sell_out(dst)
    assert picked.slice_rows == ref["slice_rows"] and np.array_equal(col, want["col_idx"]) and np.array_equal(offs, want["offsets"])
    lib.csamd_host_sliced_ell_clear(C.byref(dst))


def test_host_sliced_ell_from_blocked_ell(gpu):
    lib = load()
    rows, cols, bs = 26, 31, 4
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=12, zeros=0.1)
    bits = f16_bits(val)
    src = HostCompressed(rows, cols, ptr, idx, bits)
    bell = A.host_blocked_ell()
    assert lib.csamd_host_blocked_ell_from_compressed(C.byref(src.m), bs, C.byref(bell)) == 1
    bci, bv, ell, _ = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    for S in (1, 8, 32):
        ref = O.sliced_ell_from_blocked_ell(rows, cols, bs, ell, bci, bv, S)
        dst = A.host_sliced_ell()
        assert lib.csamd_host_sliced_ell_from_blocked_ell(C.byref(bell), S, C.byref(dst)) == 1
        offs, wid, col, v = _sell_out(dst)
        assert np.array_equal(offs, ref["offsets"]) and np.array_equal(wid, ref["widths"])
        assert np.array_equal(col, ref["col_idx"]) and np.array_equal(v, ref["val"].view(np.int16))
        lib.csamd_host_sliced_ell_clear(C.byref(dst))
    lib.csamd_host_blocked_ell_clear(C.byref(bell))
