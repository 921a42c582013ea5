"""COO -> Blocked-ELL (and the COO -> CSR sort it starts with) on the GPU vs
the oracle: the COO, stably sorted by (row, column), is a CSR whose
Blocked-ELL the oracle builds (its converter is checked against the
reference's own compiled one).  Bit-exact."""
import numpy as np
import pytest
import torch

from cellshard_amd.device import to_dev_u32
from cellshard_amd.ops import blocked_ell_from_coo, choose_blocked_ell_block_size_from_coo, compressed_from_coo
from helpers import F16, F32, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _shuffled_coo(rows, cols, seed, duplicates=0.0):
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=seed, duplicates=duplicates, zeros=0.05, empty_rows=0.1)
    r = np.repeat(np.arange(rows, dtype=np.uint32), np.diff(ptr.astype(np.int64)))
    perm = np.random.default_rng(seed).permutation(len(idx))
    return r[perm], idx[perm], val[perm]


def _as_csr(rows, r, c, v):
    order = np.lexsort((c, r))  # stable: by row, then column, ties in input order
    ptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=rows))]).astype(np.uint32)
    return ptr, c[order].astype(np.uint32), v[order]


@pytest.mark.parametrize("dup", [0.0, 0.2])
def test_blocked_ell_from_coo_matches_oracle(gpu, dup):
    rows, cols = 333, 700
    r, c, v = _shuffled_coo(rows, cols, 5, dup)
    bits = encode(v, F16)
    for bs in (8, 16):
        bci, bval, ell_cols, nnz = blocked_ell_from_coo(rows, cols, r, c, bits, cols, bs)
        ptr, idx, vb = _as_csr(rows, r, c, bits)
        rbci, rval, rell, rnnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, vb, bs)
        assert (ell_cols, nnz) == (rell, rnnz)
        assert np.array_equal(bci, rbci) and np.array_equal(bval, rval.view(np.uint16))
    got = choose_blocked_ell_block_size_from_coo(rows, cols, r, c, bits, cols, [8, 16, 32])
    ptr, idx, vb = _as_csr(rows, r, c, bits)
    ref = O.choose_blocked_ell_block_size(rows, cols, ptr, idx, [8, 16, 32], 2)
    assert got[0] == ref[0] and got[2] == ref[2] and abs(got[1] - ref[1]) <= 1e-12


def test_feature_remap_and_errors(gpu):
    rows, cols_src, cols = 120, 90, 200
    r, c, v = _shuffled_coo(rows, cols_src, 8)
    bits = encode(v, F16)
    fmap = np.random.default_rng(2).choice(cols, cols_src, replace=False).astype(np.uint32)
    bci, bval, ell_cols, nnz = blocked_ell_from_coo(rows, cols_src, r, c, bits, cols, 16, feature_to_global=fmap)
    ptr, idx, vb = _as_csr(rows, r, fmap[c], bits)
    rbci, rval, rell, rnnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, vb, 16)
    assert (ell_cols, nnz) == (rell, rnnz) and np.array_equal(bci, rbci) and np.array_equal(bval, rval.view(np.uint16))
    bad = c.copy()
    bad[3] = cols_src + 500
    with pytest.raises(Exception):
        blocked_ell_from_coo(rows, cols_src, r, bad, bits, cols_src, 16)  # column outside the matrix
    badr = r.copy()
    badr[0] = rows
    with pytest.raises(Exception):
        blocked_ell_from_coo(rows, cols_src, badr, c, bits, cols_src, 16)  # row outside the matrix


def test_compressed_from_coo_device_f32(gpu):
    rows, cols = 1000, 5000
    r, c, v = _shuffled_coo(rows, cols, 11, 0.1)
    bits = encode(v, F32)
    d_val = torch.from_numpy(bits.view(np.float32).copy()).cuda()
    csr = compressed_from_coo(rows, cols, to_dev_u32(r), to_dev_u32(c), d_val, F32)
    gptr, gidx, gval = csr.to_numpy()
    ptr, idx, vb = _as_csr(rows, r, c, bits)
    assert np.array_equal(gptr, ptr) and np.array_equal(gidx, idx)
    assert np.array_equal(gval.view(np.uint32), vb.view(np.uint32))
    # no entries
    e = compressed_from_coo(7, 3, to_dev_u32(np.zeros(0, np.uint32)), to_dev_u32(np.zeros(0, np.uint32)),
                            torch.zeros(0, dtype=torch.float32, device="cuda"), F32)
    assert np.array_equal(e.to_numpy()[0], np.zeros(8, np.uint32))
