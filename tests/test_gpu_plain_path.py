"""The plain path at its full size: CSR 20,000 x 33,538, fp32 integer counts
from the generator (row length LogNormal(median 1,500, sigma 0.5), Zipf(1.1)
columns, NegBinomial counts); CSR -> Blocked-ELL (block 16) and CSR -> CSC,
then the masked grouped row-sum + nnz-count over 16 uniform labels, row
mask Bernoulli(0.9), 2,000-feature mask.

Blocked-ELL and the transpose are bit-exact vs the oracle (the oracle's
Blocked-ELL converter is itself checked against the reference's own
compiled converter); the reduce's counts are bit-exact and its sums, being
sums of integers below 2^53, are exact too (tolerance 0).
"""
import numpy as np
import pytest

from cellshard_amd.device import to_dev_u32, to_dev_u8
from cellshard_amd.ops import (ConvertWorkspace, LabelReduce, TransposeScratch, blocked_ell_from_compressed,
                               choose_blocked_ell_block_size, compressed_transpose, synth_csr)
from helpers import F32
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROWS, COLS, G = 20_000, 33_538, 16


@pytest.fixture(scope="module")
def plain(gpu):
    ws = ConvertWorkspace()
    src = synth_csr(ROWS, COLS, median=1500.0, sigma=0.5, seed=1, value_type=F32, ws=ws)
    ptr, idx, val = src.to_numpy()
    rng = np.random.default_rng(1)
    labels = rng.integers(0, G, size=ROWS).astype(np.uint32)
    rk = (rng.random(ROWS) < 0.9).astype(np.uint8)
    fk = np.zeros(COLS, np.uint8)
    fk[rng.choice(COLS, 2000, replace=False)] = 1
    yield dict(ws=ws, src=src, ptr=ptr, idx=idx, val=val, labels=labels, rk=rk, fk=fk)
    ws.close()


def test_blocked_ell_16_bit_exact(plain):
    p = plain
    bell = blocked_ell_from_compressed(p["src"], 16, p["ws"])
    bci, bv = bell.to_numpy()
    rbci, rval, ell_cols, nnz = O.blocked_ell_from_compressed(ROWS, COLS, p["ptr"], p["idx"], p["val"], 16)
    assert bell.ell_cols == ell_cols and bell.nnz == nnz
    assert np.array_equal(bci, rbci)
    assert np.array_equal(bv.view(np.uint32), rval.view(np.uint32))
    # the tuner's pick among {8, 16, 32} equals the oracle's
    got = choose_blocked_ell_block_size(p["src"], [8, 16, 32], p["ws"])
    ref = O.choose_blocked_ell_block_size(ROWS, COLS, p["ptr"], p["idx"], [8, 16, 32], 4)
    assert got[0] == ref[0] and got[2] == ref[2] and abs(got[1] - ref[1]) <= 1e-12


def test_transpose_bit_exact(plain):
    p = plain
    optr, oidx, oval = compressed_transpose(p["src"], TransposeScratch()).to_numpy()
    rptr, ridx, rval = O.transpose(ROWS, COLS, p["ptr"], p["idx"], p["val"])
    assert np.array_equal(optr, rptr) and np.array_equal(oidx, ridx)
    assert np.array_equal(oval.view(np.uint32), np.ascontiguousarray(rval).view(np.uint32))


def test_masked_group_sum_and_nnz(plain):
    p = plain
    lr = LabelReduce()
    got = LabelReduce.fetch(lr.run(p["src"], to_dev_u32(p["labels"]), G, to_dev_u8(p["rk"]), to_dev_u8(p["fk"])))
    lr.close()
    ref = O.label_reduce_csr(ROWS, COLS, p["ptr"], p["idx"], p["val"], F32, p["labels"], p["rk"], p["fk"], G)
    assert np.array_equal(got["kept_feature_ids"], ref["kept_feature_ids"])
    assert np.array_equal(got["nnz"], ref["nnz"]) and np.array_equal(got["group_rows"], ref["group_rows"])
    assert np.array_equal(got["sum"], ref["sum"])  # integer counts: exact
    assert np.array_equal(got["sumsq"], ref["sumsq"])
