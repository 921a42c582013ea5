"""Column / row orders against Blocked-ELL padding on the GPU against the
oracle restatements (row-block support signatures; rank barycenters)."""
import numpy as np
import pytest

from cellshard_amd import CellShardError
from cellshard_amd.device import DeviceCSR
from cellshard_amd.ops import (ConvertWorkspace, blocked_ell_from_compressed, build_column_order_from_sampled_coo,
                               build_row_order_from_coo, build_shard_column_maps)
from helpers import F16, F32, FP8, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NP_BITS = {F16: np.uint16, F32: np.uint32, FP8: np.uint8}


@pytest.fixture(scope="module")
def ws():
    return ConvertWorkspace()


@pytest.mark.parametrize("vt", [F16, F32, FP8])
def test_shard_column_maps_match_oracle(ws, vt):
    cols = 700
    parts, oparts = [], []
    for k, (rows, bs) in enumerate(((130, 8), (77, 16), (260, 4))):
        ptr, idx, val = random_csr(rows, cols, 0.02 + 0.01 * k, seed=70 + k + vt, zeros=0.2, empty_rows=0.1, fp8_safe=(vt == FP8))
        bell = blocked_ell_from_compressed(DeviceCSR.from_numpy(rows, cols, ptr, idx, encode(val, vt), vt), bs, ws)
        gb, gv = bell.to_numpy()
        parts.append(bell)
        oparts.append({"rows": rows, "block_size": bs, "ell_cols": bell.ell_cols, "bci": gb,
                       "val": np.ascontiguousarray(gv).view(NP_BITS[vt])})
    e2c, c2e = build_shard_column_maps(parts, cols, ws)
    r2c, rc2e = O.shard_column_maps(oparts, vt, cols)
    assert np.array_equal(e2c, r2c) and np.array_equal(c2e, rc2e)


@pytest.mark.parametrize("seed", [1, 2])
def test_barycenter_orders_match_oracle(seed):
    rng = np.random.default_rng(seed)
    rows, cols, nnz = 3000, 900, 60000
    ri = rng.integers(0, rows, nnz).astype(np.uint32)
    ci = (rng.zipf(1.3, nnz) % cols).astype(np.uint32)
    rank = rng.permutation(rows).astype(np.uint32)
    rank[rng.random(rows) < 0.05] = 0xFFFFFFFF
    e2c, c2e = build_column_order_from_sampled_coo(rows, cols, ri, ci, rank)
    r2c, rc2e = O.barycenter_order(rows, cols, ri, ci, rank, by_col=True)
    assert np.array_equal(e2c, r2c) and np.array_equal(c2e, rc2e)
    re2c, rec2e = build_row_order_from_coo(rows, cols, ri, ci, c2e)
    o2c, oc2e = O.barycenter_order(rows, cols, ri, ci, c2e, by_col=False)
    assert np.array_equal(re2c, o2c) and np.array_equal(rec2e, oc2e)


def test_barycenter_rejects_out_of_range_indices():
    with pytest.raises(CellShardError, match="exceeds"):
        build_column_order_from_sampled_coo(4, 3, np.array([0, 4], np.uint32), np.array([1, 2], np.uint32), np.arange(4, dtype=np.uint32))
