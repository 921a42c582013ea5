"""compute_sparse_group_reduce over Blocked-ELL through the streaming path
(block sizes 8, 16, 32): windows of row blocks whose block columns are staged
in LDS, value loads skipped for slots of padding and for blocks with no kept
feature, and the kept-entry group accumulation shared with the row-major
path.  Every case is checked against the oracle (mask_groups.cu:107-381
semantics): detected counts and keep flags exactly, sums within float32
reassociation.

The cases drive each branch of the path: row counts that are not a multiple
of the block size, several windows per workgroup (wide rows), heavy padding,
no feature mask with more columns than the LDS group table holds (masks read
from memory), no groups, every row masked, and each value type.
"""
import numpy as np
import pytest

from cellshard_amd import abi as A
from cellshard_amd import runtime as csr_rt
from cellshard_amd.device import DeviceBlockedELL, ptr, to_dev_u32, to_dev_u8
from helpers import BF16, F16, F32, FP8, encode
from oracle import oracle as O
from test_gpu_mask_groups import _compare, fetch, ws  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF

# name: block size, value type, groups, feature keep fraction (None: no mask), rows, cols, width (slots), padding fraction
CASES = {
    "bs32_f32_g32": (32, F32, 32, 0.3, 3001, 5000, 40, 0.1),
    "bs8_fp8_g5_windows": (8, FP8, 5, 0.1, 3000, 120_000, 3000, 0.05),
    "bs16_bf16_global_masks": (16, BF16, 32, None, 2000, 40_000, 300, 0.0),
    "bs16_f16_g0": (16, F16, 0, 0.5, 1500, 3000, 64, 0.2),
    "bs8_f16_g12_padding": (8, F16, 12, 0.2, 2500, 9000, 200, 0.5),
    "bs32_bf16_all_masked": (32, BF16, 7, 0.4, 700, 4000, 30, 0.0),
}


def _bell(rows, cols, bs, width, pad, vt, seed):
    rng = np.random.default_rng(seed)
    row_blocks = -(-rows // bs)
    nblocks = -(-cols // bs)
    bci = np.empty((row_blocks, width), np.uint32)
    for rb in range(row_blocks):  # distinct block columns per row block, ascending
        bci[rb] = np.sort(rng.choice(nblocks, width, replace=False)).astype(np.uint32)
    bci[rng.random(bci.shape) < pad] = INVALID
    ell = width * bs
    # values 0, 1, 2, 3, 5 with probabilities .70 .15 .07 .05 .03, encoded once
    # (values of padding slots and of columns past `cols` are never entries)
    table = encode(np.array([0.0, 1.0, 2.0, 3.0, 5.0]), vt)
    pick = np.searchsorted(np.array([70, 85, 92, 97]), rng.integers(0, 100, size=rows * ell, dtype=np.int16), side="right")
    return bci.ravel(), table[pick], ell


@pytest.mark.parametrize("case", list(CASES))
def test_bell_stream_parity(ws, case):  # noqa: F811
    bs, vt, G, fkeep, rows, cols, width, pad = CASES[case]
    print(f"bell stream case {case}", flush=True)
    seed = 40 + list(CASES).index(case)
    bci, bits, ell = _bell(rows, cols, bs, width, pad, vt, seed)
    rng = np.random.default_rng(seed + 1)
    rk = np.zeros(rows, np.uint8) if case.endswith("all_masked") else (rng.random(rows) < 0.85).astype(np.uint8)
    fk = None if fkeep is None else (rng.random(cols) < fkeep).astype(np.uint8)
    fgm = rng.integers(0, 2 ** 32, size=cols, dtype=np.uint64).astype(np.uint32)
    fgm[rng.random(cols) < 0.3] = 0
    bell = DeviceBlockedELL.from_numpy(rows, cols, 0, bs, ell, bci, bits, vt)
    d_rk, d_fgm = to_dev_u8(rk), to_dev_u32(fgm)
    d_fk = None if fk is None else to_dev_u8(fk)
    groups = A.group_mask_config_view(group_count=G, feature_group_masks=ptr(d_fgm) if G else None)
    masks = A.row_feature_mask_view(row_keep=ptr(d_rk), feature_keep=ptr(d_fk) if fk is not None else None)
    filt_t = [1.0, 2, 0.7, 0 if G == 0 else G - 1]
    filt = A.sparse_group_filter_params(*filt_t)
    res = A.sparse_group_reduce_result()
    assert csr_rt.compute_sparse_group_reduce(bell.view(), ws, groups, masks, filt, res, vt) == 1
    want = O.group_reduce_bell(rows, cols, bs, ell, bci, bits, vt, rk, fk, fgm if G else None, G, filt_t)
    got = fetch(ws, res)
    _compare(got, want, G, vt)
    if case.endswith("all_masked"):
        assert not got["keep"].any() and not got["row_totals"].any()
    else:
        assert got["detected"].any()
