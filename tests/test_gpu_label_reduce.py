"""Row-label grouped reduction on the GPU vs the FP64 oracle.

Tolerance: counts (nnz, group rows, kept ids) bit-exact; sum and sumsq within
1e-6 relative of the FP64 oracle per result (FP64 partials make the observed
error ~1e-15; integer-valued data is exact).
"""
import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceBlockedELL, DeviceCSR, DeviceSlicedELL, to_dev_f32, to_dev_u32, to_dev_u8
from cellshard_amd.ops import ConvertWorkspace, LabelReduce, sliced_ell_from_compressed, synth_csr
from helpers import BF16, F16, F32, FP8, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-6


def check(got, ref, rtol=RTOL):
    assert got["kept_features"] == len(ref["kept_feature_ids"])
    assert np.array_equal(got["kept_feature_ids"], ref["kept_feature_ids"])
    assert np.array_equal(got["group_rows"], ref["group_rows"])
    assert np.array_equal(got["nnz"], ref["nnz"])
    for k in ("sum", "sumsq"):
        r = ref[k]
        err = np.abs(got[k] - r) / np.maximum(np.abs(r), 1e-300)
        err[r == 0] = np.abs(got[k][r == 0])
        assert err.max(initial=0.0) <= rtol, (k, err.max())
    if "mean" in got:
        n = np.maximum(ref["group_rows"], 1)[:, None].astype(np.float64)
        mean = ref["sum"] / n
        var = np.maximum(ref["sumsq"] / n - mean ** 2, 0)
        np.testing.assert_allclose(got["mean"], mean, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(got["var"], var, rtol=1e-9, atol=1e-9 * (np.abs(mean).max(initial=0) ** 2 + 1))


@pytest.fixture()
def lr(gpu):
    r = LabelReduce()
    yield r
    r.close()


def _inputs(rows, cols, G, seed, vt, density=0.1, **kw):
    ptr, idx, val = random_csr(rows, cols, density, seed=seed, fp8_safe=(vt == FP8), **kw)
    rng = np.random.default_rng(seed + 1)
    labels = rng.integers(0, G + 2, size=rows).astype(np.uint32)  # some labels >= G drop rows
    rk = (rng.random(rows) < 0.8).astype(np.uint8)
    fk = (rng.random(cols) < 0.3).astype(np.uint8)
    return ptr, idx, encode(val, vt), labels, rk, fk


@pytest.mark.parametrize("vt", [F16, F32, BF16, FP8])
def test_csr_parity(lr, vt):
    rows, cols, G = 2000, 700, 16
    ptr, idx, bits, labels, rk, fk = _inputs(rows, cols, G, 40 + vt, vt, zeros=0.02, empty_rows=0.05)
    scale = np.random.default_rng(2).uniform(0.1, 3.0, cols).astype(np.float32) if vt == FP8 else None
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits, vt)
    res = lr.run(src, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk), to_dev_f32(scale) if scale is not None else None)
    got = LabelReduce.fetch(res)
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, vt, labels, rk, fk, G, scale)
    check(got, ref)


def test_sell_and_bell_parity(lr):
    rows, cols, G = 777, 300, 5
    ptr, idx, bits, labels, rk, fk = _inputs(rows, cols, G, 7, F16, density=0.15)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, 32)
    sell = DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                      s["col_idx"], s["val"])
    got = LabelReduce.fetch(lr.run(sell, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, rk, fk, G)
    check(got, ref)
    bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, 16)
    bell = DeviceBlockedELL.from_numpy(rows, cols, nnz, 16, ell, bci, bv)
    got = LabelReduce.fetch(lr.run(bell, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
    check(got, ref)


def test_edge_cases(lr):
    cols = 97
    # empty matrix rows, all rows masked, no kept features, G = 1
    ptr = np.zeros(11, np.uint32)
    src = DeviceCSR.from_numpy(10, cols, ptr, np.zeros(0, np.uint32), np.zeros(0, np.int16))
    labels = np.zeros(10, np.uint32)
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), 1))
    assert got["sum"].shape == (1, cols) and not got["sum"].any() and got["group_rows"][0] == 10
    ptr, idx, bits, labels, rk, fk = _inputs(300, cols, 3, 9, F16)
    src = DeviceCSR.from_numpy(300, cols, ptr, idx, bits)
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), 3, to_dev_u8(np.zeros(300, np.uint8)), to_dev_u8(fk)))
    assert got["kept_rows"] == 0 and not got["sum"].any() and not got["group_rows"].any()
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), 3, None, to_dev_u8(np.zeros(cols, np.uint8))))
    assert got["kept_features"] == 0
    # unaligned device pointers take the scalar path
    t_idx = to_dev_u32(np.concatenate([[0], idx]))[1:]
    t_val = torch.from_numpy(np.concatenate([[0], bits]).astype(np.int16)).cuda()[1:]
    src2 = DeviceCSR(300, cols, to_dev_u32(ptr), t_idx, t_val)
    got = LabelReduce.fetch(lr.run(src2, to_dev_u32(labels), 3, to_dev_u8(rk), to_dev_u8(fk)))
    check(got, O.label_reduce_csr(300, cols, ptr, idx, bits, F16, labels, rk, fk, 3))


def test_many_groups_and_feature_tiling(lr):
    # G = 512 labels and 12,000 kept features: exercises several LDS feature tiles
    rows, cols, G = 3000, 14000, 512
    ptr, idx, val = random_csr(rows, cols, 0.01, seed=77)
    bits = encode(val, F16)
    rng = np.random.default_rng(5)
    labels = rng.integers(0, G, size=rows).astype(np.uint32)
    fk = (rng.random(cols) < 0.86).astype(np.uint8)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), G, None, to_dev_u8(fk)))
    check(got, O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, None, fk, G))


def test_repeatable_and_alone_equals_mixed(lr):
    """Same query twice gives identical bits; a query's result does not depend
    on which other query ran before it on the same workspace."""
    rows, cols, G = 1500, 500, 8
    ptr, idx, bits, labels, rk, fk = _inputs(rows, cols, G, 3, F16)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    a = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
    other_fk = to_dev_u8((np.arange(cols) % 2).astype(np.uint8))
    LabelReduce.fetch(lr.run(src, to_dev_u32(labels[::-1].copy()), 3, None, other_fk))
    b = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
    for k in ("sum", "sumsq", "nnz", "group_rows"):
        assert np.array_equal(a[k], b[k])


def test_synthetic_counts_parity(lr):
    """Generator output (fp16 counts, LogNormal row lengths, Zipf columns) through the reduce vs oracle."""
    ws = ConvertWorkspace()
    src = synth_csr(4000, 5000, median=300, sigma=0.6, seed=12, value_type=F16, ws=ws)
    ptr, idx, bits = src.to_numpy()
    rng = np.random.default_rng(12)
    G = 64
    labels = rng.integers(0, G, size=4000).astype(np.uint32)
    rk = (rng.random(4000) < 0.8).astype(np.uint8)
    fk = np.zeros(5000, np.uint8)
    fk[rng.choice(5000, 400, replace=False)] = 1
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
    check(got, O.label_reduce_csr(4000, 5000, ptr, idx, bits, F16, labels, rk, fk, G))
    # and the SELL produced on the GPU from it gives the same numbers
    sell = sliced_ell_from_compressed(src, 32, ws)
    got2 = LabelReduce.fetch(lr.run(sell, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
    for k in ("sum", "sumsq", "nnz"):
        assert np.array_equal(got[k], got2[k])
    ws.close()


@pytest.mark.parametrize("extra", [0, 1, 2, 3])
def test_tail_entries(lr, extra):
    """nnz % 4 == extra: the entries past the last whole 4-entry unit are
    resolved as tail records; put them in a kept row on kept features, for
    CSR and for the SELL built from it, and for matrices with fewer than 4
    entries."""
    rows, cols, G = 41, 64, 3
    rng = np.random.default_rng(100 + extra)
    per_row = rng.integers(0, 9, size=rows)
    per_row[-1] = 4 + extra
    total = int(per_row.sum())
    per_row[-2] += (extra - total) % 4  # nnz % 4 == extra
    ptr = np.concatenate([[0], np.cumsum(per_row)]).astype(np.uint32)
    idx = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in per_row]).astype(np.uint32)
    val = rng.integers(1, 50, size=len(idx)).astype(np.float64)
    bits = encode(val, F16)
    assert len(idx) % 4 == extra
    labels = rng.integers(0, G, size=rows).astype(np.uint32)
    rk = np.ones(rows, np.uint8)
    fk = np.ones(cols, np.uint8)
    fk[::5] = 0
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, rk, fk, G)
    check(lr_fetch(lr, src, labels, G, rk, fk), ref)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, 8)
    sell = DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                      s["col_idx"], s["val"])
    check(lr_fetch(lr, sell, labels, G, rk, fk), ref)
    # a matrix with only `extra` entries: everything is tail
    if extra:
        ptr2 = np.zeros(rows + 1, np.uint32)
        ptr2[-1] = extra
        idx2 = idx[-extra:].copy()
        bits2 = bits[-extra:].copy()
        src2 = DeviceCSR.from_numpy(rows, cols, ptr2, idx2, bits2)
        check(lr_fetch(lr, src2, labels, G, rk, fk), O.label_reduce_csr(rows, cols, ptr2, idx2, bits2, F16, labels, rk, fk, G))


def lr_fetch(lr, src, labels, G, rk, fk):
    return LabelReduce.fetch(lr.run(src, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk)))
