"""Reduced-precision inputs: the log1p(count / row_total x 1e4) generator and
its BF16 / FP8-e4m3 (per-feature absmax scale) encodings, then the masked
grouped reduce on them vs the FP64 oracle.

Encodings are pinned independently: BF16 by the round-to-nearest-even bit
formula, FP8 by brute force over all finite e4m3 codes (nearest value, ties
to the even code, saturating at 448).  The FP32 log1p values are within 2
ulp of numpy's float32 computation (log1pf vs numpy's log1p).
"""
import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceCSR, to_dev_u32, to_dev_u8
from cellshard_amd.ops import LabelReduce, log1p_normalize
from helpers import BF16, F32, FP8, bf16_bits, decode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fp8_rne_bruteforce(x):
    codes = np.arange(0x7F, dtype=np.uint8)  # finite non-negative codes 0x00..0x7e
    vals = decode(codes, FP8)
    a = np.minimum(np.abs(np.asarray(x, np.float64)), 448.0)
    d = np.abs(a[:, None] - vals[None, :])
    best = d.min(axis=1, keepdims=True)
    cand = d == best
    even = cand & ((codes[None, :] & 1) == 0)
    pick = np.where(even.any(axis=1), np.argmax(even, axis=1), np.argmax(cand, axis=1)).astype(np.uint8)
    return np.where(np.asarray(x) < 0, pick | 0x80, pick).astype(np.uint8)


def test_log1p_encodings(gpu):
    rows, cols = 300, 400
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=9, max_count=300, empty_rows=0.05)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, val.astype(np.float32), F32)
    f32v, _ = log1p_normalize(src, F32)
    got = f32v.val.cpu().numpy().view(np.float32)
    tot = np.zeros(rows)
    np.add.at(tot, np.repeat(np.arange(rows), np.diff(ptr)), val)
    t = np.repeat(tot.astype(np.float32), np.diff(ptr))
    ref = np.log1p(val.astype(np.float32) / t * np.float32(1e4)).astype(np.float32)
    ulp = np.spacing(np.maximum(np.abs(ref), np.float32(1e-30)))
    assert np.all(np.abs(got - ref) <= 2 * ulp)

    bf, _ = log1p_normalize(src, BF16)
    assert np.array_equal(bf.val.cpu().numpy().view(np.int16), bf16_bits(got))

    f8, scale = log1p_normalize(src, FP8)
    sc = scale.cpu().numpy()
    amax = np.zeros(cols, np.float32)
    np.maximum.at(amax, idx, np.abs(got))
    assert np.array_equal(sc, np.where(amax > 0, amax / np.float32(448.0), np.float32(1.0)).astype(np.float32))
    x = (got / sc[idx]).astype(np.float32)
    assert np.array_equal(f8.val.cpu().numpy().view(np.uint8), fp8_rne_bruteforce(x))


@pytest.mark.parametrize("vt", [BF16, FP8])
def test_reduce_on_log1p_values(gpu, vt):
    rows, cols, G = 2000, 900, 128
    ptr, idx, val = random_csr(rows, cols, 0.04, seed=3, max_count=60)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, val.astype(np.float32), F32)
    enc, scale = log1p_normalize(src, vt)
    rng = np.random.default_rng(1)
    labels = rng.integers(0, G, size=rows).astype(np.uint32)
    rk = (rng.random(rows) < 0.8).astype(np.uint8)
    fk = (rng.random(cols) < 0.4).astype(np.uint8)
    lr = LabelReduce()
    got = LabelReduce.fetch(lr.run(enc, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk), scale))
    lr.close()
    bits = enc.val.cpu().numpy()
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, vt, labels, rk, fk, G,
                             scale.cpu().numpy() if scale is not None else None)
    assert np.array_equal(got["nnz"], ref["nnz"]) and np.array_equal(got["group_rows"], ref["group_rows"])
    for k in ("sum", "sumsq"):
        r = ref[k]
        err = np.abs(got[k] - r) / np.where(r != 0, np.abs(r), 1.0)
        assert err.max() <= 1e-6, (k, err.max())  # FP64 partials: observed ~1e-15


@pytest.mark.parametrize("vt", [3, BF16, FP8])
def test_device_value_decoders_all_codes(gpu, vt):
    """The kernels' decoders on every code: f16 vs numpy float16, bf16 vs the
    bit shift, fp8 e4m3 (hardware convert) vs the OCP closed form."""
    from cellshard_amd._lib import check, load
    n = 256 if vt == FP8 else 65536
    codes = np.arange(n, dtype=np.uint32).astype(np.uint8 if vt == FP8 else np.uint16)
    d_in = torch.from_numpy(codes.view(np.uint8 if vt == FP8 else np.int16).copy()).cuda()
    d_out = torch.empty(n, dtype=torch.float32, device="cuda")
    check(load().csamd_decode_values(vt, d_in.data_ptr(), n, d_out.data_ptr(), torch.cuda.current_stream().cuda_stream),
          "decode_values")
    got = d_out.cpu().numpy().astype(np.float64)
    ref = decode(codes.view(np.int16) if vt != FP8 else codes, vt)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan], ref[~nan])
    assert np.array_equal(np.signbit(got[~nan]), np.signbit(ref[~nan]))  # -0 stays -0
