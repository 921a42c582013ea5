"""Parity at the full sizes of the measured workloads.

The 1,000,000 x 36,601 fp16 matrix holds about 2.39e9 stored entries, more
than 2^31, so every 32-bit entry index in the kernels runs past the signed
range here.  Checks:

* row-label grouped reduce (64 Dirichlet(0.5) labels, row keep 0.8, 3,000
  features) vs the C oracle over every row (chunked over the host cores):
  nnz and group rows bit-exact, sum / sumsq within 1e-6 relative per result;
* CSR -> CSC -> CSR returns the source bit for bit; the CSC's segment
  pointers equal the column counts and every segment is strictly ascending;
* the nnz bucket plan (12 buckets) is a stable ascending-nnz permutation and
  rebuilt row i equals source row order[i] bit for bit;
* Sliced-ELL (slice height 32) of the rebuilt CSR: each row's first nnz slots
  are its CSR entries, every other slot is padding (invalid column, value 0);
  the reduce over it (3.1e9 slots) equals the CSR reduce with the labels and
  masks permuted the same way;
* 500,000 x 36,601 log1p values as BF16 and as FP8 e4m3 with per-feature
  scale, 128 labels: counts bit-exact, sums within 1e-6 vs the FP64 oracle.

Device checks use torch on the device in row chunks; the oracle runs on the
host from copies of the device arrays.
"""
import time

import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceCSR, to_dev_u32, to_dev_u8
from cellshard_amd.ops import (ConvertWorkspace, LabelReduce, TransposeScratch, build_major_nnz_bucket_plan, compressed_transpose,
                               log1p_normalize, rebuild_compressed_major_order, sliced_ell_from_compressed, synth_csr)
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-6
INVALID = -1  # 0xFFFFFFFF as int32 bits
ROWS, COLS, GROUPS, FEATURES = 1_000_000, 36_601, 64, 3000
ROW_CHUNK = 100_000


def u64(t):
    """u32 array held as int32 bits -> int64 values."""
    return t.long() & 0xFFFFFFFF


def say(msg):
    print(f"[full-size] {msg}", flush=True)


def rel_err(got, ref):
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    err[ref == 0] = np.abs(got[ref == 0])
    return float(err.max(initial=0.0))


def check_label(got, ref):
    assert np.array_equal(got["kept_feature_ids"], ref["kept_feature_ids"])
    assert np.array_equal(got["group_rows"], ref["group_rows"])
    assert np.array_equal(got["nnz"], ref["nnz"])
    errs = {k: rel_err(got[k], ref[k]) for k in ("sum", "sumsq")}
    assert max(errs.values()) <= RTOL, errs
    return errs


@pytest.fixture(scope="module")
def big(gpu):
    t0 = time.time()
    ws = ConvertWorkspace()
    src = synth_csr(ROWS, COLS, median=2000.0, sigma=0.6, seed=1, value_type=A.VALUE_F16, ws=ws)
    say(f"generated {ROWS} x {COLS} fp16, nnz {src.nnz} in {time.time() - t0:.1f} s")
    assert src.nnz > 2 ** 31
    rng = np.random.default_rng(1)
    labels = rng.choice(GROUPS, size=ROWS, p=rng.dirichlet(np.full(GROUPS, 0.5))).astype(np.uint32)
    rk = (rng.random(ROWS) < 0.8).astype(np.uint8)
    fk = np.zeros(COLS, np.uint8)
    fk[rng.choice(COLS, FEATURES, replace=False)] = 1
    yield dict(src=src, ws=ws, labels=labels, rk=rk, fk=fk)
    ws.close()
    del src
    torch.cuda.empty_cache()


def test_label_reduce_matches_oracle_full(big):
    src, labels, rk, fk = big["src"], big["labels"], big["rk"], big["fk"]
    lr = LabelReduce()
    t0 = time.time()
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), GROUPS, to_dev_u8(rk), to_dev_u8(fk)))
    lr.close()
    p, idx, val = src.major_ptr.cpu().numpy().view(np.uint32), src.minor_idx.cpu().numpy().view(np.uint32), src.val.cpu().numpy()
    say(f"reduce + copies {time.time() - t0:.1f} s")
    t0 = time.time()
    ref = O.label_reduce_csr_chunked(ROWS, COLS, p, idx, val, A.VALUE_F16, labels, rk, fk, GROUPS)
    say(f"oracle {time.time() - t0:.1f} s on {O.worker_threads()} threads")
    errs = check_label(got, ref)
    say(f"label reduce: counts exact, max rel err {errs}")


def _row_gather_check(dst, dst_rows, src, src_rows_of):
    """Entries of dst rows [dst_rows] equal those of src rows src_rows_of[...]
    in the same order (device, int64 gather indices)."""
    dp = u64(dst.major_ptr)
    sp = u64(src.major_ptr)
    lens = dp[dst_rows + 1] - dp[dst_rows]
    assert torch.equal(lens, sp[src_rows_of + 1] - sp[src_rows_of])
    n = int(lens.sum())
    if n == 0:
        return
    rep = torch.repeat_interleave(torch.arange(dst_rows.numel(), device=dp.device), lens, output_size=n)
    off = torch.arange(n, device=dp.device) - torch.repeat_interleave(torch.cumsum(lens, 0) - lens, lens, output_size=n)
    dpos = dp[dst_rows][rep] + off
    spos = sp[src_rows_of][rep] + off
    assert torch.equal(dst.minor_idx[dpos], src.minor_idx[spos])
    assert torch.equal(dst.val[dpos], src.val[spos])


def test_transpose_round_trip_full(big):
    src = big["src"]
    scratch = TransposeScratch()
    t0 = time.time()
    csc = compressed_transpose(src, scratch)
    torch.cuda.synchronize()
    say(f"CSR -> CSC {time.time() - t0:.2f} s")
    # segment pointers = column counts
    counts = torch.bincount(src.minor_idx.long(), minlength=COLS)
    assert torch.equal(u64(csc.major_ptr)[1:] - u64(csc.major_ptr)[:-1], counts)
    assert int(u64(csc.major_ptr)[-1]) == src.nnz
    # strictly ascending source rows inside each segment (row ids < 2^31:
    # int32 differences are exact)
    ri = csc.minor_idx
    step = ri[1:] - ri[:-1]
    starts = torch.zeros(src.nnz + 1, dtype=torch.bool, device=ri.device)
    starts[u64(csc.major_ptr)[1:-1]] = True
    assert bool(((step > 0) | starts[1:-1]).all())
    del ri, step, starts, counts
    t0 = time.time()
    back = compressed_transpose(csc, scratch)
    torch.cuda.synchronize()
    say(f"CSC -> CSR {time.time() - t0:.2f} s")
    del csc, scratch
    assert torch.equal(back.major_ptr, src.major_ptr)
    assert torch.equal(back.minor_idx, src.minor_idx)
    assert torch.equal(back.val, src.val)
    del back
    torch.cuda.empty_cache()
    say("transpose round trip bit-exact")


@pytest.fixture(scope="module")
def rebuilt(big):
    src = big["src"]
    plan = build_major_nnz_bucket_plan(src, 12)
    reb, inv = rebuild_compressed_major_order(src, plan["major_order"])
    torch.cuda.synchronize()
    yield dict(plan=plan, reb=reb, inv=inv)
    del reb, inv, plan
    torch.cuda.empty_cache()


def test_bucket_rebuild_full(big, rebuilt):
    src = big["src"]
    plan, reb, inv = rebuilt["plan"], rebuilt["reb"], rebuilt["inv"]
    order = u64(plan["major_order"])
    nnz = (u64(src.major_ptr)[1:] - u64(src.major_ptr)[:-1])
    # a permutation, stable and ascending by row nnz
    assert torch.equal(torch.sort(order).values, torch.arange(ROWS, device=order.device))
    key = nnz[order] * ROWS + order
    assert bool((key[1:] > key[:-1]).all())
    assert torch.equal(u64(plan["major_nnz_sorted"]), nnz[order])
    assert torch.equal(u64(inv)[order], torch.arange(ROWS, device=order.device))
    for a in range(0, ROWS, ROW_CHUNK):
        rows = torch.arange(a, min(a + ROW_CHUNK, ROWS), device=order.device)
        _row_gather_check(reb, rows, src, order[rows])
    say("bucket rebuild bit-exact")


def test_sliced_ell_full(big, rebuilt):
    reb, order = rebuilt["reb"], rebuilt["plan"]["major_order"].long()
    sell = sliced_ell_from_compressed(reb, 32, big["ws"])
    torch.cuda.synchronize()
    say(f"SELL-32: {sell.total_slots} slots ({sell.total_slots / reb.nnz:.3f} per entry)")
    assert sell.total_slots > 2 ** 31
    dp = u64(reb.major_ptr)
    offs, wid, slot = u64(sell.slice_row_offsets), u64(sell.slice_widths), u64(sell.slice_slot_offsets)
    assert torch.equal(offs, torch.clamp(torch.arange(sell.slice_count + 1, device=dp.device) * 32, max=ROWS))
    nnz = dp[1:] - dp[:-1]
    slice_max = torch.zeros(sell.slice_count, dtype=torch.long, device=dp.device).scatter_reduce(
        0, torch.arange(ROWS, device=dp.device) // 32, nnz, reduce="amax")
    assert torch.equal(wid, slice_max)
    assert int(slot[0]) == 0
    assert torch.equal(slot[1:], torch.cumsum((offs[1:] - offs[:-1]) * wid, 0)[:-1])
    covered = torch.zeros(sell.total_slots, dtype=torch.bool, device=dp.device)
    for a in range(0, ROWS, ROW_CHUNK):
        rows = torch.arange(a, min(a + ROW_CHUNK, ROWS), device=dp.device)
        lens = nnz[rows]
        n = int(lens.sum())
        rep = torch.repeat_interleave(torch.arange(rows.numel(), device=dp.device), lens, output_size=n)
        off = torch.arange(n, device=dp.device) - torch.repeat_interleave(torch.cumsum(lens, 0) - lens, lens, output_size=n)
        r = rows[rep]
        s = r // 32
        spos = slot[s] + (r - 32 * s) * wid[s] + off
        cpos = dp[r] + off
        assert torch.equal(sell.col_idx[spos], reb.minor_idx[cpos])
        assert torch.equal(sell.val[spos], reb.val[cpos])
        covered[spos] = True
    # (no boolean-mask indexing: its index tensor cannot address > 2^31 slots)
    assert bool(((sell.col_idx == INVALID) | covered).all())
    assert bool(((sell.val == 0) | covered).all())
    del covered
    say("SELL-32 slots bit-exact")
    # the reduce over the Sliced-ELL = the reduce over the source CSR
    labels, rk, fk = big["labels"], big["rk"], big["fk"]
    o = order.cpu().numpy()
    lr = LabelReduce()
    d_fk = to_dev_u8(fk)
    got = LabelReduce.fetch(lr.run(sell, to_dev_u32(labels[o]), GROUPS, to_dev_u8(rk[o]), d_fk))
    ref = LabelReduce.fetch(lr.run(big["src"], to_dev_u32(labels), GROUPS, to_dev_u8(rk), d_fk))
    lr.close()
    del sell
    torch.cuda.empty_cache()
    assert np.array_equal(got["nnz"], ref["nnz"]) and np.array_equal(got["group_rows"], ref["group_rows"])
    errs = {k: rel_err(got[k], ref[k]) for k in ("sum", "sumsq")}
    assert max(errs.values()) <= 1e-12, errs
    say(f"SELL-32 reduce == CSR reduce ({errs})")


@pytest.mark.parametrize("vt", [A.VALUE_BF16, A.VALUE_FP8_E4M3])
def test_log1p_reduce_matches_oracle_full(gpu, vt):
    R, G = 500_000, 128
    ws = ConvertWorkspace()
    counts = synth_csr(R, COLS, median=1500.0, sigma=0.5, seed=4, value_type=A.VALUE_F32, ws=ws)
    src, scale = log1p_normalize(counts, vt, ws=ws)
    del counts
    rng = np.random.default_rng(4)
    labels = rng.integers(0, G, size=R).astype(np.uint32)
    rk = (rng.random(R) < 0.8).astype(np.uint8)
    fk = np.zeros(COLS, np.uint8)
    fk[rng.choice(COLS, FEATURES, replace=False)] = 1
    lr = LabelReduce()
    got = LabelReduce.fetch(lr.run(src, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk), scale))
    lr.close()
    p, idx, val = src.major_ptr.cpu().numpy().view(np.uint32), src.minor_idx.cpu().numpy().view(np.uint32), src.val.cpu().numpy()
    sc = scale.cpu().numpy() if scale is not None else None
    ws.close()
    del src, scale
    torch.cuda.empty_cache()
    ref = O.label_reduce_csr_chunked(R, COLS, p, idx, val, vt, labels, rk, fk, G, sc)
    errs = check_label(got, ref)
    say(f"log1p {'bf16' if vt == A.VALUE_BF16 else 'fp8'} {len(idx)} entries: counts exact, max rel err {errs}")
