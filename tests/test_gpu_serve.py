"""Serving loop on the GPU: every query's result equals the CPU oracle
(restated on the row-major CSR) and is the same alone and in the mix.

With integer counts FP64 sums are exact: results are compared bit for bit
(tolerance 0), nnz counts included.  With real values (log1p-normalised
counts stored as BF16 or FP8 e4m3) the per-(feature, group) FP64 sums are
accumulated with LDS atomics in arrival order, so alone and mixed runs may
differ in the last bits: they are compared with each other and with the
oracle within 1e-6 relative, nnz counts exactly.
"""
import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceCSR, to_dev_u32, to_dev_u8
from cellshard_amd.serve import NNZ, SUM, SUMSQ, Server, make_query, materialize
from helpers import BF16, F16, FP8, decode, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _log1p(ptr, val):
    """log1p(count / row total x 10^4) per entry (0 stays 0)."""
    rows = len(ptr) - 1
    tot = np.add.reduceat(np.asarray(val, np.float64), ptr[:-1].astype(np.int64)) if len(val) else np.zeros(rows)
    tot[np.diff(ptr.astype(np.int64)) == 0] = 1.0
    per = np.repeat(tot, np.diff(ptr.astype(np.int64)))
    return np.log1p(np.asarray(val, np.float64) / np.maximum(per, 1.0) * 1e4)


def _shards(n=3, cols=700, seed=0, vt=F16, real=False):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n):
        rows = int(rng.integers(300, 1500))
        ptr, idx, val = random_csr(rows, cols, 0.03, seed=100 + s, zeros=0.02, empty_rows=0.05)
        vals = _log1p(ptr, val) if real else val
        if real and vt == FP8:  # nearest e4m3 value (the encoder takes representable values only)
            grid = np.unique(decode(np.arange(0x7F, dtype=np.uint8), FP8))
            i = np.clip(np.searchsorted(grid, vals), 1, len(grid) - 1)
            vals = np.where(vals - grid[i - 1] <= grid[i] - vals, grid[i - 1], grid[i])
        bits = encode(vals, vt)
        cptr, cidx, cval = O.transpose(rows, cols, ptr, idx, bits)
        csc = DeviceCSR.from_numpy(rows, cols, cptr, cidx, cval, vt, axis=A.COMPRESSED_BY_COL)
        out.append(dict(rows=rows, ptr=ptr, idx=idx, bits=bits, csc=csc, vt=vt))
    return out


def _queries(shards, n, cols, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for i in range(n):
        s = int(rng.integers(len(shards)))
        rows = shards[s]["rows"]
        k = int(np.exp(rng.uniform(np.log(1), np.log(cols))))
        feats = np.sort(rng.choice(cols, size=k, replace=False)).astype(np.uint32)
        G = int(np.exp(rng.uniform(np.log(1), np.log(600))))
        op = int(rng.integers(3))
        spec = dict(shard=s, op=op, G=G, feats=feats, label_seed=int(rng.integers(1 << 32)),
                    mask_seed=int(rng.integers(1 << 32)), frac=float(rng.uniform(0.05, 1.0)))
        if i % 5 == 0:  # explicit arrays instead of seeds; some labels >= G drop rows
            spec["labels"] = rng.integers(0, G + 2, size=rows).astype(np.uint32)
            spec["keep"] = (rng.random(rows) < spec["frac"]).astype(np.uint8)
        qs.append(spec)
    return qs


def _struct(spec, keepalive):
    f = to_dev_u32(spec["feats"])
    lab = to_dev_u32(spec["labels"]) if "labels" in spec else None
    kp = to_dev_u8(spec["keep"]) if "keep" in spec else None
    keepalive += [f, lab, kp]
    return make_query(spec["shard"], spec["op"], spec["G"], f, lab, kp, spec["label_seed"], spec["mask_seed"], spec["frac"])


def _oracle(shards, spec, cols):
    sh = shards[spec["shard"]]
    return O.serve_query(sh["rows"], cols, sh["ptr"], sh["idx"], sh["bits"], sh["vt"], spec["feats"], spec["G"], spec["op"],
                         labels=spec.get("labels"), keep=spec.get("keep"), label_seed=spec["label_seed"],
                         mask_seed=spec["mask_seed"], thr24=int(min(max(spec["frac"], 0), 1) * (1 << 24) + 0.5))


def test_serve_alone_mixed_and_oracle(gpu):
    cols = 700
    shards = _shards(cols=cols)
    specs = _queries(shards, 40, cols, seed=7)
    keepalive = []
    structs = [_struct(s, keepalive) for s in specs]
    torch.cuda.synchronize()

    # alone: one query per launch, one launch at a time
    alone = Server(max_slots=4, max_batch_queries=1, max_batches_in_flight=1, slot_bytes=cols * 600 * 8, budget_bytes=1 << 30)
    for sh in shards:
        alone.add_shard(sh["csc"])
    res_alone = []
    for spec, q in zip(specs, structs):
        t = alone.submit(q)
        alone.drain()
        r = alone.result(t)
        assert r["status"] == 2
        ref = _oracle(shards, spec, cols)
        assert np.array_equal(r["out"], ref), (spec["op"], spec["G"], len(spec["feats"]))
        res_alone.append(r["out"])
        alone.release(t)
    alone.close()

    # mixed: continuous batching, staggered arrivals, several launches in flight
    mixed = Server(max_slots=64, max_batch_queries=16, max_batches_in_flight=3, slot_bytes=cols * 600 * 8, budget_bytes=1 << 30)
    for sh in shards:
        mixed.add_shard(sh["csc"])
    arrivals = np.cumsum(np.random.default_rng(3).exponential(20.0, size=len(structs)))
    tickets, lat, wall = mixed.replay(structs, arrivals, hold=True)
    assert len(tickets) == len(structs) and np.all(lat >= 0) and wall > 0
    for t, ref in zip(tickets, res_alone):
        r = mixed.result(t)
        assert r["status"] == 2 and np.array_equal(r["out"], ref)
        mixed.release(t)
    mixed.close()


def test_serve_admission_and_errors(gpu):
    cols = 300
    shards = _shards(n=2, cols=cols, seed=4)
    need = shards[0]["csc"].format_bytes()
    # budget: the slots plus one shard, so the second registration fails
    srv = Server(max_slots=2, max_batch_queries=4, max_batches_in_flight=2, slot_bytes=4096, budget_bytes=2 * 4096 + need)
    srv.add_shard(shards[0]["csc"])
    with pytest.raises(Exception):
        srv.add_shard(shards[1]["csc"])
    f = to_dev_u32(np.arange(10, dtype=np.uint32))
    torch.cuda.synchronize()
    with pytest.raises(Exception):  # output 10 x 100 x 8 B > slot
        srv.submit(make_query(0, SUM, 100, f))
    with pytest.raises(Exception):  # unknown shard
        srv.submit(make_query(5, SUM, 4, f))
    # more queries than slots: they wait for slots and finish once results are released
    ts = [srv.submit(make_query(0, op, 16, f, label_seed=i, mask_seed=i + 1)) for i, op in enumerate([SUM, SUMSQ, NNZ, SUM])]
    for _ in range(100000):
        srv.pump()
        if all(srv.result(t, fetch=False)["status"] == 2 for t in ts[:2]):
            break
    assert [srv.result(t, fetch=False)["status"] for t in ts[2:]] == [0, 0]  # no slot yet
    sh = shards[0]
    got = {}
    for t in ts[:2]:
        got[t] = srv.result(t)["out"]
        srv.release(t)
        assert srv.result(t, fetch=False)["status"] == 3
    srv.drain()
    for t in ts[2:]:
        got[t] = srv.result(t)["out"]
    for i, t in enumerate(ts):
        op = [SUM, SUMSQ, NNZ, SUM][i]
        ref = O.serve_query(sh["rows"], cols, sh["ptr"], sh["idx"], sh["bits"], F16, np.arange(10), 16, op, label_seed=i,
                            mask_seed=i + 1)
        assert np.array_equal(got[t], ref)
    srv.close()


def test_serve_materialize_matches_oracle(gpu):
    lab, keep = materialize(5000, 37, 12345, 999, 0.3)
    torch.cuda.synchronize()
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), O.serve_label(12345, 5000, 37))
    assert np.array_equal(keep.cpu().numpy(), O.serve_keep(999, 5000, int(0.3 * (1 << 24) + 0.5)))


@pytest.mark.parametrize("vt", [BF16, FP8])
def test_serve_real_values_alone_vs_mixed(gpu, vt):
    cols = 500
    shards = _shards(n=3, cols=cols, seed=11, vt=vt, real=True)
    for sh in shards:  # the values are real: not all integers
        v = decode(sh["bits"], vt)
        assert np.any(v != np.round(v))
    specs = _queries(shards, 30, cols, seed=13)
    keepalive = []
    structs = [_struct(s, keepalive) for s in specs]
    torch.cuda.synchronize()

    def close(a, b, op):
        if op == NNZ:
            return np.array_equal(a, b)
        return np.allclose(a, b, rtol=1e-6, atol=1e-12)

    alone = Server(max_slots=4, max_batch_queries=1, max_batches_in_flight=1, slot_bytes=cols * 600 * 8, budget_bytes=1 << 30)
    for sh in shards:
        alone.add_shard(sh["csc"])
    res_alone = []
    for spec, q in zip(specs, structs):
        t = alone.submit(q)
        alone.drain()
        r = alone.result(t)
        assert r["status"] == 2
        assert close(r["out"], _oracle(shards, spec, cols), spec["op"]), (spec["op"], spec["G"], len(spec["feats"]))
        res_alone.append(r["out"])
        alone.release(t)
    alone.close()

    mixed = Server(max_slots=64, max_batch_queries=16, max_batches_in_flight=3, slot_bytes=cols * 600 * 8, budget_bytes=1 << 30)
    for sh in shards:
        mixed.add_shard(sh["csc"])
    arrivals = np.cumsum(np.random.default_rng(5).exponential(20.0, size=len(structs)))
    tickets, lat, wall = mixed.replay(structs, arrivals, hold=True)
    for t, ref, spec in zip(tickets, res_alone, specs):
        r = mixed.result(t)
        assert r["status"] == 2 and close(r["out"], ref, spec["op"])
        mixed.release(t)
    mixed.close()
