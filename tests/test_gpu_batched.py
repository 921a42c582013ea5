"""Batched partitions: one reduce over uneven CSR / Sliced-ELL partitions
equals the FP64 oracle on the row-wise concatenation of the partitions.

Tolerance: counts bit-exact; sum / sumsq within 1e-6 relative of the FP64
oracle (FP64 partials: observed ~1e-15; integer-valued data is exact).
"""
import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceCSR, DeviceSlicedELL, to_dev_f32, to_dev_u32, to_dev_u8
from cellshard_amd.ops import LabelReduce
from helpers import F16, F32, FP8, encode, random_csr
from oracle import oracle as O
from test_gpu_label_reduce import check

pytestmark = pytest.mark.gpu


@pytest.fixture()
def lr(gpu):
    r = LabelReduce()
    yield r
    r.close()


def _concat(parts):
    """Row-wise concatenation of (rows, ptr, idx, bits) partitions."""
    ptrs, idxs, vals, off = [np.zeros(1, np.uint64)], [], [], 0
    for rows, ptr, idx, bits in parts:
        ptrs.append(ptr[1:].astype(np.uint64) + off)
        off += int(ptr[-1])
        idxs.append(idx)
        vals.append(bits)
    return (sum(p[0] for p in parts), np.concatenate(ptrs).astype(np.uint32), np.concatenate(idxs).astype(np.uint32),
            np.concatenate(vals))


def _sell(rows, cols, ptr, idx, bits, S, vt=F16):
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, S)
    return DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                      s["col_idx"], s["val"], vt)


def _partitions(cols, vt, seed, sizes, kinds):
    host, dev = [], []
    for i, (rows, kind) in enumerate(zip(sizes, kinds)):
        ptr, idx, val = random_csr(rows, cols, 0.08, seed=seed + i, fp8_safe=(vt == FP8), zeros=0.02, empty_rows=0.1)
        bits = encode(val, vt)
        host.append((rows, ptr, idx, bits))
        if kind == "csr":
            dev.append(DeviceCSR.from_numpy(rows, cols, ptr, idx, bits, vt))
        else:
            dev.append(_sell(rows, cols, ptr, idx, bits, kind, vt))
    return host, dev


@pytest.mark.parametrize("vt", [F16, F32, FP8])
def test_batched_mixed_partitions(lr, vt):
    cols, G = 500, 9
    sizes = [37, 0, 300, 1, 129, 64, 250, 7]
    kinds = ["csr", "csr", 32, "csr", 64, 8, "csr", 32]
    host, dev = _partitions(cols, vt, 10 * vt, sizes, kinds)
    rows, ptr, idx, bits = _concat(host)
    rng = np.random.default_rng(5)
    labels = rng.integers(0, G + 1, size=rows).astype(np.uint32)
    rk = (rng.random(rows) < 0.85).astype(np.uint8)
    fk = (rng.random(cols) < 0.5).astype(np.uint8)
    scale = rng.uniform(0.5, 2.0, cols).astype(np.float32) if vt == FP8 else None
    res = lr.run_batched(dev, to_dev_u32(labels), G, to_dev_u8(rk), to_dev_u8(fk),
                         to_dev_f32(scale) if scale is not None else None)
    got = LabelReduce.fetch(res)
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, vt, labels, rk, fk, G, scale)
    check(got, ref)


def test_batched_one_partition_equals_single_call(lr):
    rows, cols, G = 3000, 800, 16
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=3)
    bits = encode(val, F16)
    src = DeviceCSR.from_numpy(rows, cols, ptr, idx, bits)
    rng = np.random.default_rng(8)
    labels = to_dev_u32(rng.integers(0, G, size=rows).astype(np.uint32))
    rk = to_dev_u8((rng.random(rows) < 0.8).astype(np.uint8))
    single = LabelReduce.fetch(lr.run(src, labels, G, rk))
    batched = LabelReduce.fetch(lr.run_batched([src], labels, G, rk))
    for k in single:  # integer counts: FP64 sums are exact, so results are identical
        assert np.array_equal(np.asarray(single[k]), np.asarray(batched[k])), k


def test_batched_unaligned_and_tail_partitions(lr):
    """Partition arrays at offsets that are not 16-byte aligned take the
    scalar-load path; aligned partitions with nnz % 4 != 0 take tail records."""
    cols, G = 300, 4
    host, dev = [], []
    for i, (rows, shift) in enumerate([(50, 0), (61, 1), (33, 0), (80, 3)]):
        ptr, idx, val = random_csr(rows, cols, 0.06, seed=70 + i)
        bits = encode(val, F16)
        host.append((rows, ptr, idx, bits))
        d = DeviceCSR.from_numpy(rows, cols, ptr, np.concatenate([np.zeros(shift, np.uint32), idx]),
                                 np.concatenate([np.zeros(shift, bits.dtype), bits]))
        dev.append(DeviceCSR(rows, cols, d.major_ptr, d.minor_idx[shift:], d.val[shift:], F16))
    rows, ptr, idx, bits = _concat(host)
    labels = np.random.default_rng(1).integers(0, G, size=rows).astype(np.uint32)
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, None, None, G)
    check(LabelReduce.fetch(lr.run_batched(dev, to_dev_u32(labels), G)), ref)  # mixed: scalar path for all
    check(LabelReduce.fetch(lr.run_batched([dev[0], dev[2]], to_dev_u32(np.concatenate([labels[:50], labels[111:144]])), G)),
          O.label_reduce_csr(83, cols, *_concat([host[0], host[2]])[1:], F16,
                             np.concatenate([labels[:50], labels[111:144]]), None, None, G))


def test_batched_rejects_bad_partitions(lr):
    a = DeviceCSR.from_numpy(3, 10, np.array([0, 1, 2, 3]), np.array([0, 1, 2]), encode(np.ones(3), F16))
    b = DeviceCSR.from_numpy(3, 11, np.array([0, 1, 2, 3]), np.array([0, 1, 2]), encode(np.ones(3), F16))
    labels = to_dev_u32(np.zeros(6, np.uint32))
    with pytest.raises(Exception):
        lr.run_batched([a, b], labels, 2)  # different column counts
    csc = DeviceCSR(3, 10, a.major_ptr, a.minor_idx, a.val, F16, A.COMPRESSED_BY_COL)
    with pytest.raises(Exception):
        lr.run_batched([a, csc], labels, 2)  # column-major partition
    lr.run_batched([a, a], labels, 2)  # the workspace still works after the errors
    torch.cuda.synchronize()


def test_graph_replays_match_eager(lr):
    """The captured batched call, replayed several times with eager work and
    synchronizations between the replays, equals the eager call every time;
    a replay after the label array changed equals the eager call on the new
    labels (the graph reads the bound arrays' current contents)."""
    from cellshard_amd.ops import LabelReduceGraph
    cols, G = 400, 12
    sizes = [90, 333, 0, 45, 210] + [17] * 30  # 35 partitions: descriptor table through staging
    kinds = ["csr", 32, "csr", 64, "csr"] + ["csr", 32] * 15
    host, dev = _partitions(cols, F16, 500, sizes, kinds)
    rows = sum(sizes)
    rng = np.random.default_rng(9)
    labels = to_dev_u32(rng.integers(0, G, size=rows).astype(np.uint32))
    rk = to_dev_u8((rng.random(rows) < 0.9).astype(np.uint8))
    fk = to_dev_u8((rng.random(cols) < 0.6).astype(np.uint8))
    for parts in (dev, dev[:5]):  # staged table and inline descriptors
        eager = LabelReduce.fetch(lr.run_batched(parts, labels, G, rk, fk))
        graph = LabelReduceGraph(parts, labels, G, rk, fk)
        for i in range(4):
            graph.launch()
            got = LabelReduce.fetch(graph.result())
            for k in eager:
                assert np.array_equal(np.asarray(eager[k]), np.asarray(got[k])), (i, k)
            # eager work between replays: another reduce, torch kernels, a sync
            other = LabelReduce.fetch(lr.run_batched(parts[::-1], labels, G))
            assert other["kept_rows"] == sum(p.rows for p in parts)  # labels < G, no row mask
            assert float(torch.arange(1000, device="cuda", dtype=torch.float64).sum()) == 499500.0
            torch.cuda.synchronize()
        new_labels = rng.integers(0, G, size=rows).astype(np.uint32)
        labels.copy_(to_dev_u32(new_labels))
        graph.launch()
        got = LabelReduce.fetch(graph.result())
        ref = LabelReduce.fetch(lr.run_batched(parts, labels, G, rk, fk))
        for k in ref:
            assert np.array_equal(np.asarray(ref[k]), np.asarray(got[k])), k
        graph.close()
