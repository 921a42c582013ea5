"""Batched layout autotuner vs the oracle's restatement of the reference's
tuners (Blocked-ELL block size among {8, 16, 32}, Sliced-ELL slice height
among {32, 64}), per partition; the row-length histogram vs numpy.  All
integer / exact: compared bit for bit (fill and score to 1e-12)."""
import numpy as np
import pytest

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceCSR
from cellshard_amd.ops import ConvertWorkspace, tune_partitions
from helpers import F16, F32, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _expect(rows, cols, ptr, idx, vb):
    bs, fill, _ = O.choose_blocked_ell_block_size(rows, cols, ptr, idx, [8, 16, 32], vb)
    w = O.blocked_ell_shape(rows, cols, ptr, idx, bs)[0] if rows and cols else 0
    rn = np.diff(ptr.astype(np.int64)).astype(np.uint32)
    sell = O.choose_sliced_ell_slice_rows(rn, rows, len(idx), [32, 64])
    hist = np.zeros(32, np.int64)
    b = np.where(rn == 0, 0, np.floor(np.log2(np.maximum(rn, 1))).astype(np.int64) + 1)
    np.add.at(hist, np.minimum(b, 31), 1)
    return bs, fill, w, sell, hist


@pytest.mark.parametrize("vt", [F16, F32])
def test_tuner_matches_reference_rules(gpu, vt):
    vb = 2 if vt == F16 else 4
    ws = ConvertWorkspace()
    cases = []
    rng = np.random.default_rng(4)
    cols = 1003  # not a multiple of 8, 16 or 32
    for i, rows in enumerate([0, 5, 31, 32, 33, 63, 64, 65, 200, 1000, 257]):
        ptr, idx, val = random_csr(rows, cols, float(rng.uniform(0.002, 0.08)), seed=200 + i, zeros=0.05, duplicates=0.02,
                                   empty_rows=0.1)
        if len(idx) > 3:
            idx[rng.integers(len(idx))] = cols + 7  # out-of-range column: ignored by the shape count
        cases.append((rows, ptr, idx, encode(val, vt)))
    dev = [DeviceCSR.from_numpy(r, cols, p, i, b, vt) for r, p, i, b in cases]
    got = tune_partitions(dev, ws)
    for (rows, ptr, idx, _), g in zip(cases, got):
        bs, fill, w, sell, hist = _expect(rows, cols, ptr, idx, vb)
        assert g["bell_block_size"] == bs and g["bell_ell_width"] == w, (rows, g, bs, w)
        assert abs(g["bell_fill_ratio"] - fill) <= 1e-12
        assert g["sell_slice_rows"] == sell["slice_rows"] and g["sell_padded_slots"] == sell["padded_slots"], (rows, g, sell)
        assert g["sell_slice_count"] == sell["slice_count"]
        assert abs(g["sell_score"] - sell["score"]) <= 1e-12 or rows == 0
        assert np.array_equal(np.asarray(g["row_length_log2_hist"]), hist)
        nnz = len(idx)
        csr = (rows + 1) * 4 + nnz * (4 + vb)
        sellb = sell["padded_slots"] * (4 + vb) + (2 * sell["slice_count"] + 1) * 4
        bellb = ((rows + bs - 1) // bs) * w * 4 + rows * w * bs * vb
        assert (g["csr_bytes"], g["sell_bytes"], g["bell_bytes"]) == (csr, sellb, bellb)
        best = min(csr, sellb, bellb)
        want = "compressed" if csr == best else "sliced_ell" if sellb == best else "blocked_ell"
        assert g["layout"] == want
    ws.close()


def test_tuner_picks_each_layout(gpu):
    """Dense 8-column blocks favour Blocked-ELL; equal row lengths favour
    Sliced-ELL (no row pointer per row); ragged rows favour CSR."""
    ws = ConvertWorkspace()
    cols = 256
    rows = 64
    dense_idx = np.tile(np.arange(8, dtype=np.uint32), rows)
    dense_ptr = np.arange(0, 8 * rows + 1, 8, dtype=np.uint32)
    eq_idx = np.concatenate([np.sort(np.random.default_rng(r).choice(cols, 20, replace=False)) for r in range(rows)]).astype(np.uint32)
    eq_ptr = np.arange(0, 20 * rows + 1, 20, dtype=np.uint32)
    rag_ptr, rag_idx, _ = random_csr(rows, cols, 0.05, seed=1)
    rag_ptr[1:] = np.maximum.accumulate(rag_ptr[1:])
    lens = np.diff(rag_ptr.astype(np.int64))
    lens[0] = 200  # one long row: Sliced-ELL pads its slice to 200
    rag_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    rag_idx = np.concatenate([np.sort(np.random.default_rng(100 + r).choice(cols, int(n), replace=False)) for r, n in
                              enumerate(lens)]).astype(np.uint32)
    parts = [DeviceCSR.from_numpy(rows, cols, p, i, encode(np.ones(len(i)), F16)) for p, i in
             [(dense_ptr, dense_idx), (eq_ptr, eq_idx), (rag_ptr, rag_idx)]]
    got = tune_partitions(parts, ws)
    assert [g["layout"] for g in got] == ["blocked_ell", "sliced_ell", "compressed"]
    ws.close()


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_tuner_on_index_arrays_off_a_16_byte_boundary(gpu, shift):
    """Partition views into the middle of an allocation: the index array
    starts 4, 8 or 12 bytes past a 16-byte boundary, so the 16-byte unit
    reads start at a different entry of every row."""
    import torch

    from cellshard_amd.device import to_dev_u32
    ws = ConvertWorkspace()
    rng = np.random.default_rng(shift)
    cols = 777
    cases, dev = [], []
    for i, rows in enumerate([1, 7, 64, 130, 333]):
        ptr, idx, val = random_csr(rows, cols, float(rng.uniform(0.01, 0.2)), seed=50 + i, duplicates=0.02, empty_rows=0.1,
                                   sorted_rows=bool(i % 2))
        cases.append((rows, ptr, idx))
        big = to_dev_u32(np.concatenate([np.full(shift, 5, np.uint32), idx]), torch.device("cuda", 0))
        d = DeviceCSR.from_numpy(rows, cols, ptr, idx, encode(val, F16), F16)
        d.minor_idx = big[shift:]
        d._keep = big
        dev.append(d)
    got = tune_partitions(dev, ws)
    for (rows, ptr, idx), g in zip(cases, got):
        bs, fill, w, sell, hist = _expect(rows, cols, ptr, idx, 2)
        assert g["bell_block_size"] == bs and g["bell_ell_width"] == w, (rows, g, bs, w)
        assert abs(g["bell_fill_ratio"] - fill) <= 1e-12
        assert np.array_equal(np.asarray(g["row_length_log2_hist"]), hist)
    ws.close()
