"""Staging -> conversion -> reduce of batched partitions in one captured graph
(csamd_label_reduce_pipeline_capture).

Partitions are chosen so that the layout tuner picks each layout (dense
8-column blocks: Blocked-ELL; equal row lengths: Sliced-ELL; ragged rows:
CSR), plus random partitions forced to each converted layout, one of them
with unsorted and repeated columns (the Blocked-ELL replay path).  Every
replay re-stages the CSRs from pinned host copies, converts on the device
and reduces; checked against the oracle:
  - each Sliced-ELL and Blocked-ELL output bit for bit against
    sliced_ell_from_compressed / blocked_ell_from_compressed (nnz included);
  - the reduce over the concatenated partitions: counts exact, sums within
    1e-6 relative of FP64 (test_gpu_label_reduce.check);
  - after the pinned host values of one partition change, the next replay
    follows them (the staging is part of the graph).
"""
import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd.device import DeviceCSR, to_dev_u32, to_dev_u8
from cellshard_amd.ops import ConvertWorkspace, LabelReduce, LabelReducePipeline, tune_partitions
from helpers import F16, encode, random_csr
from oracle import oracle as O
from test_gpu_batched import _concat
from test_gpu_label_reduce import check

pytestmark = pytest.mark.gpu


def _cases(cols):
    rows = 64
    rng = np.random.default_rng(4)
    dense_idx = np.tile(np.arange(8, dtype=np.uint32), rows)
    dense_ptr = np.arange(0, 8 * rows + 1, 8, dtype=np.uint32)
    eq_idx = np.concatenate([np.sort(rng.choice(cols, 20, replace=False)) for _ in range(rows)]).astype(np.uint32)
    eq_ptr = np.arange(0, 20 * rows + 1, 20, dtype=np.uint32)
    rag_ptr, rag_idx, _ = random_csr(rows, cols, 0.05, seed=1)
    lens = np.diff(rag_ptr.astype(np.int64))
    lens[0] = 200
    rag_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    rag_idx = np.concatenate([np.sort(rng.choice(cols, int(n), replace=False)) for n in lens]).astype(np.uint32)
    out = [(rows, dense_ptr, dense_idx), (rows, eq_ptr, eq_idx), (rows, rag_ptr, rag_idx)]
    for k, n in enumerate((300, 77, 150)):
        ptr, idx, _ = random_csr(n, cols, 0.04, seed=20 + k, empty_rows=0.1, duplicates=0.05 if k == 1 else 0.0,
                                 sorted_rows=k != 1)
        out.append((n, np.asarray(ptr, np.uint32), np.asarray(idx, np.uint32)))
    return out


def test_pipeline_stage_convert_reduce(gpu):
    cols, G = 256, 7
    host = []
    for k, (rows, ptr, idx) in enumerate(_cases(cols)):
        vals = np.random.default_rng(30 + k).choice([0.0, 1.0, 2.0, 3.0, 5.0], size=len(idx), p=[0.05, 0.5, 0.2, 0.15, 0.1])
        host.append((rows, ptr, idx, encode(vals, F16)))
    parts = [DeviceCSR.from_numpy(r, cols, p, i, b) for r, p, i, b in host]
    ws = ConvertWorkspace()
    tune = tune_partitions(parts, ws)
    assert [t["layout"] for t in tune[:3]] == ["blocked_ell", "sliced_ell", "compressed"]
    for t, forced in zip(tune[3:], ("sliced_ell", "blocked_ell", "blocked_ell")):
        t["layout"] = forced
    tune[5]["bell_block_size"], tune[5]["bell_ell_width"] = 16, None  # re-derived below
    # the ell width of a forced block size: the most block columns of any row block
    rows5, ptr5, idx5, _ = host[5]
    _bci, _val, ell5, _nnz = O.blocked_ell_from_compressed(rows5, cols, ptr5, idx5, host[5][3], 16)
    tune[5]["bell_ell_width"] = ell5 // 16
    rows_t, ptr_t, idx_t, bits_t = _concat(host)
    rng = np.random.default_rng(6)
    labels = rng.integers(0, G + 1, size=rows_t).astype(np.uint32)
    rk = (rng.random(rows_t) < 0.85).astype(np.uint8)
    fk = (rng.random(cols) < 0.6).astype(np.uint8)
    d_labels, d_rk, d_fk = to_dev_u32(labels), to_dev_u8(rk), to_dev_u8(fk)
    pipe = LabelReducePipeline(parts, tune, d_labels, G, d_rk, d_fk, staged=True)
    # the device CSRs are restaged from the pinned copies by every launch: clear them first
    for p in parts:
        p.minor_idx.zero_()
        p.val.zero_()
    stream = torch.cuda.current_stream().cuda_stream
    ref = O.label_reduce_csr(rows_t, cols, ptr_t, idx_t, bits_t, F16, labels, rk, fk, G)
    for rep in range(3):  # replays with eager work in between
        pipe.launch(stream)
        got = LabelReduce.fetch(pipe.result())
        check(got, ref)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()
    for i, t in enumerate(tune):
        rows, ptr, idx, bits = host[i]
        if t["layout"] == "sliced_ell":
            s = pipe.sell(i)
            want = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, s.slice_rows)
            assert np.array_equal(s.slice_row_offsets.cpu().numpy().view(np.uint32), want["offsets"])
            assert np.array_equal(s.slice_widths.cpu().numpy().view(np.uint32), want["widths"])
            assert np.array_equal(s.slice_slot_offsets.cpu().numpy().view(np.uint32), want["slot_offsets"])
            slots = len(want["col_idx"])
            assert np.array_equal(s.col_idx.cpu().numpy().view(np.uint32)[:slots], want["col_idx"])
            assert np.array_equal(s.val.cpu().numpy().view(np.uint16)[:slots], np.asarray(want["val"]).view(np.uint16))
        elif t["layout"] == "blocked_ell":
            b = pipe.bell(i)
            bci, val, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, b.block_size)
            assert (b.ell_cols, b.nnz) == (ell, nnz), i
            assert np.array_equal(b.block_col_idx.cpu().numpy().view(np.uint32)[:len(bci)], np.asarray(bci, np.uint32)), i
            assert np.array_equal(b.val.cpu().numpy().view(np.uint16)[:rows * ell], np.asarray(val).view(np.uint16)), i
    # new values in a pinned staging source reach the next replay
    hv = pipe._host[3][2]
    hv.copy_(torch.from_numpy(encode(np.full(hv.numel(), 2.0), F16).view(np.int16)))
    host[3] = (host[3][0], host[3][1], host[3][2], encode(np.full(len(host[3][2]), 2.0), F16))
    rows_t, ptr_t, idx_t, bits_t = _concat(host)
    pipe.launch(stream)
    check(LabelReduce.fetch(pipe.result()), O.label_reduce_csr(rows_t, cols, ptr_t, idx_t, bits_t, F16, labels, rk, fk, G))
    pipe.close()
    ws.close()
