"""repack_sharded_compressed_to_blocked_ell(_auto)
(convert/blocked_ell_from_compressed.cuh:480-600): a sharded CSR becomes a
sharded Blocked-ELL whose parts hold target_block_rows x block_size rows
each (the last the rest).  Every part must equal the oracle's
blocked_ell_from_compressed of the same rows (stitched across source
partitions, including an empty one and parts that cut through partitions),
bit for bit; the automatic variant must pick the oracle tuner's block size
over the whole matrix.
"""
import numpy as np
import pytest

from cellshard_amd import abi as A
from cellshard_amd.sharded import Sharded, new_host_compressed, repack_to_blocked_ell
from helpers import F16, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _arr(addr, n, dtype):
    import ctypes as C
    if n == 0:
        return np.zeros(0, dtype)
    return np.frombuffer((C.c_char * (n * np.dtype(dtype).itemsize)).from_address(addr), dtype=dtype).copy()


def _source(shapes, cols, seed):
    m = Sharded(A.FORMAT_COMPRESSED)
    ptrs, idxs, vals, base = [np.zeros(1, np.int64)], [], [], 0
    for k, rows in enumerate(shapes):
        if rows:
            ptr, idx, val = random_csr(rows, cols, 0.06, seed=seed + k, empty_rows=0.1, duplicates=0.0)
        else:
            ptr, idx, val = np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0)
        bits = encode(val, F16)
        m.append_partition(new_host_compressed(rows, cols, np.asarray(ptr, np.uint32), np.asarray(idx, np.uint32), bits))
        ptrs.append(np.asarray(ptr[1:], np.int64) + base)
        idxs.append(np.asarray(idx, np.uint32))
        vals.append(bits)
        base += len(idx)
    whole = (np.concatenate(ptrs).astype(np.uint32), np.concatenate(idxs), np.concatenate(vals))
    return m, whole


def _check_parts(dst, whole, rows_total, cols, bs, rows_per_part):
    ptr, idx, bits = whole
    assert dst.m.cols == cols and dst.m.rows == rows_total
    r0 = 0
    for p in range(dst.num_partitions):
        r1 = min(rows_total, r0 + rows_per_part) if rows_per_part else rows_total
        h = dst.part(p)
        sub_ptr = (ptr[r0:r1 + 1].astype(np.int64) - ptr[r0]).astype(np.uint32)
        sub_idx, sub_val = idx[ptr[r0]:ptr[r1]], bits[ptr[r0]:ptr[r1]]
        bci, val, ell, nnz = O.blocked_ell_from_compressed(r1 - r0, cols, sub_ptr, sub_idx, sub_val, bs)
        assert (h.rows, h.cols, h.block_size, h.ell_cols, h.nnz) == (r1 - r0, cols, bs, ell, nnz)
        rb, w = -(-h.rows // bs), ell // bs
        assert np.array_equal(_arr(h.blockColIdx, rb * w, np.uint32), np.asarray(bci, np.uint32))
        assert np.array_equal(_arr(h.val, h.rows * ell, np.uint16), np.asarray(val).view(np.uint16))
        r0 = r1
    assert r0 == rows_total


def test_repack_fixed_block_size(gpu):
    cols = 90
    shapes = [37, 0, 64, 15, 100]
    m, whole = _source(shapes, cols, seed=3)
    for bs, target in ((4, 5), (8, 3), (16, 0)):  # 20-row parts, 24-row parts, one part
        dst, picked = repack_to_blocked_ell(m, block_size=bs, target_block_rows=target)
        assert picked is None
        assert dst.num_partitions == (-(-sum(shapes) // (bs * target)) if target else 1)
        _check_parts(dst, whole, sum(shapes), cols, bs, bs * target)
        dst.close()
    m.close()


def test_repack_auto_picks_whole_matrix_block_size(gpu):
    cols = 120
    shapes = [50, 33, 0, 81]
    m, whole = _source(shapes, cols, seed=9)
    cand = [2, 4, 8, 16]
    dst, tune = repack_to_blocked_ell(m, target_block_rows=4, candidates=cand)
    want_bs, _fill, _pad = O.choose_blocked_ell_block_size(sum(shapes), cols, whole[0], whole[1], cand)
    assert tune.block_size == want_bs
    _check_parts(dst, whole, sum(shapes), cols, tune.block_size, tune.block_size * 4)
    dst.close()
    m.close()
