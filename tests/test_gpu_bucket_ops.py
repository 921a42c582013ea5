"""Workspace-owned major-nnz bucketing (bucket/operators/major_nnz.cuh:11-213)
and its shard form (sharded_major_nnz.cuh:137-273): the plan (stable
ascending nnz order, sorted nnz, equal-count bucket offsets) and the rebuilt
compressed arrays with the inverse order must equal the oracle's bit for bit
(oracle.bucket_plan / oracle.rebuild; the shard is checked against the
oracle over its partitions concatenated in order).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd._lib import last_error, load
from cellshard_amd.device import DeviceCSR
from cellshard_amd.ops import MajorBucketWorkspace, device_to_numpy
from cellshard_amd.sharded import Sharded, ShardedDevice, ShardMajorBucket, new_host_compressed
from helpers import F16, F32, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _u32(addr, n):
    return device_to_numpy(addr, n, np.uint32) if n else np.zeros(0, np.uint32)


def _check(result, ptr, idx, bits, requested, vdtype):
    n = len(ptr) - 1
    plan = O.bucket_plan(ptr, n, requested)
    p = result.plan
    assert (p.major_dim, p.bucket_count) == (n, plan["bucket_count"])
    assert np.array_equal(_u32(p.major_order, n), plan["major_order"])
    assert np.array_equal(_u32(p.major_nnz_sorted, n), plan["major_nnz_sorted"])
    assert np.array_equal(_u32(p.bucket_offsets, p.bucket_count + 1), plan["bucket_offsets"])
    dptr, didx, dval, inv = O.rebuild(ptr, idx, bits, plan["major_order"])
    r = result.rebuilt
    assert r.nnz == len(idx)
    assert np.array_equal(_u32(r.majorPtr, n + 1), dptr)
    if len(idx):
        assert np.array_equal(_u32(r.minorIdx, len(idx)), didx)
        assert np.array_equal(device_to_numpy(r.val, len(idx), vdtype), np.asarray(dval).view(vdtype))
    assert np.array_equal(_u32(p.inverse_major_order, n), inv)


def test_major_bucket_workspace_operators(gpu):
    ws = MajorBucketWorkspace()
    # plan, then rebuild, of a CSR with empty rows and repeated lengths (stable ties)
    ptr, idx, val = random_csr(700, 300, 0.05, seed=1, empty_rows=0.1)
    bits = encode(val, F16)
    src = DeviceCSR.from_numpy(700, 300, ptr, idx, bits)
    plan = ws.build_plan(src, 7)
    assert plan.inverse_major_order is None
    _check(ws.rebuild(src, plan), ptr, idx, bits, 7, np.uint16)
    # one call, f32 values, a larger matrix (the workspace grows) ...
    ptr2, idx2, val2 = random_csr(2500, 900, 0.03, seed=2, empty_rows=0.05)
    bits2 = encode(val2, F32)
    src2 = DeviceCSR.from_numpy(2500, 900, ptr2, idx2, bits2, value_type=A.VALUE_F32)
    _check(ws.build_bucketed_major_view(src2, 12), ptr2, idx2, bits2, 12, np.uint32)
    # ... then a smaller one, and a CSC (major axis = columns), more buckets than majors
    cptr, cidx, cval = O.transpose(700, 300, ptr, idx, bits)
    csc = DeviceCSR.from_numpy(700, 300, cptr, cidx, cval, axis=A.COMPRESSED_BY_COL)
    torch.cuda.synchronize()
    _check(ws.build_bucketed_major_view(csc, 1000), cptr, cidx, cval, 1000, np.uint16)
    # an empty matrix: one bucket, an all-zero row pointer
    empty = DeviceCSR.from_numpy(5, 4, np.zeros(6, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint16))
    r = ws.build_bucketed_major_view(empty, 3)
    torch.cuda.synchronize()
    assert r.rebuilt.nnz == 0 and not _u32(r.rebuilt.majorPtr, 6).any()
    ws.close()


def _parts(shapes, cols, seed):
    out = []
    for k, rows in enumerate(shapes):
        ptr, idx, val = random_csr(rows, cols, 0.08, seed=seed + k, empty_rows=0.1) if rows else (
            np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
        out.append((rows, np.asarray(ptr, np.uint32), np.asarray(idx, np.uint32), encode(val, F16)))
    return out


def _concat(parts):
    ptrs, idxs, vals, base = [np.zeros(1, np.uint32)], [], [], 0
    for rows, ptr, idx, bits in parts:
        ptrs.append((ptr[1:].astype(np.int64) + base).astype(np.uint32))
        idxs.append(idx)
        vals.append(bits)
        base += len(idx)
    return np.concatenate(ptrs), np.concatenate(idxs), np.concatenate(vals)


def test_shard_major_bucket_view(gpu):
    cols = 60
    parts = _parts([300, 0, 120, 500, 77], cols, seed=20)
    m = Sharded(A.FORMAT_COMPRESSED)
    for rows, ptr, idx, bits in parts:
        m.append_partition(new_host_compressed(rows, cols, ptr, idx, bits))
    assert m.reshard([0, 420, 997])  # shard 0: partitions 0-2 (one empty), shard 1: partitions 3-4
    st = ShardedDevice(len(parts))
    st.upload_shard(m, 0)  # one packed allocation
    st.upload_partition(m, 3)  # shard 1 partition by partition
    st.upload_partition(m, 4)
    torch.cuda.synchronize()
    sb = ShardMajorBucket()
    for shard, members, requested in ((0, parts[0:3], 4), (1, parts[3:5], 9)):
        r = sb.build(st, m, shard, requested)
        torch.cuda.synchronize()
        assert r.shard_id == shard and r.rebuilt.cols == cols and r.rebuilt.axis == A.COMPRESSED_BY_ROW
        ptr, idx, bits = _concat(members)
        _check(r, ptr, idx, bits, requested, np.uint16)
    # a shard whose partitions are not resident is refused
    st.release_shard(m, 1)
    out = A.sharded_major_bucket_result()
    assert load().csamd_build_bucketed_shard_major_view(C.byref(st.s), C.byref(m.m), 1, 4, C.byref(sb.ws), C.byref(out)) == 0
    assert "resident" in last_error()
    sb.close()
    st.close()
    m.close()
