"""sharded_device<MatrixT> residency on the GPU (runtime/device/sharded_device.cuh):
partition and packed shard uploads, release through the allocation cache,
stage / swap, device bytes; and staging a shard whose host parts are not
materialized from its CSPACK01 pack, rebuilt in HBM.

Every resident partition is read back through its device view copy and
compared bit for bit with the host part it came from (or, for pack-staged
partitions, with the canonical Blocked-ELL the pack was written from).
"""
import ctypes as C
import os
import tempfile
from pathlib import Path

import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd._lib import load
from cellshard_amd.ops import (ConvertWorkspace, CspackFile, CspackWorkspace, blocked_ell_from_compressed,
                               build_bucketed_blocked_ell_device, device_to_numpy, synth_csr, write_cspack_blocked_ell)
from cellshard_amd.sharded import Sharded, ShardedDevice, host_part, new_host_compressed
from helpers import f16_bits, random_csr

pytestmark = pytest.mark.gpu


def _view(addr, cls):
    return cls.from_buffer_copy(device_to_numpy(addr, C.sizeof(cls), np.uint8).tobytes())


def _host_arr(addr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.frombuffer((C.c_char * (n * np.dtype(dtype).itemsize)).from_address(addr), dtype=dtype).copy()


def _csr_parts(shapes, seed=0):
    out = []
    for k, (rows, cols) in enumerate(shapes):
        ptr, idx, val = random_csr(rows, cols, 0.2, seed=seed + k, empty_rows=0.1)
        out.append((rows, cols, ptr, idx, f16_bits(val)))
    return out


def _check_compressed_record(rec, want):
    rows, cols, ptr, idx, bits = want
    v = _view(rec.view, A.compressed_view)
    assert (v.rows, v.cols, v.nnz) == (rows, cols, len(idx))
    assert np.array_equal(device_to_numpy(v.majorPtr, rows + 1, np.uint32), ptr)
    if len(idx):
        assert np.array_equal(device_to_numpy(v.minorIdx, len(idx), np.uint32), idx)
        assert np.array_equal(device_to_numpy(v.val, len(idx), np.uint16), bits.view(np.uint16))


def test_compressed_partitions_and_shards(gpu):
    parts = _csr_parts([(40, 30), (0, 30), (25, 30), (60, 30)])
    m = Sharded(A.FORMAT_COMPRESSED)
    for p in parts:
        m.append_partition(new_host_compressed(*p))
    assert m.reshard([0, 40, 125])  # shard 0: partition 0; shard 1: partitions 1-3
    st = ShardedDevice(m.num_partitions)
    # one partition: its own allocation, exactly the device bytes
    st.upload_partition(m, 0)
    r0 = st.record(0)
    assert r0.device_id == 0 and (r0.group_begin, r0.group_end) == (0, 1)
    assert r0.allocation_bytes == m.device_partition_bytes(0)
    _check_compressed_record(r0, parts[0])
    a0 = r0.allocation
    st.release_partition(0)
    assert st.record(0).view is None and st.s.cache_count == 1
    st.upload_partition(m, 0)  # the cached block comes back
    assert st.record(0).allocation == a0 and st.s.cache_count == 0
    # a packed shard: one allocation owned by its first partition
    st.upload_shard(m, 1)
    for p in (1, 2, 3):
        r = st.record(p)
        assert (r.group_begin, r.group_end) == (1, 4) and (r.allocation is not None) == (p == 1)
        _check_compressed_record(r, parts[p])
    assert st.record(1).allocation_bytes <= m.device_shard_bytes(1)
    # releasing one partition of the shard releases the group
    st.release_partition(3)
    assert all(st.record(p).view is None for p in (1, 2, 3)) and st.s.cache_count == 1
    # swap: shard 1 in (from the host parts), shard 0 out; drop the outgoing host part
    st.swap_shard(m, 0, 1, drop_host_after_release=True)
    assert st.record(0).view is None and all(st.record(p).view is not None for p in (1, 2, 3))
    assert m.m.parts[0] is None
    _check_compressed_record(st.record(2), parts[2])
    st.release_shard(m, 1)
    with pytest.raises(Exception):
        st.stage_shard(m, 0)  # partition 0 is gone from the host and there is no pack
    st.close()
    m.close()


def test_blocked_and_sliced_shards(gpu):
    lib = load()
    parts = _csr_parts([(33, 50), (16, 50), (7, 50)], seed=5)
    for fmt in (A.FORMAT_BLOCKED_ELL, A.FORMAT_SLICED_ELL):
        m = Sharded(fmt)
        hosts = []
        for rows, cols, ptr, idx, bits in parts:
            src = A.host_compressed(rows=rows, cols=cols, nnz=len(idx), axis=0)
            keep = [np.ascontiguousarray(ptr, np.uint32), np.ascontiguousarray(idx, np.uint32), np.ascontiguousarray(bits)]
            src.majorPtr, src.minorIdx, src.val = (a.ctypes.data for a in keep)
            addr = lib.csamd_host_part_alloc(fmt)
            if fmt == A.FORMAT_BLOCKED_ELL:
                assert lib.csamd_host_blocked_ell_from_compressed(C.byref(src), 4, A.host_blocked_ell.from_address(addr)) == 1
            else:
                assert lib.csamd_host_sliced_ell_from_compressed(C.byref(src), 8, A.host_sliced_ell.from_address(addr)) == 1
            hosts.append(host_part(fmt, addr))
            m.append_partition(addr)
        m.set_shards_by_device_bytes(sum(m.device_partition_bytes(p) for p in range(2)))
        assert m.shard_parts() == [0, 2, 3]
        st = ShardedDevice(3)
        st.upload_shard(m, 0)
        for p in (0, 1):
            r, h = st.record(p), hosts[p]
            if fmt == A.FORMAT_BLOCKED_ELL:
                v = _view(r.view, A.blocked_ell_view)
                rb, w = -(-h.rows // h.block_size), h.ell_cols // h.block_size
                assert (v.rows, v.block_size, v.ell_cols, v.nnz) == (h.rows, h.block_size, h.ell_cols, h.nnz)
                assert np.array_equal(device_to_numpy(v.blockColIdx, rb * w, np.uint32), _host_arr(h.blockColIdx, rb * w, np.uint32))
                assert np.array_equal(device_to_numpy(v.val, h.rows * h.ell_cols, np.uint16),
                                      _host_arr(h.val, h.rows * h.ell_cols, np.uint16))
                assert (r.group_begin, r.group_end) == (0, 2)
            else:
                v = _view(r.view, A.sliced_ell_view)
                offs = _host_arr(h.slice_row_offsets, h.slice_count + 1, np.uint32)
                wid = _host_arr(h.slice_widths, h.slice_count, np.uint32)
                heights = np.diff(offs.astype(np.int64))
                slots = int((heights * wid).sum())
                assert v.slice_count == h.slice_count and v.slice_rows == (8 if h.rows > 8 else h.rows)
                assert np.array_equal(device_to_numpy(v.slice_slot_offsets, h.slice_count, np.uint32),
                                      np.concatenate([[0], np.cumsum(heights * wid)[:-1]]).astype(np.uint32))
                assert np.array_equal(device_to_numpy(v.col_idx, slots, np.uint32), _host_arr(h.col_idx, slots, np.uint32))
                assert (r.group_begin, r.group_end) == (p, p + 1)  # uploaded one partition at a time
                assert r.allocation_bytes == m.device_partition_bytes(p)
        st.swap_shard(m, 0, 1)
        assert st.record(2).view is not None and st.record(0).view is None
        st.close()
        m.close()


def test_stage_shard_from_pack_rebuilds_canonical_partitions(gpu):
    ws = ConvertWorkspace()
    bells, parts = [], []
    for k, rows in enumerate((300, 129, 64)):
        src = synth_csr(rows, 900, median=60.0, sigma=0.5, seed=30 + k, value_type=A.VALUE_F16, ws=ws)
        bell = blocked_ell_from_compressed(src, 8, ws)
        bells.append(bell)
        parts.append(build_bucketed_blocked_ell_device(bell, 3, ws))
    cws = CspackWorkspace()
    tmp = Path(tempfile.mkdtemp(prefix="stage_"))
    path = tmp / "shard.cspack"
    try:
        write_cspack_blocked_ell(path, 7, parts, A.VALUE_F16, cws)
        for p in parts:
            p.close()
        # the matrix knows its partitions only by metadata: rows, nnz, aux
        m = Sharded(A.FORMAT_BLOCKED_ELL)
        m.define_partitions(900, [b.rows for b in bells], [b.nnz for b in bells],
                            [(b.width << 16) | b.block_size for b in bells])
        assert m.reshard([0, sum(b.rows for b in bells)])
        st = ShardedDevice(3)
        with CspackFile(path, 7) as f:
            st.stage_partition(m, 1, pack=f.f)
            assert st.record(1).view is not None and st.record(0).view is None
            st.stage_shard(m, 0, pack=f.f)
        for p, b in enumerate(bells):
            r = st.record(p)
            v = _view(r.view, A.blocked_ell_view)
            assert (v.rows, v.cols, v.block_size, v.ell_cols, v.nnz) == (b.rows, b.cols, b.block_size, b.ell_cols, b.nnz)
            assert torch.equal(torch.from_numpy(device_to_numpy(v.blockColIdx, b.row_blocks * b.width, np.uint32).view(np.int32)),
                               b.block_col_idx.cpu())
            assert torch.equal(torch.from_numpy(device_to_numpy(v.val, b.rows * b.ell_cols, np.uint16).view(np.int16)), b.val.cpu())
            assert r.allocation_bytes >= m.device_partition_bytes(p)
        st.release_shard(m, 0)
        assert st.s.cache_count == 3
        # a pack of the wrong shard size is refused before any device work
        m2 = Sharded(A.FORMAT_BLOCKED_ELL)
        m2.define_partitions(900, [bells[0].rows], [bells[0].nnz], [(bells[0].width << 16) | 8])
        with CspackFile(path, 7) as f, pytest.raises(Exception):
            st.stage_shard(m2, 0, pack=f.f)
        m2.close()
        st.close()
        m.close()
    finally:
        if path.exists():
            path.unlink()
        os.rmdir(tmp)
        cws.close()
        ws.close()
