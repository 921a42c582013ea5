"""Device staging (packed partition / shard uploads, device view copies,
allocation cache) on the GPU.

Uploads are byte copies: every device array equals its host array bit for
bit; Sliced-ELL slot offsets and the uniform slice height are the closed
forms of formats/sliced_ell.cuh (running sum of rows x width; the common
height when every slice but a shorter last one has it); the cache returns
the smallest block that fits and frees past 16 entries.  A reduce over the
staged copy equals the oracle on the host arrays.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from cellshard_amd import abi as A
from cellshard_amd._lib import check, load
from cellshard_amd.device import DeviceCSR, to_dev_u32
from cellshard_amd.ops import LabelReduce, _hip, device_to_numpy
from helpers import F16, encode, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data if a.size else None


def _host_csr(rows, cols, ptr, idx, bits, keep):
    h = A.host_compressed()
    h.rows, h.cols, h.nnz, h.axis = rows, cols, len(idx), A.COMPRESSED_BY_ROW
    arrs = [np.ascontiguousarray(ptr, np.uint32), np.ascontiguousarray(idx, np.uint32), np.ascontiguousarray(bits)]
    keep += arrs
    h.majorPtr, h.minorIdx, h.val = _p(arrs[0]), _p(arrs[1]), _p(arrs[2])
    return h


def _view_struct(ptr_, cls):
    """The device copy of a view, read back."""
    raw = device_to_numpy(ptr_, C.sizeof(cls), np.uint8)
    return cls.from_buffer_copy(raw.tobytes())


@pytest.mark.parametrize("async_", [False, True])
def test_upload_compressed_roundtrip_and_reduce(gpu, async_):
    lib = load()
    rows, cols, G = 700, 333, 6
    ptr, idx, val = random_csr(rows, cols, 0.05, seed=11, empty_rows=0.1)
    bits = encode(val, F16)
    keep = []
    h = _host_csr(rows, cols, ptr, idx, bits, keep)
    rec, view = A.partition_record(), A.compressed_view()
    side = torch.cuda.Stream() if async_ else None  # async copies on a stream of their own
    check(lib.csamd_upload_compressed(C.byref(h), C.byref(rec), C.byref(view), None, side.cuda_stream if side else None),
          "upload")
    torch.cuda.synchronize()
    assert np.array_equal(device_to_numpy(view.majorPtr, rows + 1, np.uint32), ptr)
    assert np.array_equal(device_to_numpy(view.minorIdx, len(idx), np.uint32), idx)
    assert np.array_equal(device_to_numpy(view.val, len(idx), np.uint16), bits.view(np.uint16))
    dv = _view_struct(rec.view, A.compressed_view)
    assert (dv.rows, dv.cols, dv.nnz, dv.majorPtr, dv.minorIdx, dv.val) == (rows, cols, len(idx), view.majorPtr, view.minorIdx, view.val)
    for a in (view.majorPtr, view.minorIdx, view.val):
        assert a % 256 == 0  # packed arrays start 256-byte aligned
    # reduce straight from the staged view
    labels = np.random.default_rng(2).integers(0, G, size=rows).astype(np.uint32)
    q = LabelReduce._query(F16, to_dev_u32(labels), G, None, None, None, True)
    lr = LabelReduce()
    res = A.label_reduce_result()
    check(lib.csamd_label_reduce_compressed(C.byref(view), C.byref(q), C.byref(lr.ws), C.byref(res)), "reduce staged")
    got = LabelReduce.fetch(res)
    ref = O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, None, None, G)
    assert np.array_equal(got["nnz"], ref["nnz"]) and np.array_equal(got["sum"], ref["sum"])  # integer counts: exact
    lr.close()
    check(lib.csamd_release_record(C.byref(rec), None), "release")
    assert rec.allocation is None


def test_upload_sliced_ell_offsets_and_height(gpu):
    lib = load()
    rows, cols = 203, 150
    ptr, idx, val = random_csr(rows, cols, 0.07, seed=5)
    bits = encode(val, F16)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, 32)
    keep = [np.ascontiguousarray(s["offsets"], np.uint32), np.ascontiguousarray(s["widths"], np.uint32),
            np.ascontiguousarray(s["col_idx"], np.uint32), np.ascontiguousarray(s["val"])]
    h = A.host_sliced_ell()
    h.rows, h.cols, h.nnz, h.slice_count = rows, cols, len(idx), len(keep[1])
    h.slice_row_offsets, h.slice_widths, h.col_idx, h.val = (_p(a) for a in keep)
    rec, view = A.partition_record(), A.sliced_ell_view()
    check(lib.csamd_upload_sliced_ell(C.byref(h), C.byref(rec), C.byref(view), None, None), "upload sell")
    n = len(keep[1])
    heights = np.diff(keep[0].astype(np.int64))
    slot_off = np.concatenate([[0], np.cumsum(heights * keep[1])])[:-1]
    assert np.array_equal(device_to_numpy(view.slice_slot_offsets, n, np.uint32), slot_off)
    assert view.slice_rows == 32  # 6 slices of 32 and a last one of 11
    total = int((heights * keep[1]).sum())
    assert np.array_equal(device_to_numpy(view.col_idx, total, np.uint32), keep[2])
    assert np.array_equal(device_to_numpy(view.val, total, np.uint16), keep[3].view(np.uint16))
    check(lib.csamd_release_record(C.byref(rec), None), "release")
    # non-uniform heights: slice_rows 0
    offs = np.array([0, 10, 30, 31], np.uint32)
    w = np.array([1, 2, 0], np.uint32)
    nslots = 10 * 1 + 20 * 2
    ci = np.zeros(nslots, np.uint32)
    vv = np.zeros(nslots, np.uint16)
    keep2 = [offs, w, ci, vv]
    h2 = A.host_sliced_ell()
    h2.rows, h2.cols, h2.nnz, h2.slice_count = 31, 4, 0, 3
    h2.slice_row_offsets, h2.slice_widths, h2.col_idx, h2.val = (_p(a) for a in keep2)
    check(lib.csamd_upload_sliced_ell(C.byref(h2), C.byref(rec), C.byref(view), None, None), "upload sell 2")
    assert view.slice_rows == 0
    assert np.array_equal(device_to_numpy(view.slice_slot_offsets, 3, np.uint32), [0, 10, 50])
    check(lib.csamd_release_record(C.byref(rec), None), "release")


def test_compressed_shard_and_cache(gpu):
    lib = load()
    cols = 90
    keep, hosts, ref = [], [], []
    for i, rows in enumerate([40, 0, 75, 13]):
        ptr, idx, val = random_csr(rows, cols, 0.1, seed=30 + i)
        bits = encode(val, F16)
        hosts.append(_host_csr(rows, cols, ptr, idx, bits, keep))
        ref.append((ptr, idx, bits))
    arr = (C.POINTER(A.host_compressed) * len(hosts))(*[C.pointer(h) for h in hosts])
    views = (A.compressed_view * len(hosts))()
    cache = A.allocation_cache()
    lib.csamd_allocation_cache_init(C.byref(cache))
    rec = A.partition_record()
    check(lib.csamd_upload_compressed_shard(arr, len(hosts), C.byref(rec), views, C.byref(cache), None), "shard upload")
    base, nbytes = rec.allocation, rec.allocation_bytes
    for v, (ptr, idx, bits) in zip(views, ref):
        assert base <= v.majorPtr < base + nbytes  # one packed allocation
        assert np.array_equal(device_to_numpy(v.majorPtr, v.rows + 1, np.uint32), ptr)
        assert np.array_equal(device_to_numpy(v.minorIdx, len(idx), np.uint32), idx)
        assert np.array_equal(device_to_numpy(v.val, len(idx), np.uint16), bits.view(np.uint16))
    dev_views = device_to_numpy(rec.view, len(hosts) * C.sizeof(A.compressed_view), np.uint8)
    assert dev_views.tobytes() == bytes(views)
    # release to the cache; the next upload that fits reuses the block
    check(lib.csamd_release_record(C.byref(rec), C.byref(cache)), "release")
    assert cache.cache_count == 1
    rec2 = A.partition_record()
    v2 = A.compressed_view()
    check(lib.csamd_upload_compressed(C.byref(hosts[3]), C.byref(rec2), C.byref(v2), C.byref(cache), None), "reupload")
    assert rec2.allocation == base and rec2.allocation_bytes == nbytes and cache.cache_count == 0
    check(lib.csamd_release_record(C.byref(rec2), C.byref(cache)), "release 2")
    lib.csamd_allocation_cache_clear(C.byref(cache))
    assert cache.cache_count == 0


def test_cache_best_fit_and_limit(gpu):
    lib = load()
    cache = A.allocation_cache()
    lib.csamd_allocation_cache_init(C.byref(cache))
    hip = _hip()
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    ptrs = []
    for size in [4096, 1024, 65536, 2048] + [512] * 14:
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), size) == 0
        ptrs.append((p.value, size))
    for p, size in ptrs:  # 18 blocks: the last two are freed, not kept
        check(lib.csamd_cache_allocation(C.byref(cache), 0, C.c_void_p(p), size), "cache")
    assert cache.cache_count == 16
    got = C.c_size_t(0)
    p = lib.csamd_take_cached_allocation(C.byref(cache), 0, 1500, C.byref(got))
    assert (p, got.value) == (ptrs[3][0], 2048)  # smallest block >= 1500
    assert lib.csamd_take_cached_allocation(C.byref(cache), 1, 10, C.byref(got)) is None  # other device
    assert lib.csamd_take_cached_allocation(C.byref(cache), 0, 1 << 20, C.byref(got)) is None  # too big
    check(lib.csamd_cache_allocation(C.byref(cache), 0, C.c_void_p(p), 2048), "cache again")
    lib.csamd_allocation_cache_clear(C.byref(cache))
    assert cache.cache_count == 0
