"""sharded<MatrixT> metadata and sharded_device<MatrixT> residency over the
C ABI (csamd_sharded_* / csamd_sharded_device_*), with the reference's names
(runtime/layout/sharded.cuh, runtime/host/sharded_host.cuh,
runtime/device/sharded_device.cuh).

Parts appended here are host structs the matrix owns (allocated by the
library, destroyed by clear / drop).  The metadata calls need no device;
the residency calls do.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi as A
from ._lib import check, load

FORMATS = {"compressed": A.FORMAT_COMPRESSED, "blocked_ell": A.FORMAT_BLOCKED_ELL, "sliced_ell": A.FORMAT_SLICED_ELL}
_HOST_TYPES = {A.FORMAT_COMPRESSED: A.host_compressed, A.FORMAT_BLOCKED_ELL: A.host_blocked_ell, A.FORMAT_SLICED_ELL: A.host_sliced_ell}


def _ul_array(values):
    vals = [int(v) for v in values]
    return (C.c_ulong * max(len(vals), 1))(*vals) if vals else None


def new_host_compressed(rows, cols, indptr, indices, val_bits, axis=A.COMPRESSED_BY_ROW) -> int:
    """A library-allocated host compressed part (the address), filled from
    numpy arrays (2-byte value bits)."""
    lib = load()
    p = lib.csamd_host_part_alloc(A.FORMAT_COMPRESSED)
    if not p:
        raise MemoryError("host_part_alloc")
    m = A.host_compressed.from_address(p)
    m.rows, m.cols, m.nnz, m.axis = int(rows), int(cols), int(len(indices)), axis
    check(lib.csamd_host_compressed_allocate(C.byref(m)), "host_compressed_allocate")
    ptr = np.ascontiguousarray(indptr, np.uint32)
    idx = np.ascontiguousarray(indices, np.uint32)
    val = np.ascontiguousarray(np.asarray(val_bits).view(np.uint16))
    C.memmove(m.majorPtr, ptr.ctypes.data, ptr.nbytes)
    if len(idx):
        C.memmove(m.minorIdx, idx.ctypes.data, idx.nbytes)
        C.memmove(m.val, val.ctypes.data, val.nbytes)
    return p


def host_part(format_: int, address: int):
    """The host struct at `address` (a part owned by a sharded matrix)."""
    return _HOST_TYPES[format_].from_address(address)


class Sharded:
    """Owns a csamd_sharded of one part format."""

    def __init__(self, format_: int):
        self.format = format_
        self.m = A.sharded()
        load().csamd_sharded_init(C.byref(self.m))

    def close(self):
        if self.m is not None:
            load().csamd_sharded_clear(C.byref(self.m), self.format)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- metadata operations (sharded_host.cuh)
    def define_partitions(self, cols, rows, nnz=None, aux=None):
        n = len(rows)
        check(load().csamd_sharded_define_partitions(C.byref(self.m), self.format, int(cols), n, _ul_array(rows),
                                                     _ul_array(nnz) if nnz is not None else None,
                                                     _ul_array(aux) if aux is not None else None), "define_partitions")

    def append_partition(self, part_address: int | None):
        check(load().csamd_sharded_append_partition(C.byref(self.m), self.format, part_address), "append_partition")

    def concatenate(self, other: "Sharded"):
        check(load().csamd_sharded_concatenate(C.byref(self.m), C.byref(other.m)), "concatenate")

    def set_shards_to_partitions(self):
        check(load().csamd_sharded_set_shards_to_partitions(C.byref(self.m)), "set_shards_to_partitions")

    def set_equal_shards(self, count):
        check(load().csamd_sharded_set_equal_shards(C.byref(self.m), int(count)), "set_equal_shards")

    def build_shard_offsets_by_rows(self, target_rows):
        lib = load()
        out = C.POINTER(C.c_ulong)()
        n = C.c_ulong(0)
        check(lib.csamd_sharded_build_shard_offsets_by_rows(C.byref(self.m), int(target_rows), C.byref(out), C.byref(n)),
              "build_shard_offsets_by_rows")
        if not out:
            return []
        vals = [out[i] for i in range(n.value + 1)]
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        libc.free(C.cast(out, C.c_void_p))
        return vals

    def reshard(self, offsets) -> bool:
        arr = _ul_array(offsets)
        return load().csamd_sharded_reshard(C.byref(self.m), len(offsets) - 1, arr) == 1

    def set_shards_by_nnz(self, max_nnz):
        check(load().csamd_sharded_set_shards_by_nnz(C.byref(self.m), int(max_nnz)), "set_shards_by_nnz")

    def set_shards_by_part_bytes(self, max_bytes):
        check(load().csamd_sharded_set_shards_by_part_bytes(C.byref(self.m), self.format, int(max_bytes)), "set_shards_by_part_bytes")

    def set_shards_by_device_bytes(self, max_bytes):
        check(load().csamd_sharded_set_shards_by_device_bytes(C.byref(self.m), self.format, int(max_bytes)),
              "set_shards_by_device_bytes")

    def drop_partition(self, p):
        check(load().csamd_sharded_drop_partition(C.byref(self.m), self.format, int(p)), "drop_partition")

    def drop_shard(self, s):
        check(load().csamd_sharded_drop_shard(C.byref(self.m), self.format, int(s)), "drop_shard")

    # --- lookups and sizes (sharded.cuh)
    def find_partition(self, row):
        return int(load().csamd_sharded_find_partition(C.byref(self.m), int(row)))

    def find_shard(self, row):
        return int(load().csamd_sharded_find_shard(C.byref(self.m), int(row)))

    def first_partition_in_shard(self, s):
        return int(load().csamd_sharded_first_partition_in_shard(C.byref(self.m), int(s)))

    def last_partition_in_shard(self, s):
        return int(load().csamd_sharded_last_partition_in_shard(C.byref(self.m), int(s)))

    def row_is_partition_boundary(self, row):
        return load().csamd_sharded_row_is_partition_boundary(C.byref(self.m), int(row)) == 1

    def partition_bytes(self, p):
        return int(load().csamd_sharded_partition_bytes(C.byref(self.m), self.format, int(p)))

    def shard_bytes(self, s):
        return int(load().csamd_sharded_shard_bytes(C.byref(self.m), self.format, int(s)))

    def bytes(self):
        return int(load().csamd_sharded_bytes(C.byref(self.m), self.format))

    def device_partition_bytes(self, p):
        return int(load().csamd_sharded_device_partition_bytes(C.byref(self.m), self.format, int(p)))

    def device_shard_bytes(self, s):
        return int(load().csamd_sharded_device_shard_bytes(C.byref(self.m), self.format, int(s)))

    # --- plain views of the metadata
    @property
    def num_partitions(self):
        return int(self.m.num_partitions)

    @property
    def num_shards(self):
        return int(self.m.num_shards)

    def partition_offsets(self):
        return [int(self.m.partition_offsets[i]) for i in range(self.num_partitions + 1)] if self.m.partition_offsets else [0]

    def shard_offsets(self):
        return [int(self.m.shard_offsets[i]) for i in range(self.num_shards + 1)] if self.m.shard_offsets else []

    def shard_parts(self):
        return [int(self.m.shard_parts[i]) for i in range(self.num_shards + 1)] if self.m.shard_parts else []

    def column(self, name):
        arr = getattr(self.m, name)
        return [int(arr[i]) for i in range(self.num_partitions)] if arr else []

    def part(self, p):
        addr = self.m.parts[p]
        return host_part(self.format, addr) if addr else None


def repack_to_blocked_ell(src: "Sharded", block_size: int | None = None, target_block_rows: int = 0, candidates=None):
    """repack_sharded_compressed_to_blocked_ell(_auto) (convert/blocked_ell_from_compressed.cuh:480-600):
    a new Sharded(FORMAT_BLOCKED_ELL) of parts of target_block_rows x block_size
    rows; with block_size None the size is chosen from `candidates` over the
    whole matrix.  Returns (dst, picked tune result or None)."""
    dst = Sharded(A.FORMAT_BLOCKED_ELL)
    lib = load()
    if block_size is not None:
        check(lib.csamd_repack_sharded_compressed_to_blocked_ell(C.byref(src.m), int(block_size), int(target_block_rows),
                                                                 C.byref(dst.m)), "repack_sharded_compressed_to_blocked_ell")
        return dst, None
    cand = np.ascontiguousarray(candidates, dtype=np.uint32)
    tune = A.blocked_ell_tune_result()
    check(lib.csamd_repack_sharded_compressed_to_blocked_ell_auto(C.byref(src.m), cand.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                                  len(cand), int(target_block_rows), C.byref(dst.m), C.byref(tune)),
          "repack_sharded_compressed_to_blocked_ell_auto")
    return dst, tune


class ShardedDevice:
    """Owns a csamd_sharded_device (residency table + allocation cache)."""

    def __init__(self, capacity: int = 0):
        self.s = A.sharded_device()
        load().csamd_sharded_device_init(C.byref(self.s))
        if capacity:
            self.reserve(capacity)

    def reserve(self, capacity):
        check(load().csamd_sharded_device_reserve(C.byref(self.s), int(capacity)), "sharded_device reserve")

    def close(self):
        if self.s is not None:
            check(load().csamd_sharded_device_clear(C.byref(self.s)), "sharded_device clear")
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record(self, p):
        return self.s.parts[p]

    def upload_partition(self, view: Sharded, p, device=0, stream=None):
        check(load().csamd_sharded_device_upload_partition(C.byref(self.s), C.byref(view.m), view.format, int(p), device, stream),
              "upload_partition")

    def release_partition(self, p):
        check(load().csamd_sharded_device_release_partition(C.byref(self.s), int(p)), "release_partition")

    def upload_shard(self, view: Sharded, s, device=0, stream=None):
        check(load().csamd_sharded_device_upload_shard(C.byref(self.s), C.byref(view.m), view.format, int(s), device, stream),
              "upload_shard")

    def release_shard(self, view: Sharded, s):
        check(load().csamd_sharded_device_release_shard(C.byref(self.s), C.byref(view.m), int(s)), "release_shard")

    def stage_shard(self, view: Sharded, s, pack=None, device=0, stream=None, drop_host_after_upload=False):
        check(load().csamd_sharded_device_stage_shard(C.byref(self.s), C.byref(view.m), view.format,
                                                      C.byref(pack) if pack is not None else None, int(s), device, stream,
                                                      1 if drop_host_after_upload else 0), "stage_shard")

    def stage_partition(self, view: Sharded, p, pack=None, device=0, stream=None, drop_host_after_upload=False):
        check(load().csamd_sharded_device_stage_partition(C.byref(self.s), C.byref(view.m), view.format,
                                                          C.byref(pack) if pack is not None else None, int(p), device, stream,
                                                          1 if drop_host_after_upload else 0), "stage_partition")

    def swap_shard(self, view: Sharded, out_shard, in_shard, pack=None, device=0, stream=None, drop_host_after_upload=False,
                   drop_host_after_release=False):
        check(load().csamd_sharded_device_swap_shard(C.byref(self.s), C.byref(view.m), view.format,
                                                     C.byref(pack) if pack is not None else None, int(out_shard), int(in_shard),
                                                     device, stream, 1 if drop_host_after_upload else 0,
                                                     1 if drop_host_after_release else 0), "swap_shard")


class ShardMajorBucket:
    """A csamd_sharded_major_bucket_workspace (bucket/operators/sharded_major_nnz.cuh:35-273):
    the nnz bucket plan and rebuild of one resident shard of a compressed
    sharded matrix.  Results alias the workspace until its next call."""

    def __init__(self, device: int = 0, stream=None):
        self.ws = A.sharded_major_bucket_workspace()
        load().csamd_sharded_major_bucket_init(C.byref(self.ws))
        check(load().csamd_sharded_major_bucket_setup(C.byref(self.ws), device, stream), "shard bucket setup")

    def build(self, state: ShardedDevice, view: Sharded, shard_id: int, requested_bucket_count: int,
              value_type: int = A.VALUE_F16) -> A.sharded_major_bucket_result:
        out = A.sharded_major_bucket_result()
        check(load().csamd_build_bucketed_shard_major_view_typed(C.byref(state.s), C.byref(view.m), int(shard_id), value_type,
                                                                 requested_bucket_count, C.byref(self.ws), C.byref(out)),
              "build_bucketed_shard_major_view")
        return out

    def close(self):
        if self.ws is not None:
            load().csamd_sharded_major_bucket_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
