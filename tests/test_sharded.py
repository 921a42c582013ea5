"""sharded<MatrixT> metadata operations (host only, no device): known answers
worked out by hand from the reference's rules (runtime/host/sharded_host.cuh,
runtime/layout/sharded.cuh) and properties every cut must satisfy."""
import ctypes as C

import numpy as np
import pytest

from cellshard_amd import abi as A
from cellshard_amd._lib import load
from cellshard_amd.sharded import Sharded, new_host_compressed
from helpers import f16_bits, random_csr

ROWS = [3, 0, 5, 2]
NNZ = [10, 0, 7, 4]


@pytest.fixture()
def m():
    s = Sharded(A.FORMAT_COMPRESSED)
    s.define_partitions(100, ROWS, NNZ, [0, 0, 0, 0])
    yield s
    s.close()


def test_define_partitions_offsets_and_lookups(m):
    assert m.partition_offsets() == [0, 3, 3, 8, 10]
    assert (m.m.rows, m.m.cols, m.m.nnz) == (10, 100, 21)
    assert m.num_shards == 4 and m.shard_offsets() == [0, 3, 3, 8, 10] and m.shard_parts() == [0, 1, 2, 3, 4]
    # row -> partition: the empty partition 1 owns no row
    assert [m.find_partition(r) for r in (0, 2, 3, 7, 8, 9, 10)] == [0, 0, 2, 2, 3, 3, 4]
    assert [m.row_is_partition_boundary(r) for r in (0, 3, 4, 8, 10)] == [True, True, False, True, True]


def test_set_equal_shards_known_answer(m):
    # target ceil(10 / 2) = 5 rows: partitions 0-2 hold 8 >= 5 rows, and one
    # partition remains for the last shard
    m.set_equal_shards(2)
    assert m.shard_offsets() == [0, 8, 10] and m.shard_parts() == [0, 3, 4]
    assert [m.find_shard(r) for r in (0, 7, 8, 9)] == [0, 0, 1, 1]
    m.set_equal_shards(9)  # more shards than partitions: one shard per partition
    assert m.num_shards == 4
    m.set_equal_shards(0)
    assert m.num_shards == 0


def test_shard_offsets_by_rows_and_reshard(m):
    # greedy, at most 5 rows per shard unless one partition is larger
    assert m.build_shard_offsets_by_rows(5) == [0, 3, 8, 10]
    assert m.build_shard_offsets_by_rows(0) == [0, 3, 3, 8, 10]
    assert m.reshard([0, 3, 10])
    assert m.shard_parts() == [0, 1, 4]
    assert m.first_partition_in_shard(1) == 1 and m.last_partition_in_shard(1) == 4
    assert not m.reshard([0, 4, 10])    # row 4 is inside partition 2
    assert not m.reshard([0, 8, 9])     # must end at the row count
    assert not m.reshard([0, 8, 3, 10])  # cuts must not decrease


def test_set_shards_by_nnz_known_answer(m):
    # 10 | (0 skipped) 7 + 4 = 11 <= 12
    m.set_shards_by_nnz(12)
    assert m.shard_offsets() == [0, 3, 10] and m.shard_parts() == [0, 1, 4]
    m.set_shards_by_nnz(0)
    assert m.num_shards == 4


@pytest.mark.parametrize("seed", range(5))
def test_greedy_cuts_properties(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    rows = rng.integers(0, 50, n)
    nnz = rows * rng.integers(0, 20, n)
    s = Sharded(A.FORMAT_COMPRESSED)
    s.define_partitions(64, rows, nnz, np.zeros(n, int))
    off = s.partition_offsets()
    for cap in (1, int(nnz.max(initial=0)) + 1, int(nnz.sum()) // 3 + 1):
        s.set_shards_by_nnz(cap)
        so, sp_ = s.shard_offsets(), s.shard_parts()
        assert so[0] == 0 and so[-1] == off[-1] and all(x in off for x in so)
        weights = [int(nnz[sp_[k]:sp_[k + 1]].sum()) for k in range(s.num_shards)]
        for k, w in enumerate(weights):
            nonzero = [int(v) for v in nnz[sp_[k]:sp_[k + 1]] if v]
            assert w <= cap or len(nonzero) == 1  # over the cap only as one big partition
            if k + 1 < s.num_shards:  # greedy: the next shard's first nonzero partition did not fit
                nxt = next(int(v) for v in nnz[sp_[k + 1]:] if v)
                assert w + nxt > cap
    s.close()


def test_parts_metadata_bytes_and_drop():
    lib = load()
    s = Sharded(A.FORMAT_COMPRESSED)
    host = []
    for k, (rows, cols) in enumerate([(6, 9), (0, 9), (4, 9)]):
        ptr, idx, val = random_csr(rows, cols, 0.3, seed=k)
        host.append((rows, ptr, idx, f16_bits(val)))
        s.append_partition(new_host_compressed(rows, cols, ptr, idx, f16_bits(val)))
    assert s.m.cols == 9 and s.partition_offsets() == [0, 6, 6, 10]
    assert s.column("partition_nnz") == [len(h[2]) for h in host]
    part = s.part(2)
    assert part.rows == 4 and list(np.ctypeslib.as_array(C.cast(part.majorPtr, C.POINTER(C.c_uint32)), (5,))) == list(host[2][1])
    # host bytes: struct + (rows + 1) x 4 + nnz x (4 + 2), materialized or not
    want = [C.sizeof(A.host_compressed) + (h[0] + 1) * 4 + len(h[2]) * 6 for h in host]
    assert [s.partition_bytes(i) for i in range(3)] == want
    s.drop_partition(0)
    assert s.m.parts[0] is None and s.partition_bytes(0) == want[0]
    # partitions 0 and 1 fit under the cap, partition 2 starts a shard at row
    # 6; shard membership follows row offsets, so the empty partition 1 (it
    # starts at row 6 too) falls to the second shard
    s.set_shards_by_part_bytes(want[0] + want[1])
    assert s.shard_offsets() == [0, 6, 10] and s.shard_parts() == [0, 1, 3]
    # device bytes: the view at 0, then each array 256-byte aligned
    def dev(rows, nnz):
        o = 0
        for b in (C.sizeof(A.compressed_view), (rows + 1) * 4, nnz * 4):
            o = (o + b + 255) // 256 * 256
        return o + nnz * 2
    assert [s.device_partition_bytes(i) for i in range(3)] == [dev(h[0], len(h[2])) for h in host]
    s.set_shards_by_device_bytes(dev(host[0][0], len(host[0][2])) + dev(0, 0))
    assert s.shard_offsets() == [0, 6, 10] and s.shard_parts() == [0, 1, 3]
    total = C.sizeof(A.sharded) + s.m.partition_capacity * (8 + 8 + 24) + 8 + (s.m.shard_capacity + 1) * 16 + sum(want)
    assert s.bytes() == total
    # concatenate moves the parts over
    t = Sharded(A.FORMAT_COMPRESSED)
    t.append_partition(new_host_compressed(2, 9, np.array([0, 1, 2]), np.array([3, 4]), f16_bits([1.0, 2.0])))
    s.concatenate(t)
    assert s.num_partitions == 4 and s.partition_offsets() == [0, 6, 6, 10, 12] and t.num_partitions == 0
    assert t.m.rows == 0 and s.num_shards == 4
    s.drop_shard(3)
    assert s.m.parts[3] is None
    t.close()
    s.close()
    assert lib.csamd_sharded_partition_bytes(None, A.FORMAT_COMPRESSED, 0) == 0


def test_aux_words_and_bytes_from_metadata():
    # Blocked-ELL aux = (width << 16) | block; Sliced-ELL aux = (slots << 32) | slices
    s = Sharded(A.FORMAT_BLOCKED_ELL)
    s.define_partitions(40, [20, 7], [50, 9], [(3 << 16) | 16, (2 << 16) | 4])
    assert s.partition_bytes(0) == C.sizeof(A.host_blocked_ell) + 2 * 3 * 4 + 20 * 48 * 2
    assert s.partition_bytes(1) == C.sizeof(A.host_blocked_ell) + 2 * 2 * 4 + 7 * 8 * 2
    s.close()
    t = Sharded(A.FORMAT_SLICED_ELL)
    t.define_partitions(40, [64], [100], [(130 << 32) | 2])
    assert t.partition_bytes(0) == C.sizeof(A.host_sliced_ell) + 3 * 4 + 2 * 4 + 130 * 6
    t.close()
    lib = load()
    b = A.host_blocked_ell(rows=5, cols=10, nnz=3, block_size=4, ell_cols=12)
    assert lib.csamd_partition_aux(A.FORMAT_BLOCKED_ELL, C.addressof(b)) == (3 << 16) | 4
    c = A.host_compressed(rows=5, cols=10, nnz=3, axis=A.COMPRESSED_BY_COL)
    assert lib.csamd_partition_aux(A.FORMAT_COMPRESSED, C.addressof(c)) == 1


def test_empty_matrix_calls_are_safe():
    s = Sharded(A.FORMAT_COMPRESSED)
    s.define_partitions(5, [], [], [])
    assert s.num_partitions == 0 and s.num_shards == 0 and s.m.rows == 0
    s.set_equal_shards(3)
    s.set_shards_by_nnz(10)
    assert s.build_shard_offsets_by_rows(4) == []
    assert s.find_partition(0) == 0
    s.close()
