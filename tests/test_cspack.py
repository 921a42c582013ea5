"""CSPACK01 shard packs on the CPU: the numpy restatement of the container
and blob codecs is pinned against the reference's own raw payload codec
(src/io/pack/packfile.cu, compiled under oracle/_ref) byte for byte and
through the reference's loader, and against a container assembled by hand
from the format document (docs/SPEC_CSPACK_V1.md:87-155); the library's
header checks (csamd_cspack_open, host-only) run on files the restatement
wrote."""
import ctypes as C
import struct

import numpy as np
import pytest

from cellshard_amd import _lib
from cellshard_amd import abi as A
from helpers import random_bell_partition, random_sell_partition
from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_pack_available(), reason="oracle/_ref/libref_pack.so not built (reference absent)")


@needs_ref
@pytest.mark.parametrize("vb", [1, 2, 4])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_raw_segments_match_reference_codec(tmp_path, vb, seed):
    rng = np.random.default_rng(seed * 10 + vb)
    bell = random_bell_partition(rng, 77, 90, 4, vb)
    for k, s in enumerate(bell["segments"]):
        mine = O.raw_blocked_ell_bytes(s["rows"], s["cols"], s["nnz"], s["block_size"], s["ell_cols"], s["block_col_idx"], s["val"])
        theirs = O.ref_raw_blocked_ell_bytes(tmp_path / f"b{k}", s["rows"], s["cols"], s["nnz"], s["block_size"], s["ell_cols"],
                                             s["block_col_idx"], s["val"])
        assert mine == theirs
    sell = random_sell_partition(rng, 65, 90, 4, vb)
    for k, s in enumerate(sell["segments"]):
        mine = O.raw_sliced_ell_bytes(s["rows"], s["cols"], s["nnz"], s["slice_row_offsets"], s["slice_widths"], s["col_idx"],
                                      s["val"])
        theirs = O.ref_raw_sliced_ell_bytes(tmp_path / f"s{k}", s["rows"], s["cols"], s["nnz"], s["slice_row_offsets"],
                                            s["slice_widths"], s["col_idx"], s["val"])
        assert mine == theirs


@needs_ref
@pytest.mark.parametrize("vb", [1, 2, 4])
def test_reference_loader_reads_restated_segments(tmp_path, vb):
    rng = np.random.default_rng(40 + vb)
    bell = random_bell_partition(rng, 61, 70, 3, vb)
    for k, s in enumerate(bell["segments"]):
        f = tmp_path / f"b{k}"
        f.write_bytes(O.raw_blocked_ell_bytes(s["rows"], s["cols"], s["nnz"], s["block_size"], s["ell_cols"], s["block_col_idx"],
                                              s["val"]))
        hdr, bci, val = O.ref_load_blocked_ell_raw(f, vb)
        assert list(hdr) == [O.FORMAT_BLOCKED_ELL, s["rows"], s["cols"], s["nnz"], s["block_size"], s["ell_cols"]]
        assert np.array_equal(bci, s["block_col_idx"]) and np.array_equal(val, s["val"])
    sell = random_sell_partition(rng, 61, 70, 3, vb)
    for k, s in enumerate(sell["segments"]):
        f = tmp_path / f"s{k}"
        f.write_bytes(O.raw_sliced_ell_bytes(s["rows"], s["cols"], s["nnz"], s["slice_row_offsets"], s["slice_widths"], s["col_idx"],
                                             s["val"]))
        hdr, off, wid, col, val = O.ref_load_sliced_ell_raw(f, vb)
        assert list(hdr[:5]) == [O.FORMAT_SLICED_ELL, s["rows"], s["cols"], s["nnz"], s["slice_count"]]
        assert np.array_equal(off, s["slice_row_offsets"]) and np.array_equal(wid, s["slice_widths"])
        assert np.array_equal(col, s["col_idx"]) and np.array_equal(val, s["val"])


def test_container_known_answer():
    # One Blocked-ELL partition, assembled field by field from the format
    # document: 3 rows, 4 cols, two segments (rows 0..2 at block size 2 with
    # one column block, row 2 at block size 1 with no blocks).
    part = dict(rows=3, cols=4, nnz=5, segment_row_offsets=np.array([0, 2, 3], np.uint32),
                exec_to_canonical=np.array([2, 0, 1], np.uint32), canonical_to_exec=np.array([1, 2, 0], np.uint32),
                col_maps=(np.array([0, 1, 2, 3], np.uint32), np.array([0, 1, 2, 3], np.uint32)),
                segments=[dict(rows=2, cols=4, nnz=4, block_size=2, ell_cols=2, block_col_idx=np.array([1], np.uint32),
                               val=np.array([0x3C00, 0x4000, 0x4200, 0x4400], np.uint16)),
                          dict(rows=1, cols=4, nnz=1, block_size=1, ell_cols=0, block_col_idx=np.zeros(0, np.uint32),
                               val=np.zeros(0, np.uint16))])
    blob = (struct.pack("<4I", 3, 4, 5, 2) + struct.pack("<3I", 0, 2, 3) + struct.pack("<3I", 2, 0, 1) + struct.pack("<3I", 1, 2, 0)
            + struct.pack("<4I", 0, 1, 2, 3) * 2
            + b"\x06" + struct.pack("<3I", 2, 4, 4) + struct.pack("<2I", 2, 2) + struct.pack("<I", 1) + struct.pack("<4H", 0x3C00, 0x4000,
                                                                                                                 0x4200, 0x4400)
            + b"\x06" + struct.pack("<3I", 1, 4, 1) + struct.pack("<2I", 1, 0))
    expect = b"CSPACK01" + struct.pack("<QQQ", 7, 1, 32) + blob
    got = O.cspack_bytes(7, [O.bell_execution_blob(part)])
    assert got == expect
    assert len(blob) == 16 + 12 + 24 + 32 + (13 + 8 + 4 + 8) + (13 + 8)
    sid, offs = O.cspack_header(got)
    assert sid == 7 and offs == [32]
    back = O.decode_bell_blob(got, 32, 2)
    assert back["end"] == len(got)
    assert np.array_equal(back["segments"][0]["val"], part["segments"][0]["val"])


@pytest.mark.parametrize("vb", [1, 2, 4])
def test_restated_codec_round_trips(vb):
    rng = np.random.default_rng(vb)
    bells = [random_bell_partition(rng, r, 50, n, vb) for r, n in [(40, 3), (0, 0), (0, 2), (17, 1), (90, 7)]]
    data = O.cspack_bytes(3, [O.bell_execution_blob(p) for p in bells])
    sid, offs = O.cspack_header(data)
    assert sid == 3 and len(offs) == len(bells)
    for p, o in zip(bells, offs):
        q = O.decode_bell_blob(data, o, vb)
        assert q["rows"] == p["rows"] and len(q["segments"]) == len(p["segments"])
        assert np.array_equal(q["segment_row_offsets"], p["segment_row_offsets"])
        assert np.array_equal(q["exec_to_canonical"], p["exec_to_canonical"])
        assert np.array_equal(q["col_maps"][1], p["col_maps"][1])
        for a, b in zip(q["segments"], p["segments"]):
            assert a["block_size"] == b["block_size"] and a["ell_cols"] == b["ell_cols"]
            assert np.array_equal(a["block_col_idx"], b["block_col_idx"]) and np.array_equal(a["val"], b["val"])
    sells = [random_sell_partition(rng, r, 50, n, vb) for r, n in [(40, 3), (0, 1), (23, 5)]]
    data = O.cspack_bytes(4, [O.sell_execution_blob(p) for p in sells])
    _, offs = O.cspack_header(data)
    for p, o in zip(sells, offs):
        q = O.decode_sell_blob(data, o, vb)
        assert np.array_equal(q["canonical_slice_row_offsets"], p["canonical_slice_row_offsets"])
        for a, b in zip(q["segments"], p["segments"]):
            assert np.array_equal(a["slice_row_offsets"], b["slice_row_offsets"])
            assert np.array_equal(a["col_idx"], b["col_idx"]) and np.array_equal(a["val"], b["val"])


def _open(path, shard_id=A.CSPACK_ANY_SHARD):
    f = A.cspack_file()
    rc = _lib.load().csamd_cspack_open(str(path).encode(), shard_id, C.byref(f))
    return rc, f


def test_open_reads_offsets_and_extents(tmp_path):
    rng = np.random.default_rng(9)
    blobs = [O.bell_execution_blob(random_bell_partition(rng, r, 30, n, 2)) for r, n in [(20, 2), (0, 0), (33, 4)]]
    path = tmp_path / "shard.3.cspack"
    path.write_bytes(O.cspack_bytes(3, blobs))
    rc, f = _open(path, 3)
    assert rc == 1, _lib.last_error()
    try:
        assert f.shard_id == 3 and f.partition_count == 3 and f.file_bytes == path.stat().st_size
        assert [f.partition_offsets[i] for i in range(3)] == O.cspack_header(path.read_bytes())[1]
        assert [f.partition_bytes[i] for i in range(3)] == [len(b) for b in blobs]
    finally:
        _lib.load().csamd_cspack_close(C.byref(f))
    # an empty pack
    path.write_bytes(O.cspack_bytes(5, []))
    rc, f = _open(path)
    assert rc == 1 and f.partition_count == 0 and f.shard_id == 5
    _lib.load().csamd_cspack_close(C.byref(f))


@pytest.mark.parametrize("case", ["magic", "shard", "short", "table", "offset_low", "offset_high", "missing"])
def test_open_rejects_bad_headers(tmp_path, case):
    rng = np.random.default_rng(1)
    data = bytearray(O.cspack_bytes(2, [O.bell_execution_blob(random_bell_partition(rng, 10, 8, 2, 2))]))
    shard = 2
    if case == "magic":
        data[7:8] = b"2"
    elif case == "shard":
        shard = 3
    elif case == "short":
        data = data[:20]
    elif case == "table":
        data[16:24] = struct.pack("<Q", 1 << 40)
    elif case == "offset_low":
        data[24:32] = struct.pack("<Q", 8)
    elif case == "offset_high":
        data[24:32] = struct.pack("<Q", len(data) + 1)
    path = tmp_path / "x.cspack"
    if case != "missing":
        path.write_bytes(bytes(data))
    rc, f = _open(path, shard)
    assert rc == 0
    msg = _lib.last_error()
    assert "cspack open" in msg
    assert f.impl is None and f.partition_count == 0


@pytest.mark.parametrize("bs,req", [(8, 3), (16, 2), (4, 5)])
def test_canonical_coo_restatement_undoes_the_bucketing(bs, req):
    """The canonical COO restatement inverts the bucketed execution layout
    (restated from the reference's builder, pinned against the GPU builder
    by the execution-layout tests): canonical Blocked-ELL -> execution
    partition -> canonical COO -> Blocked-ELL gives back the original when
    no stored value is zero."""
    from helpers import F16, encode, random_csr
    rows, cols = 203, 300
    ptr, idx, _ = random_csr(rows, cols, 0.05, seed=bs + req)
    bits = encode(np.random.default_rng(bs).integers(1, 9, size=len(idx)).astype(np.float64), F16)
    bci, val, ell_cols, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    lay = O.bell_execution_layout(rows, bs, ell_cols, bci, req)
    segs, e2c = O.bell_fill_reference(rows, bs, ell_cols, bci, val, lay)
    c2e = np.empty(rows, np.uint32)
    c2e[e2c] = np.arange(rows, dtype=np.uint32)
    sro = np.concatenate([[0], np.cumsum([s[0] for s in segs])]).astype(np.uint32)
    part = dict(rows=rows, cols=cols, nnz=nnz, segment_row_offsets=sro, exec_to_canonical=e2c, canonical_to_exec=c2e,
                col_maps=(np.arange(cols, dtype=np.uint32), np.arange(cols, dtype=np.uint32)),
                segments=[dict(rows=r, cols=cols, nnz=0, block_size=bs, ell_cols=w * bs, block_col_idx=b, val=v)
                          for (r, b, v), w in zip(segs, lay["width"])])
    cptr, cidx, cval = O.canonical_csr_of_bucketed_bell(part, 2)
    assert np.array_equal(cptr, ptr) and len(cidx) == len(idx)
    b2, v2, e2, n2 = O.blocked_ell_from_compressed(rows, cols, cptr, cidx, cval, bs)
    assert (e2, n2) == (ell_cols, nnz)
    assert np.array_equal(b2, bci) and np.array_equal(v2, val)


def test_canonical_sliced_restatement_undoes_a_row_permutation():
    """A canonical Sliced-ELL moved into one permuted exec segment (one slice
    as wide as its widest row) is rebuilt slot for slot by the restatement
    of rebuild_canonical_sliced_partition."""
    from helpers import F16, encode, random_csr
    rows, cols = 90, 120
    ptr, idx, _ = random_csr(rows, cols, 0.08, seed=3, empty_rows=0.1)
    bits = encode(np.random.default_rng(3).integers(1, 9, size=len(idx)).astype(np.float64), F16)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, 16)
    off, wid, slots = s["offsets"].astype(np.int64), s["widths"].astype(np.int64), s["slot_offsets"].astype(np.int64)
    W = int(wid.max())
    perm = np.random.default_rng(4).permutation(rows).astype(np.uint32)
    col = np.full(rows * W, 0xFFFFFFFF, np.uint32)
    val = np.zeros(rows * W, np.uint16)
    for x, c in enumerate(perm):
        q = int(np.searchsorted(off[:-1], c, side="right") - 1)
        b = int(slots[q] + (c - off[q]) * wid[q])
        cc, vv = s["col_idx"][b:b + int(wid[q])], s["val"][b:b + int(wid[q])]
        k = cc != 0xFFFFFFFF
        col[x * W:x * W + int(k.sum())] = cc[k]
        val[x * W:x * W + int(k.sum())] = vv[k]
    inv = np.empty(rows, np.uint32)
    inv[perm] = np.arange(rows, dtype=np.uint32)
    part = dict(rows=rows, cols=cols, nnz=len(idx), segment_row_offsets=np.array([0, rows], np.uint32), exec_to_canonical=perm,
                canonical_to_exec=inv, canonical_slice_row_offsets=s["offsets"], canonical_slice_widths=s["widths"],
                segments=[dict(rows=rows, cols=cols, nnz=len(idx), slice_count=1, slice_row_offsets=np.array([0, rows], np.uint32),
                               slice_widths=np.array([W], np.uint32), col_idx=col, val=val)])
    got_slots, got_col, got_val = O.canonical_sliced_of_bucketed_sell(part, 2)
    assert np.array_equal(got_slots, s["slot_offsets"])
    assert np.array_equal(got_col, s["col_idx"]) and np.array_equal(got_val, s["val"])
