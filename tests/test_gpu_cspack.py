"""CSPACK01 shard packs on the GPU: files the library writes from device
partitions equal the restated encoder's bytes (which the CPU tests pin to
the reference's own raw payload codec); staging from files the restatement
wrote returns every array bit for bit (Blocked-ELL and Sliced-ELL, 1-, 2-
and 4-byte values, empty partitions and segments, many segments, pieces
smaller than a segment so reads and copies split arrays at odd places);
the shard call equals the per-partition calls; damaged blobs are refused
on the host before any device work, and the workspace stays usable."""
import struct

import numpy as np
import pytest

from cellshard_amd import _lib
from cellshard_amd.device import DeviceBlockedELL, DeviceSlicedELL
from cellshard_amd.ops import (ConvertWorkspace, CspackFile, CspackWorkspace, build_bucketed_blocked_ell_device,
                               build_bucketed_sliced_ell_device, stage_cspack_blocked_ell, stage_cspack_shard_blocked_ell,
                               stage_cspack_sliced_ell, write_cspack_blocked_ell, write_cspack_sliced_ell)
from helpers import F16, encode, random_bell_partition, random_sell_partition
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VT = {1: 7, 2: 3, 4: 4}  # FP8 e4m3, F16, F32 value types by width


@pytest.fixture()
def cws(gpu):
    w = ConvertWorkspace()
    yield w
    w.close()


def _bell_dev(rows, cols, bs, seed):
    rng = np.random.default_rng(seed)
    ptr, idx = [0], []
    for r in range(rows):
        k = int(rng.integers(0, cols // (2 if (r // bs) % 3 == 0 else 12)))
        idx.append(np.sort(rng.choice(cols, k, replace=False)))
        ptr.append(ptr[-1] + k)
    ptr = np.asarray(ptr, np.uint32)
    idx = np.concatenate(idx).astype(np.uint32)
    bits = encode(rng.integers(0, 9, size=len(idx)).astype(np.float64), F16)
    bci, val, ell_cols, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    return DeviceBlockedELL.from_numpy(rows, cols, nnz, bs, ell_cols, bci, val)


def _sell_dev(rows, cols, S, seed):
    rng = np.random.default_rng(seed)
    lens = np.where(rng.random(rows) < 0.7, rng.integers(0, 6, rows), rng.integers(20, 60, rows))
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    idx = np.concatenate([np.sort(rng.choice(cols, int(n), replace=False)) for n in lens]).astype(np.uint32)
    bits = encode(rng.integers(1, 9, size=len(idx)).astype(np.float64), F16)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, S)
    return DeviceSlicedELL.from_numpy(rows, cols, len(idx), s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"],
                                      s["col_idx"], s["val"])


def _same_bell(got, want):
    for k in ("rows", "cols", "nnz"):
        assert got[k] == want[k], k
    assert np.array_equal(got["segment_row_offsets"], want["segment_row_offsets"])
    assert np.array_equal(got["exec_to_canonical"], want["exec_to_canonical"])
    assert np.array_equal(got["canonical_to_exec"], want["canonical_to_exec"])
    assert np.array_equal(got["col_maps"][0], want["col_maps"][0]) and np.array_equal(got["col_maps"][1], want["col_maps"][1])
    assert len(got["segments"]) == len(want["segments"])
    for a, b in zip(got["segments"], want["segments"]):
        for k in ("rows", "cols", "nnz", "block_size", "ell_cols"):
            assert a[k] == b[k], k
        assert np.array_equal(a["block_col_idx"], b["block_col_idx"])
        assert a["val"].dtype == b["val"].dtype and np.array_equal(a["val"], b["val"])


def _same_sell(got, want):
    for k in ("rows", "cols", "nnz"):
        assert got[k] == want[k], k
    for k in ("segment_row_offsets", "exec_to_canonical", "canonical_to_exec", "canonical_slice_row_offsets",
              "canonical_slice_widths"):
        assert np.array_equal(got[k], want[k]), k
    assert len(got["segments"]) == len(want["segments"])
    for a, b in zip(got["segments"], want["segments"]):
        for k in ("rows", "cols", "nnz", "slice_count"):
            assert a[k] == b[k], k
        for k in ("slice_row_offsets", "slice_widths", "col_idx", "val"):
            assert np.array_equal(a[k], b[k]), k
        off, wid = b["slice_row_offsets"].astype(np.int64), b["slice_widths"].astype(np.int64)
        slots = np.concatenate([[0], np.cumsum(np.diff(off) * wid)])[:-1] if b["slice_count"] else np.zeros(0)
        assert np.array_equal(a["slice_slot_offsets"], slots)
        h = np.diff(off)
        uniform = b["slice_count"] > 0 and all(x == h[0] for x in h[:-1]) and h[-1] <= h[0]
        assert a["slice_rows"] == (int(h[0]) if uniform else 0)


def _bell_bytes(p):
    return sum(O.bell_packed_bytes(s["rows"], s["ell_cols"], s["block_size"]) for s in p["segments"])


def test_written_blocked_ell_pack_matches_restatement_and_stages_back(cws, tmp_path):
    parts = [build_bucketed_blocked_ell_device(_bell_dev(rows, 300, bs, rows + bs), k, cws)
             for rows, bs, k in [(203, 8, 3), (97, 16, 8), (40, 8, 1)]]
    want = [p.to_host() for p in parts]
    path = tmp_path / "shard.11.cspack"
    ws = CspackWorkspace()
    write_cspack_blocked_ell(path, 11, parts, 3, ws)
    assert path.read_bytes() == O.cspack_bytes(11, [O.bell_execution_blob(w) for w in want])
    assert not (tmp_path / "shard.11.cspack.tmp").exists()
    with CspackFile(path, 11) as f:
        assert f.partition_count == 3
        for i, w in enumerate(want):
            got = stage_cspack_blocked_ell(f, i, 3, ws)
            g = got.to_host()
            _same_bell(g, w)
            assert g["bytes"] == w["bytes"] == _bell_bytes(w)
            got.close()
    for p in parts:
        p.close()
    ws.close()


@pytest.mark.parametrize("vb", [1, 2, 4])
@pytest.mark.parametrize("threads,chunk", [(0, 0), (4, 4096), (3, 1000)])
def test_stage_blocked_ell_from_restated_pack(gpu, tmp_path, vb, threads, chunk):
    rng = np.random.default_rng(vb * 100 + threads)
    shapes = [(300, 500, 3), (0, 20, 0), (0, 20, 2), (1, 5, 1), (257, 64, 20), (1000, 2000, 5)]
    want = [random_bell_partition(rng, r, c, n, vb) for r, c, n in shapes]
    path = tmp_path / "shard.5.cspack"
    path.write_bytes(O.cspack_bytes(5, [O.bell_execution_blob(w) for w in want]))
    ws = CspackWorkspace(reader_threads=threads, chunk_bytes=chunk)
    with CspackFile(path, 5) as f:
        for i, w in enumerate(want):
            print(f"partition {i} rows {w['rows']} segments {len(w['segments'])}", flush=True)
            got = stage_cspack_blocked_ell(f, i, VT[vb], ws)
            g = got.to_host()
            _same_bell(g, w)
            assert g["bytes"] == _bell_bytes(w)
            got.close()
        shard = stage_cspack_shard_blocked_ell(f, VT[vb], ws)
        assert len(shard) == len(want)
        for got, w in zip(shard, want):
            _same_bell(got.to_host(), w)
            got.close()
    ws.close()


def test_written_sliced_ell_pack_matches_restatement_and_stages_back(cws, tmp_path):
    parts = [build_bucketed_sliced_ell_device(_sell_dev(rows, 200, S, rows * 7 + S), k, cws)
             for rows, S, k in [(150, 32, 3), (64, 8, 8), (1, 32, 4)]]
    want = [p.to_host() for p in parts]
    path = tmp_path / "shard.2.cspack"
    ws = CspackWorkspace()
    write_cspack_sliced_ell(path, 2, parts, 3, ws)
    assert path.read_bytes() == O.cspack_bytes(2, [O.sell_execution_blob(w) for w in want])
    with CspackFile(path, 2) as f:
        for i, w in enumerate(want):
            got = stage_cspack_sliced_ell(f, i, 3, ws)
            g = got.to_host()
            _same_sell(g, w)
            assert g["bytes"] == sum(O.sell_packed_bytes(s["slice_count"], len(s["col_idx"])) for s in w["segments"])
            got.close()
    for p in parts:
        p.close()
    ws.close()


@pytest.mark.parametrize("vb", [1, 2, 4])
def test_stage_sliced_ell_from_restated_pack(gpu, tmp_path, vb):
    rng = np.random.default_rng(vb + 7)
    want = [random_sell_partition(rng, r, c, n, vb) for r, c, n in [(200, 300, 4), (0, 9, 1), (31, 40, 12), (500, 1000, 2)]]
    path = tmp_path / "s.cspack"
    path.write_bytes(O.cspack_bytes(1, [O.sell_execution_blob(w) for w in want]))
    ws = CspackWorkspace(reader_threads=2, chunk_bytes=777)
    with CspackFile(path) as f:
        for i, w in enumerate(want):
            print(f"partition {i} rows {w['rows']} segments {len(w['segments'])}", flush=True)
            got = stage_cspack_sliced_ell(f, i, VT[vb], ws)
            _same_sell(got.to_host(), w)
            got.close()
    ws.close()


def _corrupt(data, offs, case):
    """Damages partition 0 (partition 1 for a truncated file); returns the
    bytes and the damaged partition."""
    d = bytearray(data)
    o = offs[0]
    rows, cols, _, nseg = struct.unpack_from("<4I", d, o)
    seg0 = o + 16 + 4 * (nseg + 1) + 8 * rows + 8 * cols  # first segment's raw header
    if case == "format":
        d[seg0] = 7
    elif case == "segment_rows":
        struct.pack_into("<I", d, o + 20, struct.unpack_from("<I", d, o + 20)[0] + 1)
    elif case == "ell_cols":
        struct.pack_into("<I", d, seg0 + 17, struct.unpack_from("<I", d, seg0 + 17)[0] + 1)
    elif case == "truncated":
        return bytes(d[:-3]), 1
    return bytes(d), 0


@pytest.mark.parametrize("case", ["format", "segment_rows", "ell_cols", "truncated"])
def test_damaged_blobs_are_refused(gpu, tmp_path, case):
    rng = np.random.default_rng(3)
    want = [random_bell_partition(rng, 64, 40, 2, 2, block_sizes=(4,)), random_bell_partition(rng, 30, 40, 1, 2)]
    want[1]["segments"][0]["ell_cols"] = 0
    want[1]["segments"][0]["block_col_idx"] = np.zeros(0, np.uint32)
    want[1]["segments"][0]["val"] = np.zeros(0, np.uint16)
    want[1]["segments"].append(dict(rows=0, cols=40, nnz=0, block_size=1, ell_cols=1, block_col_idx=np.zeros(0, np.uint32),
                                    val=np.zeros(0, np.uint16)))
    want[1]["segment_row_offsets"] = np.array([0, 30, 30], np.uint32)
    data = O.cspack_bytes(0, [O.bell_execution_blob(w) for w in want])
    _, offs = O.cspack_header(data)
    bad, victim = _corrupt(data, offs, case)
    path = tmp_path / "bad.cspack"
    path.write_bytes(bad)
    ws = CspackWorkspace()
    with CspackFile(path) as f:
        with pytest.raises(RuntimeError):
            stage_cspack_blocked_ell(f, victim, 3, ws)
        assert f"cspack partition {victim}" in _lib.last_error()
        with pytest.raises(RuntimeError):
            stage_cspack_shard_blocked_ell(f, 3, ws)
    good = tmp_path / "good.cspack"
    good.write_bytes(data)
    with CspackFile(good) as f:
        for i, w in enumerate(want):
            got = stage_cspack_blocked_ell(f, i, 3, ws)
            _same_bell(got.to_host(), w)
            got.close()
    ws.close()


# ---------------------------------------------------------------------------
# Canonical reconstruction (the reference's fetch of a Blocked-ELL partition
# from a pack: exec partition -> canonical COO -> Blocked-ELL)
# ---------------------------------------------------------------------------
def _exec_partition_no_repeats(rng, rows, cols, nseg, vb, block_sizes=(4, 8, 16)):
    """Random exec partition whose row blocks never repeat a block column
    (so every (row, column) appears once), with zero and negative-zero
    values, invalid slots, and non-identity row and column maps."""
    p = random_bell_partition(rng, rows, cols, nseg, vb, block_sizes)
    for s in p["segments"]:
        bs, w = s["block_size"], s["ell_cols"] // s["block_size"] if s["block_size"] else 0
        nb = max(cols // bs, 1)
        rb = (s["rows"] + bs - 1) // bs
        w = min(w, nb)
        s["ell_cols"] = w * bs
        bci = np.full(rb * w, 0xFFFFFFFF, np.uint32)
        for b in range(rb):
            pick = rng.choice(nb, w, replace=False).astype(np.uint32)
            pick[rng.random(w) < 0.2] = 0xFFFFFFFF
            bci[b * w:(b + 1) * w] = pick
        s["block_col_idx"] = bci
        val = rng.integers(0, 1 << (8 * vb), s["rows"] * w * bs, dtype=np.uint64)
        val[rng.random(val.size) < 0.3] = 0
        val[rng.random(val.size) < 0.05] = 1 << (8 * vb - 1)  # -0
        s["val"] = val.astype({1: np.uint8, 2: np.uint16, 4: np.uint32}[vb])
    return p


@pytest.mark.parametrize("vb", [1, 2, 4])
def test_canonical_csr_matches_restatement(cws, tmp_path, vb):
    rng = np.random.default_rng(60 + vb)
    from cellshard_amd.ops import canonical_compressed_from_bucketed_blocked_ell
    want = [_exec_partition_no_repeats(rng, r, c, n, vb) for r, c, n in [(150, 200, 3), (33, 47, 1), (0, 10, 0), (64, 500, 5)]]
    path = tmp_path / "c.cspack"
    path.write_bytes(O.cspack_bytes(0, [O.bell_execution_blob(w) for w in want]))
    ws = CspackWorkspace()
    with CspackFile(path) as f:
        for i, w in enumerate(want):
            print(f"partition {i} rows {w['rows']}", flush=True)
            part = stage_cspack_blocked_ell(f, i, VT[vb], ws)
            csr = canonical_compressed_from_bucketed_blocked_ell(part, cws)
            p_, i_, v_ = csr.to_numpy()
            ptr, idx, val = O.canonical_csr_of_bucketed_bell(w, vb)
            assert np.array_equal(p_.view(np.uint32), ptr)
            assert np.array_equal(i_.view(np.uint32), idx)
            assert np.array_equal(v_.view(val.dtype), val)
            if vb == 2 and w["rows"]:
                from cellshard_amd.ops import reconstruct_canonical_blocked_ell
                bell = reconstruct_canonical_blocked_ell(part, 8, cws)
                bci, bval, ell, nnz = O.blocked_ell_from_compressed(w["rows"], w["cols"], ptr, idx, val, 8)
                gb, gv = bell.to_numpy()
                assert bell.ell_cols == ell and bell.nnz == nnz
                assert np.array_equal(gb.view(np.uint32), bci) and np.array_equal(gv.view(np.uint16), bval)
            part.close()
    ws.close()


def test_reconstructed_partition_equals_the_original(cws, tmp_path):
    """canonical Blocked-ELL -> bucketed exec partition -> pack -> staged ->
    reconstructed: the original Blocked-ELL bit for bit (no stored zeros)."""
    from cellshard_amd.ops import reconstruct_canonical_blocked_ell
    rng = np.random.default_rng(8)
    rows, cols, bs = 301, 400, 8
    ptr, idx = [0], []
    for r in range(rows):
        k = int(rng.integers(0, cols // (2 if (r // bs) % 3 == 0 else 12)))
        idx.append(np.sort(rng.choice(cols, k, replace=False)))
        ptr.append(ptr[-1] + k)
    ptr = np.asarray(ptr, np.uint32)
    idx = np.concatenate(idx).astype(np.uint32)
    bits = encode(rng.integers(1, 9, size=len(idx)).astype(np.float64), F16)
    bci, val, ell_cols, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    src = DeviceBlockedELL.from_numpy(rows, cols, nnz, bs, ell_cols, bci, val)
    parts = [build_bucketed_blocked_ell_device(src, 3, cws)]
    path = tmp_path / "r.cspack"
    ws = CspackWorkspace()
    write_cspack_blocked_ell(path, 4, parts, 3, ws)
    with CspackFile(path, 4) as f:
        staged = stage_cspack_blocked_ell(f, 0, 3, ws)
    back = reconstruct_canonical_blocked_ell(staged, bs, cws)
    gb, gv = back.to_numpy()
    assert back.ell_cols == ell_cols and back.nnz == nnz
    assert np.array_equal(gb.view(np.uint32), bci) and np.array_equal(gv.view(np.uint16), val)
    staged.close()
    parts[0].close()
    ws.close()


@pytest.mark.parametrize("case", ["row_map", "col_map"])
def test_reconstruction_rejects_bad_maps(cws, tmp_path, case):
    from cellshard_amd.ops import canonical_compressed_from_bucketed_blocked_ell
    rng = np.random.default_rng(2)
    w = _exec_partition_no_repeats(rng, 40, 64, 2, 2)
    if case == "row_map":
        w["exec_to_canonical"][1] = w["exec_to_canonical"][0]
    else:
        w["col_maps"][0][:] = 64  # every mapped column past the end
    path = tmp_path / "bad.cspack"
    path.write_bytes(O.cspack_bytes(0, [O.bell_execution_blob(w)]))
    ws = CspackWorkspace()
    with CspackFile(path) as f:
        part = stage_cspack_blocked_ell(f, 0, 3, ws)
    with pytest.raises(RuntimeError):
        canonical_compressed_from_bucketed_blocked_ell(part, cws)
    assert ("permutation" if case == "row_map" else "mapped column") in _lib.last_error()
    part.close()
    ws.close()


def _with_canonical_slices(rng, p):
    """Gives a random Sliced-ELL exec partition a canonical slice layout
    that covers its rows: uniform heights (short last slice), each slice as
    wide as its rows' valid slots need plus 0-2 spare."""
    rows = p["rows"]
    need = np.zeros(rows, np.int64)
    sro = p["segment_row_offsets"]
    for k, s in enumerate(p["segments"]):
        so = s["slice_row_offsets"].astype(np.int64)
        sw = s["slice_widths"].astype(np.int64)
        slots = O._slot_offsets(so, sw)
        for local in range(s["rows"]):
            q = int(np.searchsorted(so[:-1], local, side="right") - 1)
            b = int(slots[q] + (local - so[q]) * sw[q])
            need[p["exec_to_canonical"][int(sro[k]) + local]] = int((s["col_idx"][b:b + int(sw[q])] != 0xFFFFFFFF).sum())
    h = int(rng.integers(1, 9))
    off = np.unique(np.concatenate([np.arange(0, rows, h), [rows]])).astype(np.uint32) if rows else np.zeros(1, np.uint32)
    wid = np.array([need[off[i]:off[i + 1]].max() + int(rng.integers(0, 3)) for i in range(len(off) - 1)], np.uint32)
    p["canonical_slice_row_offsets"] = off
    p["canonical_slice_widths"] = wid
    return p


@pytest.mark.parametrize("vb", [1, 2, 4])
def test_canonical_sliced_rebuild_matches_restatement(cws, tmp_path, vb):
    from cellshard_amd.ops import rebuild_canonical_sliced_ell
    rng = np.random.default_rng(70 + vb)
    want = [_with_canonical_slices(rng, random_sell_partition(rng, r, c, n, vb)) for r, c, n in [(200, 300, 4), (31, 40, 7), (1, 5, 1)]]
    path = tmp_path / "s.cspack"
    path.write_bytes(O.cspack_bytes(0, [O.sell_execution_blob(w) for w in want]))
    ws = CspackWorkspace()
    with CspackFile(path) as f:
        for i, w in enumerate(want):
            print(f"partition {i} rows {w['rows']}", flush=True)
            part = stage_cspack_sliced_ell(f, i, VT[vb], ws)
            got = rebuild_canonical_sliced_ell(part, cws)
            slots, col, val = O.canonical_sliced_of_bucketed_sell(w, vb)
            o, wd, so, gc, gv = got.to_numpy()
            assert np.array_equal(o.view(np.uint32), w["canonical_slice_row_offsets"])
            assert np.array_equal(wd.view(np.uint32), w["canonical_slice_widths"])
            assert np.array_equal(so.view(np.uint32), slots)
            assert np.array_equal(gc.view(np.uint32), col) and np.array_equal(gv.view(val.dtype), val)
            part.close()
    ws.close()


def test_canonical_sliced_round_trip(cws, tmp_path):
    """canonical Sliced-ELL -> bucketed exec partition -> pack -> staged ->
    rebuilt: the original, slot for slot."""
    from cellshard_amd.ops import rebuild_canonical_sliced_ell
    dev = _sell_dev(150, 200, 32, 99)
    orig = dev.to_numpy()
    part = build_bucketed_sliced_ell_device(dev, 3, cws)
    path = tmp_path / "rt.cspack"
    ws = CspackWorkspace()
    write_cspack_sliced_ell(path, 1, [part], 3, ws)
    with CspackFile(path, 1) as f:
        staged = stage_cspack_sliced_ell(f, 0, 3, ws)
    back = rebuild_canonical_sliced_ell(staged, cws)
    assert back.slice_rows == dev.slice_rows and back.nnz == dev.nnz
    for a, b in zip(back.to_numpy(), orig):
        assert np.array_equal(a, b)
    staged.close()
    part.close()
    ws.close()


def test_canonical_sliced_rebuild_rejects_narrow_slices(cws, tmp_path):
    from cellshard_amd.ops import rebuild_canonical_sliced_ell
    rng = np.random.default_rng(5)
    w = _with_canonical_slices(rng, random_sell_partition(rng, 60, 50, 2, 2))
    w["canonical_slice_widths"] = np.maximum(w["canonical_slice_widths"].astype(np.int64) - 3, 0).astype(np.uint32)
    with pytest.raises(ValueError):
        O.canonical_sliced_of_bucketed_sell(w, 2)
    path = tmp_path / "n.cspack"
    path.write_bytes(O.cspack_bytes(0, [O.sell_execution_blob(w)]))
    ws = CspackWorkspace()
    with CspackFile(path) as f:
        part = stage_cspack_sliced_ell(f, 0, 3, ws)
    with pytest.raises(RuntimeError):
        rebuild_canonical_sliced_ell(part, cws)
    assert "exceed" in _lib.last_error()
    part.close()
    ws.close()
