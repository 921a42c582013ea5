"""numpy front end of the CPU oracle (oracle.c) and of the reference CPU
converter built under oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libref_convert.so"

VT_F16, VT_F32, VT_BF16, VT_FP8 = 3, 4, 6, 7
VB = {VT_F16: 2, VT_F32: 4, VT_BF16: 2, VT_FP8: 1}
INVALID = 0xFFFFFFFF

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing; run __graft_entry__.build()")
        _lib = C.CDLL(str(LIB_PATH))
        _lib.oracle_half_to_float.restype = C.c_float
        _lib.oracle_half_to_float.argtypes = [C.c_uint16]
        _lib.oracle_fp8_e4m3_to_float.restype = C.c_float
        _lib.oracle_fp8_e4m3_to_float.argtypes = [C.c_uint8]
        _lib.oracle_kept_features.restype = C.c_uint32
        _lib.oracle_log2_bucket.restype = C.c_uint32
        _lib.oracle_log2_bucket.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    return _lib


def ref_available() -> bool:
    return REF_PATH.exists()


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(str(REF_PATH))
    return _ref


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _u32(a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _u8(a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _f(x):
    return C.c_float(float(x))


def _ok(rc, what):
    if rc != 1:
        raise RuntimeError(f"oracle {what} failed")


# ---------------------------------------------------------------------------
def _reduce_outputs(rows, G):
    return dict(row_totals=np.zeros(rows, np.float64), detected=np.zeros(rows, np.uint32),
                max_values=np.zeros(rows, np.float32), keep=np.zeros(rows, np.uint8),
                group_counts=np.zeros(rows * max(G, 1), np.float64), group_pct=np.zeros(rows * max(G, 1), np.float32))


def _filter_args(filt):
    return [_f(filt[0]), C.c_uint32(int(filt[1])), _f(filt[2]), C.c_uint32(int(filt[3]))]


def _finish(o, rows, G):
    o["group_counts"] = o["group_counts"][: rows * G].reshape(rows, G)
    o["group_pct"] = o["group_pct"][: rows * G].reshape(rows, G)
    return o


def group_reduce_csr(rows, cols, ptr, idx, val, vt, row_keep, feature_keep, fgm, G, filt):
    ptr, idx, row_keep, feature_keep, fgm = _u32(ptr), _u32(idx), _u8(row_keep), _u8(feature_keep), _u32(fgm)
    val = np.ascontiguousarray(val)
    o = _reduce_outputs(rows, G)
    _ok(lib().oracle_group_reduce_csr(C.c_uint32(rows), C.c_uint32(cols), _p(ptr), _p(idx), _p(val), C.c_uint32(vt),
                                      _p(row_keep), _p(feature_keep), _p(fgm), C.c_uint32(G), *_filter_args(filt),
                                      *[_p(o[k]) for k in ("row_totals", "detected", "max_values", "keep", "group_counts",
                                                           "group_pct")]), "group_reduce_csr")
    return _finish(o, rows, G)


def group_reduce_sell(rows, cols, slice_rows, offsets, widths, slot_offsets, col_idx, val, vt, row_keep, feature_keep, fgm, G,
                      filt):
    offsets, widths, slot_offsets, col_idx = _u32(offsets), _u32(widths), _u32(slot_offsets), _u32(col_idx)
    row_keep, feature_keep, fgm = _u8(row_keep), _u8(feature_keep), _u32(fgm)
    val = np.ascontiguousarray(val)
    o = _reduce_outputs(rows, G)
    _ok(lib().oracle_group_reduce_sell(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(len(offsets) - 1), C.c_uint32(slice_rows),
                                       _p(offsets), _p(widths), _p(slot_offsets), _p(col_idx), _p(val), C.c_uint32(vt),
                                       _p(row_keep), _p(feature_keep), _p(fgm), C.c_uint32(G), *_filter_args(filt),
                                       *[_p(o[k]) for k in ("row_totals", "detected", "max_values", "keep", "group_counts",
                                                            "group_pct")]), "group_reduce_sell")
    return _finish(o, rows, G)


def group_reduce_bell(rows, cols, block_size, ell_cols, bci, val, vt, row_keep, feature_keep, fgm, G, filt):
    bci, row_keep, feature_keep, fgm = _u32(bci), _u8(row_keep), _u8(feature_keep), _u32(fgm)
    val = np.ascontiguousarray(val)
    o = _reduce_outputs(rows, G)
    _ok(lib().oracle_group_reduce_bell(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(block_size), C.c_uint32(ell_cols), _p(bci),
                                       _p(val), C.c_uint32(vt), _p(row_keep), _p(feature_keep), _p(fgm), C.c_uint32(G),
                                       *_filter_args(filt),
                                       *[_p(o[k]) for k in ("row_totals", "detected", "max_values", "keep", "group_counts",
                                                            "group_pct")]), "group_reduce_bell")
    return _finish(o, rows, G)


# ---------------------------------------------------------------------------
def kept_features(cols, feature_keep):
    fk = _u8(feature_keep)
    ids = np.zeros(max(cols, 1), np.uint32)
    F = lib().oracle_kept_features(C.c_uint32(cols), _p(fk), _p(ids))
    return ids[:F]


def _label_outputs(G, F):
    return (np.zeros(G * max(F, 1), np.float64), np.zeros(G * max(F, 1), np.float64), np.zeros(G * max(F, 1), np.uint32),
            np.zeros(G, np.uint32))


def _label_finish(G, F, s, q, n, gr, ids):
    return dict(sum=s[: G * F].reshape(G, F), sumsq=q[: G * F].reshape(G, F), nnz=n[: G * F].reshape(G, F), group_rows=gr,
                kept_feature_ids=ids)


def label_reduce_csr(rows, cols, ptr, idx, val, vt, labels, row_keep, feature_keep, G, scale=None):
    ids = kept_features(cols, feature_keep)
    F = len(ids)
    s, q, n, gr = _label_outputs(G, F)
    ptr, idx, labels, row_keep, feature_keep = _u32(ptr), _u32(idx), _u32(labels), _u8(row_keep), _u8(feature_keep)
    scale = None if scale is None else np.ascontiguousarray(scale, dtype=np.float32)
    val = np.ascontiguousarray(val)
    _ok(lib().oracle_label_reduce_csr(C.c_uint32(rows), C.c_uint32(cols), _p(ptr), _p(idx), _p(val), C.c_uint32(vt), _p(scale),
                                      _p(labels), _p(row_keep), _p(feature_keep), C.c_uint32(G), _p(s), _p(q), _p(n), _p(gr)),
        "label_reduce_csr")
    return _label_finish(G, F, s, q, n, gr, ids)


def worker_threads() -> int:
    """Host threads for the chunked oracle: the cores this process may run on,
    capped by OMP_NUM_THREADS when that is set (the CPU share of a GPU box)."""
    import os

    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    try:
        cap = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        cap = 0
    return max(1, min(n, cap) if cap > 0 else n)


def label_reduce_csr_chunked(rows, cols, ptr, idx, val, vt, labels, row_keep, feature_keep, G, scale=None, threads=None,
                             chunks_per_thread=4):
    """label_reduce_csr over contiguous row chunks on `threads` host threads
    (ctypes releases the GIL during each call).  Chunk partials are added in
    FP64 in chunk order: counts are exact, sums equal the one-pass oracle's
    up to FP64 rounding.  ptr may hold offsets up to 2^32 - 1 (any integer
    dtype); idx / val are the whole arrays."""
    from concurrent.futures import ThreadPoolExecutor

    threads = threads or worker_threads()
    ptr = np.asarray(ptr).astype(np.int64, copy=False)
    n = max(1, min(rows, threads * chunks_per_thread)) if rows else 1
    cuts = np.linspace(0, rows, n + 1).astype(np.int64)
    labels, row_keep = _u32(labels), _u8(row_keep)

    def chunk(k):
        a, b = int(cuts[k]), int(cuts[k + 1])
        pa = (ptr[a:b + 1] - ptr[a]).astype(np.uint32)
        return label_reduce_csr(b - a, cols, pa, idx[ptr[a]:ptr[b]], val[ptr[a]:ptr[b]], vt, labels[a:b],
                                None if row_keep is None else row_keep[a:b], feature_keep, G, scale)

    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(chunk, range(n)))
    out = dict(parts[0])
    for key in ("sum", "sumsq"):
        out[key] = np.sum([p[key] for p in parts], axis=0, dtype=np.float64) if len(parts) > 1 else parts[0][key]
    for key in ("nnz", "group_rows"):
        out[key] = np.sum([p[key].astype(np.uint64) for p in parts], axis=0).astype(np.uint32)
    return out


def label_reduce_sell(rows, cols, slice_rows, offsets, widths, slot_offsets, col_idx, val, vt, labels, row_keep, feature_keep, G,
                      scale=None):
    ids = kept_features(cols, feature_keep)
    F = len(ids)
    s, q, n, gr = _label_outputs(G, F)
    offsets, widths, slot_offsets, col_idx = _u32(offsets), _u32(widths), _u32(slot_offsets), _u32(col_idx)
    labels, row_keep, feature_keep = _u32(labels), _u8(row_keep), _u8(feature_keep)
    scale = None if scale is None else np.ascontiguousarray(scale, dtype=np.float32)
    val = np.ascontiguousarray(val)
    _ok(lib().oracle_label_reduce_sell(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(len(offsets) - 1), C.c_uint32(slice_rows),
                                       _p(offsets), _p(widths), _p(slot_offsets), _p(col_idx), _p(val), C.c_uint32(vt), _p(scale),
                                       _p(labels), _p(row_keep), _p(feature_keep), C.c_uint32(G), _p(s), _p(q), _p(n), _p(gr)),
        "label_reduce_sell")
    return _label_finish(G, F, s, q, n, gr, ids)


def label_reduce_bell(rows, cols, block_size, ell_cols, bci, val, vt, labels, row_keep, feature_keep, G, scale=None):
    ids = kept_features(cols, feature_keep)
    F = len(ids)
    s, q, n, gr = _label_outputs(G, F)
    bci, labels, row_keep, feature_keep = _u32(bci), _u32(labels), _u8(row_keep), _u8(feature_keep)
    scale = None if scale is None else np.ascontiguousarray(scale, dtype=np.float32)
    val = np.ascontiguousarray(val)
    _ok(lib().oracle_label_reduce_bell(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(block_size), C.c_uint32(ell_cols), _p(bci),
                                       _p(val), C.c_uint32(vt), _p(scale), _p(labels), _p(row_keep), _p(feature_keep),
                                       C.c_uint32(G), _p(s), _p(q), _p(n), _p(gr)), "label_reduce_bell")
    return _label_finish(G, F, s, q, n, gr, ids)


# ---------------------------------------------------------------------------
def blocked_ell_shape(rows, cols, ptr, idx, bs):
    ptr, idx = _u32(ptr), _u32(idx)
    w, t, f = C.c_uint32(), C.c_uint64(), C.c_double()
    _ok(lib().oracle_blocked_ell_shape(C.c_uint32(rows), C.c_uint32(cols), _p(ptr), _p(idx), C.c_uint32(bs), C.byref(w),
                                       C.byref(t), C.byref(f)), "blocked_ell_shape")
    return w.value, t.value, f.value


def choose_blocked_ell_block_size(rows, cols, ptr, idx, candidates, value_bytes=2):
    ptr, idx, cands = _u32(ptr), _u32(idx), _u32(candidates)
    bs, fill, pad = C.c_uint32(), C.c_double(), C.c_uint64()
    _ok(lib().oracle_choose_blocked_ell_block_size(C.c_uint32(rows), C.c_uint32(cols), _p(ptr), _p(idx), _p(cands),
                                                   C.c_uint32(len(cands)), C.c_uint32(value_bytes), C.byref(bs), C.byref(fill),
                                                   C.byref(pad)), "choose_blocked_ell_block_size")
    return bs.value, fill.value, pad.value


def blocked_ell_from_compressed(rows, cols, ptr, idx, val, bs):
    """Returns (block_col_idx, val_bits, ell_cols, nnz)."""
    w, _, _ = blocked_ell_shape(rows, cols, ptr, idx, bs)
    ptr, idx = _u32(ptr), _u32(idx)
    val = np.ascontiguousarray(val)
    vb = val.dtype.itemsize
    rb = (rows + bs - 1) // bs
    bci = np.zeros(max(rb * w, 1), np.uint32)
    out = np.zeros(max(rows * w * bs, 1), val.dtype)
    nnz = C.c_uint32()
    _ok(lib().oracle_blocked_ell_from_compressed(C.c_uint32(rows), C.c_uint32(cols), _p(ptr), _p(idx), _p(val), C.c_uint32(vb),
                                                 C.c_uint32(bs), C.c_uint32(w), _p(bci), _p(out), C.byref(nnz)),
        "blocked_ell_from_compressed")
    return bci[: rb * w], out[: rows * w * bs], w * bs, nnz.value


def compressed_row_nnz(rows, cols, axis, ptr, idx):
    ptr, idx = _u32(ptr), _u32(idx)
    out = np.zeros(max(rows, 1), np.uint32)
    _ok(lib().oracle_compressed_row_nnz(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(axis), _p(ptr), _p(idx), _p(out)),
        "compressed_row_nnz")
    return out[:rows]


def sliced_ell_layout(row_nnz, rows, slice_rows):
    rn = _u32(row_nnz) if rows else np.zeros(1, np.uint32)
    s = min(max(slice_rows, 1), rows) if rows else 0
    count = (rows + s - 1) // s if rows else 0
    offs = np.zeros(count + 1, np.uint32)
    widths = np.zeros(max(count, 1), np.uint32)
    slots = np.zeros(max(count, 1), np.uint32)
    c, t = C.c_uint32(), C.c_uint64()
    rc = lib().oracle_sliced_ell_layout(_p(rn), C.c_uint32(rows), C.c_uint32(slice_rows), _p(offs), _p(widths), _p(slots),
                                        C.byref(c), C.byref(t))
    if rc != 1:
        raise OverflowError("sliced_ell layout: total slots exceed u32")
    return s, offs, widths[:count], slots[:count], t.value


def sliced_ell_from_compressed(rows, cols, axis, ptr, idx, val, slice_rows):
    rn = compressed_row_nnz(rows, cols, axis, ptr, idx)
    s, offs, widths, slots, total = sliced_ell_layout(rn, rows, slice_rows)
    ptr, idx = _u32(ptr), _u32(idx)
    val = np.ascontiguousarray(val)
    col = np.zeros(max(total, 1), np.uint32)
    out = np.zeros(max(total, 1), val.dtype)
    _ok(lib().oracle_sliced_ell_from_compressed(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(axis), _p(ptr), _p(idx), _p(val),
                                                C.c_uint32(val.dtype.itemsize), C.c_uint32(len(offs) - 1), _p(offs), _p(widths),
                                                _p(slots), C.c_uint64(total), _p(col), _p(out)), "sliced_ell_from_compressed")
    return dict(slice_rows=s, offsets=offs, widths=widths, slot_offsets=slots, col_idx=col[:total], val=out[:total])


def blocked_ell_row_nnz(rows, cols, bs, ell_cols, bci, val):
    bci = _u32(bci)
    val = np.ascontiguousarray(val)
    out = np.zeros(max(rows, 1), np.uint32)
    _ok(lib().oracle_blocked_ell_row_nnz(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(bs), C.c_uint32(ell_cols), _p(bci),
                                         _p(val), C.c_uint32(val.dtype.itemsize), _p(out)), "blocked_ell_row_nnz")
    return out[:rows]


def sliced_ell_from_blocked_ell(rows, cols, bs, ell_cols, bci, val, slice_rows):
    rn = blocked_ell_row_nnz(rows, cols, bs, ell_cols, bci, val)
    s, offs, widths, slots, total = sliced_ell_layout(rn, rows, slice_rows)
    bci = _u32(bci)
    val = np.ascontiguousarray(val)
    col = np.zeros(max(total, 1), np.uint32)
    out = np.zeros(max(total, 1), val.dtype)
    _ok(lib().oracle_sliced_ell_from_blocked_ell(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(bs), C.c_uint32(ell_cols), _p(bci),
                                                 _p(val), C.c_uint32(val.dtype.itemsize), C.c_uint32(len(offs) - 1), _p(offs),
                                                 _p(widths), _p(slots), C.c_uint64(total), _p(col), _p(out)),
        "sliced_ell_from_blocked_ell")
    return dict(slice_rows=s, offsets=offs, widths=widths, slot_offsets=slots, col_idx=col[:total], val=out[:total])


def choose_sliced_ell_slice_rows(row_nnz, rows, nnz, candidates):
    rn, cands = _u32(row_nnz) if rows else np.zeros(1, np.uint32), _u32(candidates)
    s, c, p, sc = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_double()
    _ok(lib().oracle_choose_sliced_ell_slice_rows(_p(rn), C.c_uint32(rows), C.c_uint32(nnz), _p(cands), C.c_uint32(len(cands)),
                                                  C.byref(s), C.byref(c), C.byref(p), C.byref(sc)), "choose_sliced_ell_slice_rows")
    return dict(slice_rows=s.value, slice_count=c.value, padded_slots=p.value, score=sc.value)


# ---------------------------------------------------------------------------
def transpose(cdim, udim, ptr, idx, val):
    ptr, idx = _u32(ptr), _u32(idx)
    val = np.ascontiguousarray(val)
    nnz = len(idx)
    optr = np.zeros(udim + 1, np.uint32)
    oidx = np.zeros(max(nnz, 1), np.uint32)
    oval = np.zeros(max(nnz, 1), val.dtype)
    _ok(lib().oracle_transpose(C.c_uint32(cdim), C.c_uint32(udim), _p(ptr), _p(idx), _p(val), C.c_uint32(val.dtype.itemsize),
                               _p(optr), _p(oidx), _p(oval)), "transpose")
    return optr, oidx[:nnz], oval[:nnz]


def bucket_plan(ptr, n, requested):
    ptr = _u32(ptr)
    nnz, srt, order = (np.zeros(max(n, 1), np.uint32) for _ in range(3))
    offs = np.zeros(max(n, 1) + 2, np.uint32)
    b = C.c_uint32()
    _ok(lib().oracle_bucket_plan(_p(ptr), C.c_uint32(n), C.c_uint32(requested), _p(nnz), _p(srt), _p(order), _p(offs), C.byref(b)),
        "bucket_plan")
    return dict(major_nnz=nnz[:n], major_nnz_sorted=srt[:n], major_order=order[:n], bucket_offsets=offs[: b.value + 1],
                bucket_count=b.value)


def log2_offsets(nnz_sorted, B=12, min_log2=5):
    s = _u32(nnz_sorted)
    offs = np.zeros(B + 1, np.uint32)
    _ok(lib().oracle_log2_offsets(_p(s) if len(s) else None, C.c_uint32(len(s)), C.c_uint32(B), C.c_uint32(min_log2), _p(offs)),
        "log2_offsets")
    return offs


def log2_bucket(v, B=12, min_log2=5):
    return lib().oracle_log2_bucket(int(v), B, min_log2)


def rebuild(ptr, idx, val, order):
    ptr, idx, order = _u32(ptr), _u32(idx), _u32(order)
    val = np.ascontiguousarray(val)
    n = len(order)
    dptr = np.zeros(n + 1, np.uint32)
    didx = np.zeros(max(len(idx), 1), np.uint32)
    dval = np.zeros(max(len(idx), 1), val.dtype)
    inv = np.zeros(max(n, 1), np.uint32)
    _ok(lib().oracle_rebuild(_p(ptr), _p(idx), _p(val), C.c_uint32(val.dtype.itemsize), C.c_uint32(n), _p(order), _p(dptr),
                             _p(didx), _p(dval), _p(inv)), "rebuild")
    return dptr, didx[: len(idx)], dval[: len(idx)], inv[:n]


def bell_bucket_keys(row_blocks, width, bci):
    bci = _u32(bci) if len(bci) else np.zeros(1, np.uint32)
    keys = np.zeros(max(row_blocks, 1), np.uint64)
    lib().oracle_bell_bucket_keys(C.c_uint32(row_blocks), C.c_uint32(width), _p(bci), _p(keys))
    return keys[:row_blocks]


def reoptimize(rows, cols, bs, ell_cols, bci, val16, row_keep, feature_keep, row_remap, feature_remap, out_rows, out_cols):
    bci, row_keep, feature_keep = _u32(bci), _u8(row_keep), _u8(feature_keep)
    row_remap, feature_remap = _u32(row_remap), _u32(feature_remap)
    val16 = np.ascontiguousarray(val16, dtype=np.uint16)
    maxn = rows * ell_cols
    mp = np.zeros(out_rows + 1, np.uint32)
    mi = np.zeros(max(maxn, 1), np.uint32)
    mv = np.zeros(max(maxn, 1), np.uint16)
    nnz = C.c_uint32()
    _ok(lib().oracle_reoptimize(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(bs), C.c_uint32(ell_cols), _p(bci), _p(val16),
                                _p(row_keep), _p(feature_keep), _p(row_remap), _p(feature_remap), C.c_uint32(out_rows),
                                C.c_uint32(out_cols), _p(mp), _p(mi), _p(mv), C.byref(nnz)), "reoptimize")
    return mp, mi[: nnz.value], mv[: nnz.value], nnz.value


def _stored_nonzero(bits, vt):
    bits = np.asarray(bits)
    if vt == VT_F32:
        return bits.view(np.float32) != 0.0
    if vt == VT_FP8:
        return (bits.view(np.uint8) & 0x7F) != 0
    return (bits.view(np.uint16) & 0x7FFF) != 0


def _filter_rows(part, vt, output_rows, output_cols, keep_rows, keep_cols, row_remap, col_remap, per, claimed, row0=0):
    """Surviving entries of one Blocked-ELL part, per output row (in place
    into `per`); returns (kept rows, kept rows' ell slots)."""
    rows, cols, bs, ell_cols = part["rows"], part["cols"], part["block_size"], part["ell_cols"]
    W = ell_cols // bs if bs else 0
    bci = np.asarray(part["bci"], np.uint32).astype(np.int64)
    vals = np.asarray(part["val"]).reshape(rows, ell_cols) if rows * ell_cols else None
    kept = 0
    e = np.arange(ell_cols)
    slot, lane = (e // bs, e % bs) if bs else (e, e)
    for r in range(rows):
        if keep_rows is not None and keep_rows[r] == 0:
            continue
        kept += 1
        o = int(row_remap[r]) if row_remap is not None else row0 + r  # shard row without a remap
        if o >= output_rows or not output_cols:
            continue
        if o in claimed:
            raise ValueError(f"two kept rows map to output row {o}")
        claimed.add(o)
        if vals is None:
            continue
        bc = bci[(r // bs) * W + slot]
        valid = bc != INVALID
        col = np.where(valid, bc * bs + lane, cols)
        m = valid & (col < cols) & _stored_nonzero(vals[r], vt)
        cs = np.minimum(col, max(cols - 1, 0))
        if keep_cols is not None:
            m &= np.asarray(keep_cols)[cs] != 0
        oc = np.asarray(col_remap, np.int64)[cs] if col_remap is not None else col
        m &= oc < output_cols
        per[o] = (oc[m].astype(np.uint32), vals[r][m])
    return kept, kept * ell_cols


def _csr_of(per, output_rows, dtype):
    counts = np.array([len(per[o][0]) if o in per else 0 for o in range(output_rows)], np.int64)
    ptr = np.zeros(output_rows + 1, np.uint32)
    ptr[1:] = np.cumsum(counts)
    idx = np.concatenate([per[o][0] for o in range(output_rows) if o in per] or [np.zeros(0, np.uint32)])
    val = np.concatenate([per[o][1] for o in range(output_rows) if o in per] or [np.zeros(0, dtype)])
    return ptr, idx.astype(np.uint32), val.astype(dtype)


def filter_blocked_ell(rows, cols, bs, ell_cols, bci, val_bits, vt, output_rows, output_cols, keep_rows=None, keep_cols=None,
                       row_remap=None, col_remap=None):
    """Masked Blocked-ELL -> CSR, restated from
    src/convert/filtered_blocked_ell_to_compressed.cuh:29-123 (which entries
    survive, in ell order) and :264-373 (stats: kept_rows = output_rows,
    slots = output_rows x ell_cols).  Row pointers are the plain exclusive
    scan of the per-row counts: row i = [ptr[i], ptr[i + 1]).
    Returns (ptr, idx, val_bits, stats dict)."""
    val_bits = np.asarray(val_bits)
    part = {"rows": rows, "cols": cols, "block_size": bs, "ell_cols": ell_cols, "bci": bci, "val": val_bits}
    per = {}
    if output_rows and output_cols:
        _filter_rows(part, vt, output_rows, output_cols, keep_rows, keep_cols, row_remap, col_remap, per, set())
    ptr, idx, val = _csr_of(per, output_rows, val_bits.dtype)
    live = len(idx)
    slots = output_rows * ell_cols if output_rows and output_cols else 0
    stats = {"kept_rows": output_rows, "output_rows": output_rows, "output_cols": output_cols, "live_nnz": live,
             "kept_row_slots": slots, "dead_slots": max(slots - live, 0),
             "live_fill_ratio": 1.0 if slots == 0 else live / slots}
    return ptr, idx, val, stats


def repack_blocked_ell(parts, vt, output_rows, output_cols, block_size, requested_bucket_count, keep_rows=None, keep_cols=None,
                       row_remap=None, col_remap=None):
    """Shard repack, restated from src/repack/operators/sharded_blocked_ell.cuh:393-723
    and repack/kernels/sharded_blocked_ell.cuh:23-128: the parts' rows in
    shard order are filtered to CSR (kept rows counted over the shard, slots
    priced at each kept row's own ell width), bucketed by stable ascending
    nnz (equal-count cuts), and the bucketed rows packed as Blocked-ELL of
    `block_size` (ascending block columns, W = widest row block).  Row
    pointers as in filter_blocked_ell."""
    dtype = np.asarray(parts[0]["val"]).dtype if parts else np.uint16
    per, claimed = {}, set()
    kept = slots = 0
    r0 = 0
    for part in parts:
        n = part["rows"]
        sl = slice(r0, r0 + n)
        k, s_ = _filter_rows(part, vt, output_rows, output_cols, None if keep_rows is None else np.asarray(keep_rows)[sl], keep_cols,
                             None if row_remap is None else np.asarray(row_remap)[sl], col_remap, per, claimed, r0)
        kept, slots, r0 = kept + k, slots + s_, r0 + n
    if not (output_rows and output_cols):
        per = {}
    ptr, idx, val = _csr_of(per, output_rows, dtype)
    live = len(idx)
    stats = {"kept_rows": kept, "output_rows": output_rows, "output_cols": output_cols, "live_nnz": live, "kept_row_slots": slots,
             "dead_slots": max(slots - live, 0), "live_fill_ratio": 1.0 if slots == 0 else live / slots, "ell_width": 0}
    out = {"ptr": ptr, "idx": idx, "val": val, "stats": stats}
    if not (output_rows and output_cols and live):
        return out
    plan = bucket_plan(ptr, output_rows, requested_bucket_count)
    order = plan["major_order"]
    lens = np.diff(ptr.astype(np.int64))[order]
    rptr = np.zeros(output_rows + 1, np.uint32)
    rptr[1:] = np.cumsum(lens)
    ridx = np.concatenate([idx[ptr[o]:ptr[o + 1]] for o in order])
    rval = np.concatenate([val[ptr[o]:ptr[o + 1]] for o in order])
    bci, bv, ell, nnz = blocked_ell_from_compressed(output_rows, output_cols, rptr, ridx, rval, block_size)
    stats["ell_width"] = ell // block_size
    out.update(plan=plan, rptr=rptr, ridx=ridx, rval=rval, bell_bci=bci, bell_val=bv, bell_ell_cols=ell)
    return out


_FNV_PRIME = 1099511628211
_M64 = (1 << 64) - 1


def _mix_signature(seed, value):
    for k in range(8):
        seed ^= (value >> (8 * k)) & 0xFF
        seed = (seed * _FNV_PRIME) & _M64
    return seed


def shard_column_maps(parts, vt, cols):
    """Column order by row-block support signature, restated from
    include/CellShard/ingest/internal/dataset/internal/convert_layout_support_part.hh:20-86.
    parts: Blocked-ELL dicts (rows, block_size, ell_cols, bci, val bits) in
    shard order.  Returns (exec_to_canonical, canonical_to_exec)."""
    support = [0] * cols
    ha = [1469598103934665603] * cols
    hb = [_FNV_PRIME] * cols
    first = [INVALID] * cols
    g = 0
    for part in parts:
        rows, bs, ell = part["rows"], part["block_size"], part["ell_cols"]
        W = ell // bs if bs else 0
        bci = np.asarray(part["bci"], np.uint32)
        vals = np.asarray(part["val"]).reshape(rows, ell) if rows * ell else None
        nzm = _stored_nonzero(vals, vt) if vals is not None else None
        for rb in range((rows + bs - 1) // bs if bs else 0):
            r0, r1 = rb * bs, min(rows, rb * bs + bs)
            for slot in range(W):
                bc = int(bci[rb * W + slot])
                if bc == INVALID:
                    continue
                for lane in range(bs):
                    col = bc * bs + lane
                    if col >= cols or not nzm[r0:r1, slot * bs + lane].any():
                        continue
                    support[col] += 1
                    ha[col] = _mix_signature(ha[col], g + 1)
                    hb[col] = _mix_signature(hb[col], ((g + 1) * 1315423911) & _M64)
                    first[col] = min(first[col], g)
            g += 1
    order = sorted(range(cols), key=lambda c: (support[c] == 0, first[c], ha[c], hb[c], -support[c], c))
    e2c = np.array(order, np.uint32)
    c2e = np.zeros(cols, np.uint32)
    c2e[e2c] = np.arange(cols, dtype=np.uint32)
    return e2c, c2e


def _ordered_f32(v):
    bits = np.asarray(v, np.float32).view(np.uint32).astype(np.uint64)
    return np.where(bits & 0x80000000, bits ^ 0xFFFFFFFF, bits ^ 0x80000000)


def barycenter_order(rows, cols, row_idx, col_idx, other_rank, by_col, float_sums=False):
    """Entities (columns when by_col, else rows) sorted by the mean rank of
    the entities they touch, then by index; FLT_MAX when they touch none.
    Restated from src/bucket/operators/blocked_ell_bipartite_optimize.cuh:223-309.
    float_sums: accumulate in float32 in COO order as the reference's atomics
    would in that order; otherwise exact sums rounded once.
    Returns (exec_to_canonical, canonical_to_exec)."""
    n = cols if by_col else rows
    ent = np.asarray(col_idx if by_col else row_idx, np.int64)
    oth = np.asarray(row_idx if by_col else col_idx, np.int64)
    rank = np.asarray(other_rank, np.uint64)[oth] if len(oth) else np.zeros(0, np.uint64)
    ok = rank != INVALID
    count = np.bincount(ent[ok], minlength=n).astype(np.uint32)
    if float_sums:
        s_ = np.zeros(n, np.float32)
        for e_, r_ in zip(ent[ok], rank[ok]):
            s_[e_] = np.float32(s_[e_] + np.float32(r_))
        sums = s_
    else:
        sums = np.bincount(ent[ok], weights=None, minlength=n) * 0
        acc = np.zeros(n, np.uint64)
        np.add.at(acc, ent[ok], rank[ok])
        sums = acc.astype(np.float32)  # one rounding of the exact integer
    with np.errstate(divide="ignore", invalid="ignore"):
        bary = np.where(count != 0, np.float32(sums) / np.float32(count), np.float32(np.finfo(np.float32).max)).astype(np.float32)
    keys = (_ordered_f32(bary) << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    e2c = np.argsort(keys, kind="stable").astype(np.uint32)
    c2e = np.zeros(n, np.uint32)
    c2e[e2c] = np.arange(n, dtype=np.uint32)
    return e2c, c2e


# ---------------------------------------------------------------------------
# Reference CPU converter (oracle/_ref), fp16 values.
def ref_choose_blocked_ell_block_size(rows, cols, ptr, idx, val16, candidates):
    ptr, idx, cands = _u32(ptr), _u32(idx), _u32(candidates)
    val16 = np.ascontiguousarray(val16).view(np.uint16)
    bs, fill, pad = C.c_uint32(), C.c_double(), C.c_ulonglong()
    _ok(ref().ref_choose_blocked_ell_block_size(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(len(idx)), _p(ptr), _p(idx),
                                                _p(val16), _p(cands), C.c_uint32(len(cands)), C.byref(bs), C.byref(fill),
                                                C.byref(pad)), "ref choose")
    return bs.value, fill.value, pad.value


def ref_blocked_ell_from_compressed(rows, cols, ptr, idx, val16, bs):
    ptr, idx = _u32(ptr), _u32(idx)
    val16 = np.ascontiguousarray(val16).view(np.uint16)
    ell, nnz = C.c_uint32(), C.c_uint32()
    r = ref()
    _ok(r.ref_blocked_ell_from_compressed(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(len(idx)), _p(ptr), _p(idx), _p(val16),
                                          C.c_uint32(bs), C.byref(ell), C.byref(nnz), None, C.c_ulonglong(0), None,
                                          C.c_ulonglong(0)), "ref convert (size)")
    w = ell.value // bs
    rb = (rows + bs - 1) // bs
    bci = np.zeros(max(rb * w, 1), np.uint32)
    out = np.zeros(max(rows * ell.value, 1), np.uint16)
    _ok(r.ref_blocked_ell_from_compressed(C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(len(idx)), _p(ptr), _p(idx), _p(val16),
                                          C.c_uint32(bs), C.byref(ell), C.byref(nnz), _p(bci), C.c_ulonglong(len(bci)), _p(out),
                                          C.c_ulonglong(len(out))), "ref convert")
    return bci[: rb * w], out[: rows * ell.value], ell.value, nnz.value


def serve_label(seed: int, rows: int, G: int) -> np.ndarray:
    L = lib()
    L.oracle_serve_label.restype = C.c_uint32
    L.oracle_serve_label.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    return np.array([L.oracle_serve_label(seed, r, G) for r in range(rows)], dtype=np.uint32)


def serve_keep(seed: int, rows: int, thr24: int) -> np.ndarray:
    L = lib()
    L.oracle_serve_keep.restype = C.c_int
    L.oracle_serve_keep.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    return np.array([L.oracle_serve_keep(seed, r, thr24) for r in range(rows)], dtype=np.uint8)


def serve_query(rows, cols, ptr, idx, val, vt, features, G, op, labels=None, keep=None, label_seed=0, mask_seed=0,
                thr24=1 << 24):
    """Feature-subset grouped query on the CSR; returns [k x G] (f64, or u32 for op 2)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.uint32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    val = np.ascontiguousarray(val)
    features = np.ascontiguousarray(features, dtype=np.uint32)
    k = len(features)
    out = np.zeros((k, G), dtype=np.uint32 if op == 2 else np.float64)
    labels = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint32)
    keep = None if keep is None else np.ascontiguousarray(keep, dtype=np.uint8)
    _ok(lib().oracle_serve_query(C.c_uint32(rows), C.c_uint32(cols), _p(ptr), _p(idx), _p(val), C.c_uint32(vt), _p(features),
                                 C.c_uint32(k), C.c_uint32(G), C.c_uint32(op), _p(labels), _p(keep), C.c_uint32(label_seed),
                                 C.c_uint32(mask_seed), C.c_uint32(thr24), _p(out)), "serve_query")
    return out


# ---------------------------------------------------------------------------
# Bucketed execution partitions (src/io/csh5/execution_builders_part.hh).
# Layout rules restated in numpy / small Python loops; byte estimators from
# include/CellShard/io/pack/packfile.cuh:110-116 and src/io/pack/packfile.cu:
# 179-188 (16-byte / 13-byte headers, 2-byte values).
# ---------------------------------------------------------------------------
INVALID = 0xFFFFFFFF


def bell_packed_bytes(rows, ell_cols, bs):
    return 16 + 4 + 4 + ((rows + bs - 1) // bs) * (ell_cols // bs) * 4 + rows * ell_cols * 2


def sell_packed_bytes(slices, slots):
    return 13 + 4 + (slices + 1) * 4 + slices * 4 + slots * 6


def bell_execution_layout(rows, bs, ell_cols, bci, requested):
    """execution_builders_part.hh:1035-1098 + the byte fallback of :1206-1246.
    Returns dict(order, begin, width, bytes)."""
    rb = (rows + bs - 1) // bs
    W = ell_cols // bs
    widths = (np.asarray(bci, np.uint32).reshape(rb, W) != INVALID).sum(axis=1).astype(np.int64) if rb * W else np.zeros(rb, np.int64)
    sortable = rb - 1 if rb and rows - (rb - 1) * bs != bs else rb
    order = np.arange(rb)
    order[:sortable] = np.argsort(widths[:sortable], kind="stable")

    def layout(req):
        if rb == 0:
            return [0], []
        nb = max(1, min(req, rb))
        begin, width = [0], []
        for b in range(nb):
            lo, hi = b * rb // nb, (b + 1) * rb // nb
            width.append(int(widths[order[lo:hi]].max()) if hi > lo else 0)
            begin.append(hi)
        if len(width) > 1 and all(w == width[0] for w in width):
            begin, width = [0, rb], [width[0]]
        return begin, width

    def nbytes(begin, width):
        t = 0
        for s, w in enumerate(width):
            seg_rows = sum(min(bs, rows - int(q) * bs) for q in order[begin[s]:begin[s + 1]])
            t += bell_packed_bytes(seg_rows, w * bs, bs)
        return t

    begin, width = layout(requested)
    b = nbytes(begin, width)
    original = bell_packed_bytes(rows, ell_cols, bs)
    if len(width) > 1 and b >= original:
        begin, width = layout(1)
        b = original
    return dict(order=order, begin=begin, width=width, bytes=b, widths=widths)


def bell_choose_bucket_count(rows, bs, ell_cols, bci):
    """execution_builders_part.hh:1248-1275: fewest bytes over 1..min(8, row
    blocks) buckets, ties to fewer."""
    rb = (rows + bs - 1) // bs
    best = None
    for b in range(1, max(1, min(8, rb)) + 1):
        nb = bell_execution_layout(rows, bs, ell_cols, bci, b)["bytes"]
        if best is None or nb < best[1]:
            best = (b, nb)
    return best


def bell_fill_reference(rows, bs, ell_cols, bci, val, lay):
    """Segments filled as execution_builders_part.hh:1155-1204 does: row blocks
    in layout order, valid slots compacted in slot order, values copied."""
    W = ell_cols // bs
    bci = np.asarray(bci, np.uint32).reshape(-1, W) if W else np.zeros(((rows + bs - 1) // bs, 0), np.uint32)
    val = np.asarray(val).reshape(rows, ell_cols) if rows else np.asarray(val)
    segs, exec_to_canon = [], []
    for s, w in enumerate(lay["width"]):
        rbs = lay["order"][lay["begin"][s]:lay["begin"][s + 1]]
        seg_rows = sum(min(bs, rows - int(q) * bs) for q in rbs)
        sb = np.full((len(rbs), w), INVALID, np.uint32)
        sv = np.zeros((seg_rows, w * bs), val.dtype)
        r0 = 0
        for j, q in enumerate(rbs):
            q = int(q)
            rin = min(bs, rows - q * bs)
            slots = np.nonzero(bci[q] != INVALID)[0]
            sb[j, :len(slots)] = bci[q, slots]
            for d, t in enumerate(slots):
                sv[j * bs:j * bs + rin, d * bs:(d + 1) * bs] = val[q * bs:q * bs + rin, t * bs:(t + 1) * bs]
            exec_to_canon.extend(range(q * bs, q * bs + rin))
            r0 += rin
        segs.append((seg_rows, sb.reshape(-1), sv.reshape(-1)))
    return segs, np.asarray(exec_to_canon, np.uint32)


def sell_row_widths(rows, offsets, widths, slot_offsets, col_idx):
    out = np.zeros(rows, np.int64)
    for s in range(len(widths)):
        for r in range(int(offsets[s]), int(offsets[s + 1])):
            b = int(slot_offsets[s]) + (r - int(offsets[s])) * int(widths[s])
            out[r] = int((np.asarray(col_idx[b:b + int(widths[s])]) != INVALID).sum())
    return out


def sell_min_split_cost(sorted_widths, k):
    """Brute-force DP (O(k n^2)): the fewest bytes of exactly k non-empty
    segments of the sorted widths, a segment (i, j] costing
    29 + 6 x w[j-1] x (j - i) (one slice, 2-byte values)."""
    w = [int(x) for x in sorted_widths]
    n = len(w)
    F = sell_packed_bytes(1, 0)
    INF = float("inf")
    dp = [0] + [INF] * n
    for _ in range(k):
        nd = [INF] * (n + 1)
        for j in range(1, n + 1):
            for i in range(j):
                if dp[i] < INF:
                    nd[j] = min(nd[j], dp[i] + F + 6 * w[j - 1] * (j - i))
        dp = nd
    return dp[n]


# ---------------------------------------------------------------------------
# CSPACK01 shard packs (docs/SPEC_CSPACK_V1.md:87-155).  Partitions are dicts
# with the keys of cellshard_amd.ops.BucketedBlockedEll / BucketedSlicedEll
# .to_host(); value arrays are raw value bits whose itemsize is the stored
# width (2 for the reference's f16 storage).
# ---------------------------------------------------------------------------
CSPACK_MAGIC = b"CSPACK01"
FORMAT_BLOCKED_ELL, FORMAT_SLICED_ELL = 6, 7   # io/common/raw_format.hh:10-24
REF_PACK_PATH = HERE / "_ref" / "libref_pack.so"
_ref_pack = None


def _le_u32(*vals) -> bytes:
    return np.asarray(vals, dtype="<u4").tobytes()


def _arr_bytes(a, dtype) -> bytes:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype)).astype(np.dtype(dtype).newbyteorder("<"), copy=False).tobytes()


def raw_blocked_ell_bytes(rows, cols, nnz, block_size, ell_cols, block_col_idx, val) -> bytes:
    """store_blocked_ell_raw (src/io/pack/packfile.cu:351-369; header
    :32-38): format u8, rows, cols, nnz, block_size, ell_cols, then
    blockColIdx[row_blocks * ell_width] and val[rows * ell_cols], unpadded."""
    rb = (rows + block_size - 1) // block_size if block_size else 0
    w = ell_cols // block_size if block_size else 0
    bci = np.asarray(block_col_idx, dtype=np.uint32).reshape(-1)
    v = np.ascontiguousarray(val).reshape(-1)
    assert bci.size == rb * w and v.size == rows * ell_cols
    return (bytes([FORMAT_BLOCKED_ELL]) + _le_u32(rows, cols, nnz, block_size, ell_cols) + _arr_bytes(bci, np.uint32)
            + v.tobytes())


def raw_sliced_ell_bytes(rows, cols, nnz, slice_row_offsets, slice_widths, col_idx, val) -> bytes:
    """store_sliced_ell_raw (packfile.cu:597-621): format u8, rows, cols, nnz,
    slice_count, slice_row_offsets[slice_count + 1], slice_widths[slice_count],
    col_idx[total_slots], val[total_slots]."""
    off = np.asarray(slice_row_offsets, dtype=np.uint32).reshape(-1)
    wid = np.asarray(slice_widths, dtype=np.uint32).reshape(-1)
    n = wid.size
    slots = int(((off[1:].astype(np.int64) - off[:-1]) * wid).sum()) if n else 0
    col = np.asarray(col_idx, dtype=np.uint32).reshape(-1)
    v = np.ascontiguousarray(val).reshape(-1)
    assert off.size == n + 1 and col.size == slots and v.size == slots
    return (bytes([FORMAT_SLICED_ELL]) + _le_u32(rows, cols, nnz, n) + _arr_bytes(off, np.uint32) + _arr_bytes(wid, np.uint32)
            + _arr_bytes(col, np.uint32) + v.tobytes())


def bell_execution_blob(part) -> bytes:
    """write_execution_partition_blob (src/io/csh5/execution_builders_part.hh
    :2116-2132)."""
    segs = part["segments"]
    out = [_le_u32(part["rows"], part["cols"], part["nnz"], len(segs)), _arr_bytes(part["segment_row_offsets"], np.uint32),
           _arr_bytes(part["exec_to_canonical"], np.uint32), _arr_bytes(part["canonical_to_exec"], np.uint32),
           _arr_bytes(part["col_maps"][0], np.uint32), _arr_bytes(part["col_maps"][1], np.uint32)]
    for s in segs:
        out.append(raw_blocked_ell_bytes(s["rows"], s["cols"], s["nnz"], s["block_size"], s["ell_cols"], s["block_col_idx"],
                                         s["val"]))
    return b"".join(out)


def sell_execution_blob(part) -> bytes:
    """write_sliced_execution_partition_blob (execution_runtime_part.hh
    :168-195)."""
    segs = part["segments"]
    cso = np.asarray(part["canonical_slice_row_offsets"], np.uint32)
    out = [_le_u32(part["rows"], part["cols"], part["nnz"], len(segs), cso.size - 1),
           _arr_bytes(part["segment_row_offsets"], np.uint32), _arr_bytes(part["exec_to_canonical"], np.uint32),
           _arr_bytes(part["canonical_to_exec"], np.uint32), _arr_bytes(cso, np.uint32),
           _arr_bytes(part["canonical_slice_widths"], np.uint32)]
    for s in segs:
        out.append(raw_sliced_ell_bytes(s["rows"], s["cols"], s["nnz"], s["slice_row_offsets"], s["slice_widths"], s["col_idx"],
                                        s["val"]))
    return b"".join(out)


def cspack_bytes(shard_id, blobs) -> bytes:
    """materialize_*_cspack (execution_runtime_part.hh:692-712): magic,
    shard_id u64, partition_count u64, offsets u64, blobs end to end."""
    off = 24 + 8 * len(blobs)
    offs = []
    for b in blobs:
        offs.append(off)
        off += len(b)
    return CSPACK_MAGIC + np.asarray([shard_id, len(blobs)] + offs, dtype="<u8").tobytes() + b"".join(blobs)


def cspack_header(data: bytes):
    """(shard_id, offsets) of a CSPACK01 file (execution_runtime_part.hh
    :586-598)."""
    if data[:8] != CSPACK_MAGIC:
        raise ValueError("bad magic")
    shard_id, count = np.frombuffer(data, "<u8", 2, 8)
    return int(shard_id), [int(x) for x in np.frombuffer(data, "<u8", int(count), 24)]


class _Cursor:
    def __init__(self, data, pos):
        self.d, self.p = data, pos

    def u8(self):
        v = self.d[self.p]
        self.p += 1
        return v

    def u32(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.d, "<u4", k, self.p).astype(np.uint32)
        self.p += 4 * k
        return int(a[0]) if n is None else a

    def raw(self, n, vb):
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[vb]
        a = np.frombuffer(self.d, np.dtype(dt).newbyteorder("<"), n, self.p).astype(dt)
        self.p += n * vb
        return a


def decode_bell_blob(data: bytes, off: int, vb: int) -> dict:
    """load_execution_partition_blob (execution_builders_part.hh:2146-2197)
    with load_blocked_ell_raw (packfile.cu:390-425) per segment."""
    c = _Cursor(data, off)
    rows, cols, nnz, nseg = c.u32(), c.u32(), c.u32(), c.u32()
    part = dict(rows=rows, cols=cols, nnz=nnz, segment_row_offsets=c.u32(nseg + 1), exec_to_canonical=c.u32(rows),
                canonical_to_exec=c.u32(rows))
    part["col_maps"] = (c.u32(cols), c.u32(cols))
    segs = []
    for _ in range(nseg):
        fmt = c.u8()
        assert fmt == FORMAT_BLOCKED_ELL
        r, cc, z, bs, ell = (c.u32() for _ in range(5))
        rb = (r + bs - 1) // bs if bs else 0
        w = ell // bs if bs else 0
        segs.append(dict(rows=r, cols=cc, nnz=z, block_size=bs, ell_cols=ell, block_col_idx=c.u32(rb * w), val=c.raw(r * ell, vb)))
    part["segments"] = segs
    part["end"] = c.p
    return part


def decode_sell_blob(data: bytes, off: int, vb: int) -> dict:
    """load_sliced_execution_partition_blob (execution_runtime_part.hh
    :197-258) with load_sliced_ell_raw (packfile.cu:643-697) per segment."""
    c = _Cursor(data, off)
    rows, cols, nnz, nseg, ncs = (c.u32() for _ in range(5))
    part = dict(rows=rows, cols=cols, nnz=nnz, segment_row_offsets=c.u32(nseg + 1), exec_to_canonical=c.u32(rows),
                canonical_to_exec=c.u32(rows), canonical_slice_row_offsets=c.u32(ncs + 1), canonical_slice_widths=c.u32(ncs))
    segs = []
    for _ in range(nseg):
        fmt = c.u8()
        assert fmt == FORMAT_SLICED_ELL
        r, cc, z, n = (c.u32() for _ in range(4))
        offs, wid = c.u32(n + 1), c.u32(n)
        slots = int(((offs[1:].astype(np.int64) - offs[:-1]) * wid).sum()) if n else 0
        segs.append(dict(rows=r, cols=cc, nnz=z, slice_count=n, slice_row_offsets=offs, slice_widths=wid, col_idx=c.u32(slots),
                         val=c.raw(slots, vb)))
    part["segments"] = segs
    part["end"] = c.p
    return part


def ref_pack_available() -> bool:
    return REF_PACK_PATH.exists()


def ref_pack():
    global _ref_pack
    if _ref_pack is None:
        _ref_pack = C.CDLL(str(REF_PACK_PATH))
        _ref_pack.ref_load_bell_partition.restype = C.c_longlong
        _ref_pack.ref_load_bell_partition.argtypes = [C.c_char_p, C.c_ulonglong]
    return _ref_pack


def ref_raw_blocked_ell_bytes(path, rows, cols, nnz, block_size, ell_cols, block_col_idx, val) -> bytes:
    """The reference's own store_blocked_ell_raw (oracle/_ref) into `path`."""
    bci = _u32(block_col_idx)
    v = np.ascontiguousarray(val)
    _ok(ref_pack().ref_store_blocked_ell_raw(str(path).encode(), C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(nnz),
                                             C.c_uint32(block_size), C.c_uint32(ell_cols), _p(bci), _p(v),
                                             C.c_size_t(v.itemsize)), "ref store blocked_ell")
    return Path(path).read_bytes()


def ref_raw_sliced_ell_bytes(path, rows, cols, nnz, slice_row_offsets, slice_widths, col_idx, val) -> bytes:
    off, wid, col = _u32(slice_row_offsets), _u32(slice_widths), _u32(col_idx)
    v = np.ascontiguousarray(val)
    _ok(ref_pack().ref_store_sliced_ell_raw(str(path).encode(), C.c_uint32(rows), C.c_uint32(cols), C.c_uint32(nnz),
                                            C.c_uint32(wid.size), _p(off), _p(wid), _p(col), _p(v), C.c_size_t(v.itemsize)),
        "ref store sliced_ell")
    return Path(path).read_bytes()


def ref_load_blocked_ell_raw(path, vb):
    """The reference's load_blocked_ell_raw on a raw payload file: (header
    [format, rows, cols, nnz, block_size, ell_cols], bci, val bits)."""
    hdr = np.zeros(6, np.uint32)
    _ok(ref_pack().ref_load_blocked_ell_raw(str(path).encode(), C.c_size_t(vb), _p(hdr), None, C.c_ulonglong(0), None,
                                            C.c_ulonglong(0)), "ref load blocked_ell (size)")
    rows, bs, ell = int(hdr[1]), int(hdr[4]), int(hdr[5])
    nb = ((rows + bs - 1) // bs) * (ell // bs) if bs else 0
    bci = np.zeros(max(nb, 1), np.uint32)
    val = np.zeros(max(rows * ell, 1), {1: np.uint8, 2: np.uint16, 4: np.uint32}[vb])
    _ok(ref_pack().ref_load_blocked_ell_raw(str(path).encode(), C.c_size_t(vb), _p(hdr), _p(bci), C.c_ulonglong(bci.size), _p(val),
                                            C.c_ulonglong(val.size)), "ref load blocked_ell")
    return hdr, bci[:nb], val[: rows * ell]


def ref_load_sliced_ell_raw(path, vb):
    """The reference's load_sliced_ell_raw: (header [format, rows, cols, nnz,
    slice_count, total_slots], slice_row_offsets, slice_widths, col, val)."""
    hdr = np.zeros(6, np.uint32)
    _ok(ref_pack().ref_load_sliced_ell_raw(str(path).encode(), C.c_size_t(vb), _p(hdr), None, None, None, None, C.c_ulonglong(0)),
        "ref load sliced_ell (size)")
    n, slots = int(hdr[4]), int(hdr[5])
    off, wid = np.zeros(n + 1, np.uint32), np.zeros(max(n, 1), np.uint32)
    col = np.zeros(max(slots, 1), np.uint32)
    val = np.zeros(max(slots, 1), {1: np.uint8, 2: np.uint16, 4: np.uint32}[vb])
    _ok(ref_pack().ref_load_sliced_ell_raw(str(path).encode(), C.c_size_t(vb), _p(hdr), _p(off), _p(wid), _p(col), _p(val),
                                           C.c_ulonglong(slots)), "ref load sliced_ell")
    return hdr, off, wid[:n], col[:slots], val[:slots]


def canonical_coo_of_bucketed_bell(part, vb):
    """bucketed_partition_to_canonical_coo (src/io/csh5/execution_builders_part
    .hh:909-968): per segment, per exec row, per slot with a valid block, per
    column in the block, the stored values that are not +-0 at exec columns
    < cols; rows mapped exec -> canonical, columns through the column map.
    Returns (row_idx, col_idx, val bits) in that order."""
    zero_mask = {1: 0x7F, 2: 0x7FFF, 4: 0x7FFFFFFF}[vb]
    cols = part["cols"]
    e2c = np.asarray(part["exec_to_canonical"], np.int64)
    cmap = part.get("col_maps")
    cmap = None if cmap is None else np.asarray(cmap[0], np.int64)
    out_r, out_c, out_v = [], [], []
    sro = part["segment_row_offsets"]
    for k, s in enumerate(part["segments"]):
        bs, ell = s["block_size"], s["ell_cols"]
        w = ell // bs if bs else 0
        bci = np.asarray(s["block_col_idx"], np.int64)
        val = np.asarray(s["val"])
        for r in range(s["rows"]):
            crow = int(e2c[int(sro[k]) + r])
            rb = r // bs
            for slot in range(w):
                bc = int(bci[rb * w + slot])
                if bc == INVALID:
                    continue
                for cib in range(bs):
                    ec = bc * bs + cib
                    v = int(val[r * ell + slot * bs + cib])
                    if (v & zero_mask) == 0 or ec >= cols:
                        continue
                    out_r.append(crow)
                    out_c.append(int(cmap[ec]) if cmap is not None else ec)
                    out_v.append(v)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[vb]
    return np.asarray(out_r, np.uint32), np.asarray(out_c, np.uint32), np.asarray(out_v, dt)


def canonical_csr_of_bucketed_bell(part, vb):
    """The canonical COO grouped by row with a stable sort (each canonical
    row comes from one exec row, so its entries keep the COO order)."""
    r, c, v = canonical_coo_of_bucketed_bell(part, vb)
    order = np.argsort(r, kind="stable")
    ptr = np.zeros(part["rows"] + 1, np.int64)
    np.add.at(ptr, r.astype(np.int64) + 1, 1)
    return np.cumsum(ptr).astype(np.uint32), c[order], v[order]


def _slot_offsets(off, wid):
    off = np.asarray(off, np.int64)
    wid = np.asarray(wid, np.int64)
    return np.concatenate([[0], np.cumsum(np.diff(off) * wid)])[:-1] if len(wid) else np.zeros(0, np.int64)


def canonical_sliced_of_bucketed_sell(part, vb):
    """rebuild_canonical_sliced_partition (src/io/csh5/execution_runtime_part
    .hh:275-318): the canonical Sliced-ELL in the partition's canonical slice
    layout; exec row x's valid slots (column != invalid), in slot order, at
    the front of canonical row exec_to_canonical[x], invalid / 0 after.  The
    exec row is found in its segment's own slices (the reference reads the
    first slice's width, the same for the one-slice segments its builder
    emits).  Returns (slot offsets, col_idx, val bits); raises when a row's
    valid slots exceed its canonical width, as the reference fails."""
    off = np.asarray(part["canonical_slice_row_offsets"], np.int64)
    wid = np.asarray(part["canonical_slice_widths"], np.int64)
    slots = _slot_offsets(off, wid)
    total = int(slots[-1] + (off[-1] - off[-2]) * wid[-1]) if len(wid) else 0
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[vb]
    col = np.full(total, INVALID, np.uint32)
    val = np.zeros(total, dt)
    sro = np.asarray(part["segment_row_offsets"], np.int64)
    e2c = np.asarray(part["exec_to_canonical"], np.int64)
    for x in range(part["rows"]):
        c = int(e2c[x])
        ds = int(np.searchsorted(off[:-1], c, side="right") - 1)
        dbase = int(slots[ds] + (c - off[ds]) * wid[ds])
        k = next(k for k in range(len(part["segments"])) if sro[k] <= x < sro[k + 1])
        s = part["segments"][k]
        local = x - int(sro[k])
        so = np.asarray(s["slice_row_offsets"], np.int64)
        sw = np.asarray(s["slice_widths"], np.int64)
        q = int(np.searchsorted(so[:-1], local, side="right") - 1)
        sbase = int(_slot_offsets(so, sw)[q] + (local - so[q]) * sw[q])
        sc = np.asarray(s["col_idx"], np.uint32)[sbase:sbase + int(sw[q])]
        sv = np.asarray(s["val"])[sbase:sbase + int(sw[q])]
        keep = sc != INVALID
        n = int(keep.sum())
        if n > wid[ds]:
            raise ValueError(f"row {c}: {n} valid slots exceed width {wid[ds]}")
        col[dbase:dbase + n] = sc[keep]
        val[dbase:dbase + n] = sv[keep]
    return slots.astype(np.uint32), col, val
