"""Shared test data builders (numpy only)."""
from __future__ import annotations

import numpy as np

F16, F32, BF16, FP8 = 3, 4, 6, 7


def f16_bits(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32).astype(np.float16).view(np.int16)


def bf16_bits(x) -> np.ndarray:
    b = np.asarray(x, dtype=np.float32).view(np.uint32)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
    return r.view(np.int16)


def decode(bits, vt) -> np.ndarray:
    bits = np.asarray(bits)
    if vt == F16:
        return bits.view(np.float16).astype(np.float64)
    if vt == BF16:
        return (bits.view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    if vt == F32:
        return bits.view(np.float32).astype(np.float64)
    if vt == FP8:
        b = bits.view(np.uint8).astype(np.int64)
        s = np.where(b >> 7, -1.0, 1.0)
        e = (b >> 3) & 0xF
        m = b & 7
        out = np.where(e == 0, m * 2.0 ** -9, (1 + m / 8.0) * 2.0 ** (e - 7))
        out = np.where((e == 15) & (m == 7), np.nan, out)
        return s * out
    raise ValueError(vt)


def encode(values, vt) -> np.ndarray:
    if vt == F16:
        return f16_bits(values)
    if vt == BF16:
        return bf16_bits(values)
    if vt == F32:
        return np.asarray(values, dtype=np.float32)
    if vt == FP8:
        # exact small integers / dyadic values only (tests pick representable values)
        table = {float(decode(np.array([i], np.uint8), FP8)[0]): i for i in range(256) if (i & 0x7F) != 0x7F}
        return np.array([table[float(v)] for v in np.asarray(values, dtype=np.float64)], dtype=np.uint8)
    raise ValueError(vt)


def random_csr(rows, cols, density=0.05, seed=0, max_count=20, sorted_rows=True, empty_rows=0.0, zeros=0.0,
               duplicates=0.0, fp8_safe=False):
    """Random CSR with integer counts.  Returns (indptr u32, indices u32, values f64)."""
    rng = np.random.default_rng(seed)
    indptr = [0]
    indices = []
    values = []
    for r in range(rows):
        if rng.random() < empty_rows:
            k = 0
        else:
            k = int(rng.binomial(cols, density))
        cs = rng.choice(cols, size=k, replace=False) if k else np.zeros(0, np.int64)
        if sorted_rows:
            cs = np.sort(cs)
        if duplicates and k:
            nd = int(rng.binomial(k, duplicates))
            if nd:
                dup = rng.choice(cs, size=nd)
                cs = np.concatenate([cs, dup])
                if sorted_rows:
                    cs = np.sort(cs, kind="stable")
        if fp8_safe:
            v = rng.choice([1.0, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 16.0], size=len(cs))
        else:
            v = rng.integers(1, max_count + 1, size=len(cs)).astype(np.float64)
        if zeros and len(cs):
            v[rng.random(len(cs)) < zeros] = 0.0
        indices.append(cs)
        values.append(v)
        indptr.append(indptr[-1] + len(cs))
    idx = np.concatenate(indices).astype(np.uint32) if indices else np.zeros(0, np.uint32)
    val = np.concatenate(values) if values else np.zeros(0)
    return np.asarray(indptr, np.uint32), idx, val


def dense_of(rows, cols, indptr, indices, values):
    d = np.zeros((rows, cols))
    for r in range(rows):
        for i in range(indptr[r], indptr[r + 1]):
            d[r, indices[i]] = values[i]  # last duplicate wins
    return d


BITS = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def _cuts(rng, rows, nseg):
    if nseg == 0:
        return np.zeros(1, np.uint32)
    inner = np.sort(rng.integers(0, rows + 1, nseg - 1)) if nseg > 1 else np.zeros(0, np.int64)
    return np.concatenate([[0], inner, [rows]]).astype(np.uint32)


def random_bell_partition(rng, rows, cols, nseg, vb, block_sizes=(1, 2, 4, 8, 16)):
    """A bucketed Blocked-ELL execution partition as a dict of numpy arrays
    (keys of BucketedBlockedEll.to_host()), with random structure: segment
    cuts, per-segment block sizes and widths, invalid block slots."""
    sro = _cuts(rng, rows, nseg if rows else nseg)
    if rows == 0:
        sro = np.zeros(nseg + 1, np.uint32)
    segs = []
    for k in range(nseg):
        r = int(sro[k + 1] - sro[k])
        bs = int(rng.choice(block_sizes))
        w = int(rng.integers(0, 4)) if r else int(rng.integers(0, 2))
        rb = (r + bs - 1) // bs
        bci = rng.integers(0, max(cols // bs, 1), rb * w).astype(np.uint32)
        bci[rng.random(rb * w) < 0.2] = 0xFFFFFFFF
        val = rng.integers(0, 1 << (8 * vb), r * w * bs, dtype=np.uint64).astype(BITS[vb])
        segs.append(dict(rows=r, cols=cols, nnz=int(rng.integers(0, r * w * bs + 1)), block_size=bs, ell_cols=w * bs,
                         block_col_idx=bci, val=val))
    perm = rng.permutation(rows).astype(np.uint32)
    inv = np.empty(rows, np.uint32)
    inv[perm] = np.arange(rows, dtype=np.uint32)
    cperm = rng.permutation(cols).astype(np.uint32)
    cinv = np.empty(cols, np.uint32)
    cinv[cperm] = np.arange(cols, dtype=np.uint32)
    return dict(rows=rows, cols=cols, nnz=sum(s["nnz"] for s in segs), segments=segs, segment_row_offsets=sro,
                exec_to_canonical=perm, canonical_to_exec=inv, col_maps=(cperm, cinv))


def random_sell_partition(rng, rows, cols, nseg, vb):
    """A bucketed Sliced-ELL execution partition dict with several slices per
    segment (uniform heights with a short last slice, or ragged heights)."""
    sro = _cuts(rng, rows, nseg) if rows else np.zeros(nseg + 1, np.uint32)
    segs = []
    for k in range(nseg):
        r = int(sro[k + 1] - sro[k])
        if r == 0:
            offs = np.zeros(1 + int(rng.integers(0, 2)), np.uint32)
        elif rng.random() < 0.5:
            h = int(rng.integers(1, 9))
            offs = np.unique(np.concatenate([np.arange(0, r, h), [r]])).astype(np.uint32)
        else:
            offs = np.unique(np.concatenate([[0], rng.integers(0, r, int(rng.integers(0, 5))), [r]])).astype(np.uint32)
        n = len(offs) - 1
        wid = rng.integers(0, 6, n).astype(np.uint32)
        slots = int(((offs[1:].astype(np.int64) - offs[:-1]) * wid).sum())
        col = rng.integers(0, cols, slots).astype(np.uint32)
        col[rng.random(slots) < 0.2] = 0xFFFFFFFF
        val = rng.integers(0, 1 << (8 * vb), slots, dtype=np.uint64).astype(BITS[vb])
        segs.append(dict(rows=r, cols=cols, nnz=int(rng.integers(0, slots + 1)), slice_count=n, slice_row_offsets=offs,
                         slice_widths=wid, col_idx=col, val=val))
    perm = rng.permutation(rows).astype(np.uint32)
    inv = np.empty(rows, np.uint32)
    inv[perm] = np.arange(rows, dtype=np.uint32)
    ncs = int(rng.integers(0, 5))
    cso = np.sort(rng.integers(0, rows + 1, ncs + 1)).astype(np.uint32)
    return dict(rows=rows, cols=cols, nnz=sum(s["nnz"] for s in segs), segments=segs, segment_row_offsets=sro,
                exec_to_canonical=perm, canonical_to_exec=inv, canonical_slice_row_offsets=cso,
                canonical_slice_widths=rng.integers(0, 9, ncs).astype(np.uint32))
