"""Pin the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: the reference's fixtures, the reference's compiled CPU converter
(golden vectors), dense numpy recomputation, scipy, or closed forms.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import BF16, F16, F32, FP8, decode, dense_of, encode, f16_bits, random_csr
from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def mg():
    return json.loads((GOLD / "mask_groups_fixture.json").read_text())


def _check_asserted(got, expected, tol):
    for k, v in expected.items():
        assert abs(float(got[int(k)]) - float(v)) <= tol, (k, got[int(k)], v)


def test_decoders_match_numpy():
    h = np.arange(0, 65536, 97, dtype=np.uint32).astype(np.uint16)
    for b in h:
        ref = np.array([b], np.uint16).view(np.float16).astype(np.float32)[0]
        got = O.lib().oracle_half_to_float(int(b))
        assert (np.isnan(ref) and np.isnan(got)) or ref == got
    # e4m3fn: closed form (1 + m/8) * 2^(e-7), subnormal m * 2^-9, 0x7f/0xff NaN
    for b in range(256):
        got = O.lib().oracle_fp8_e4m3_to_float(b)
        ref = decode(np.array([b], np.uint8), FP8)[0]
        assert (np.isnan(ref) and np.isnan(got)) or ref == got
    assert O.lib().oracle_fp8_e4m3_to_float(0x7E) == 448.0


def test_group_reduce_reference_fixture_blocked(mg):
    b = mg["blocked_ell"]
    e = mg["expected"]["blocked_ell_masked"]
    tol = mg["expected"]["tolerance_abs"]
    o = O.group_reduce_bell(mg["rows"], mg["cols"], b["block_size"], b["ell_cols"], b["block_col_idx"], f16_bits(b["values"]), F16,
                            mg["row_keep"], mg["feature_keep"], mg["group_masks"], mg["group_count"], mg["filter"])
    _check_asserted(o["row_totals"], e["row_totals"], tol)
    _check_asserted(o["detected"], e["detected"], 0)
    _check_asserted(o["keep"], e["keep"], 0)
    _check_asserted(o["group_counts"].ravel(), e["group_counts"], tol)
    _check_asserted(o["group_pct"].ravel(), e["group_pct"], tol)
    u = mg["expected"]["blocked_ell_unmasked_no_groups"]
    o2 = O.group_reduce_bell(mg["rows"], mg["cols"], b["block_size"], b["ell_cols"], b["block_col_idx"], f16_bits(b["values"]), F16,
                             None, None, None, 0, mg["filter"])
    _check_asserted(o2["row_totals"], u["row_totals"], tol)
    _check_asserted(o2["detected"], u["detected"], 0)


def test_group_reduce_reference_fixture_compressed_and_sliced(mg):
    c = mg["compressed"]
    tol = mg["expected"]["tolerance_abs"]
    o = O.group_reduce_csr(mg["rows"], mg["cols"], c["ptr"], c["idx"], f16_bits(c["values"]), F16, mg["row_keep"],
                           mg["feature_keep"], mg["group_masks"], mg["group_count"], mg["filter"])
    _check_asserted(o["row_totals"], mg["expected"]["compressed_masked"]["row_totals"], tol)
    _check_asserted(o["group_counts"].ravel(), mg["expected"]["compressed_masked"]["group_counts"], tol)
    s = mg["sliced_ell"]
    o = O.group_reduce_sell(mg["rows"], mg["cols"], s["slice_rows"], s["slice_row_offsets"], s["slice_widths"],
                            s["slice_slot_offsets"], s["col_idx"], f16_bits(s["values"]), F16, mg["row_keep"], mg["feature_keep"],
                            mg["group_masks"], mg["group_count"], mg["filter"])
    _check_asserted(o["row_totals"], mg["expected"]["sliced_ell_masked"]["row_totals"], tol)


@pytest.mark.parametrize("vt", [F16, F32, BF16, FP8])
def test_group_reduce_matches_dense_numpy(vt):
    rows, cols, G = 60, 45, 7
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=11, zeros=0.05, fp8_safe=(vt == FP8))
    rng = np.random.default_rng(3)
    rk = (rng.random(rows) < 0.8).astype(np.uint8)
    fk = (rng.random(cols) < 0.7).astype(np.uint8)
    fgm = rng.integers(0, 1 << G, size=cols).astype(np.uint32)
    bits = encode(val, vt)
    o = O.group_reduce_csr(rows, cols, ptr, idx, bits, vt, rk, fk, fgm, G, [5.0, 2, 0.5, 1])
    d = dense_of(rows, cols, ptr, idx, decode(bits, vt)) * fk[None, :] * rk[:, None]
    np.testing.assert_allclose(o["row_totals"], d.sum(1), rtol=1e-12)
    assert np.array_equal(o["detected"], (d != 0).sum(1))
    for g in range(G):
        cols_g = ((fgm >> g) & 1).astype(bool)
        np.testing.assert_allclose(o["group_counts"][:, g], d[:, cols_g].sum(1), rtol=1e-12)
    np.testing.assert_allclose(o["max_values"], np.maximum(d.max(1), 0), rtol=0)
    s = d.sum(1).astype(np.float32)
    frac = np.where(s > 0, d[:, ((fgm >> 1) & 1).astype(bool)].sum(1).astype(np.float32) / np.where(s > 0, s, 1), 0)
    keep = (s >= 5.0) & ((d != 0).sum(1) >= 2) & (frac <= 0.5) & (rk == 1)
    assert np.array_equal(o["keep"].astype(bool), keep)


@pytest.mark.parametrize("vt", [F16, BF16, FP8, F32])
def test_label_reduce_matches_np_add_at(vt):
    rows, cols, G = 80, 50, 5
    ptr, idx, val = random_csr(rows, cols, 0.25, seed=21, zeros=0.05, fp8_safe=(vt == FP8))
    rng = np.random.default_rng(4)
    labels = rng.integers(0, G + 1, size=rows).astype(np.uint32)  # label G drops the row
    rk = (rng.random(rows) < 0.85).astype(np.uint8)
    fk = (rng.random(cols) < 0.6).astype(np.uint8)
    scale = rng.uniform(0.5, 2.0, cols).astype(np.float32) if vt == FP8 else None
    bits = encode(val, vt)
    o = O.label_reduce_csr(rows, cols, ptr, idx, bits, vt, labels, rk, fk, G, scale)
    kept = np.flatnonzero(fk)
    v = decode(bits, vt).astype(np.float32).astype(np.float64)
    if scale is not None:
        v = v * scale[idx].astype(np.float64)
    rows_of = np.repeat(np.arange(rows), np.diff(ptr))
    ok = (rk[rows_of] == 1) & (labels[rows_of] < G) & (fk[idx] == 1) & (v != 0)
    rank = np.full(cols, -1)
    rank[kept] = np.arange(len(kept))
    S = np.zeros((G, len(kept)))
    Q = np.zeros((G, len(kept)))
    N = np.zeros((G, len(kept)), np.int64)
    np.add.at(S, (labels[rows_of][ok], rank[idx][ok]), v[ok])
    np.add.at(Q, (labels[rows_of][ok], rank[idx][ok]), v[ok] ** 2)
    np.add.at(N, (labels[rows_of][ok], rank[idx][ok]), 1)
    np.testing.assert_allclose(o["sum"], S, rtol=1e-12)
    np.testing.assert_allclose(o["sumsq"], Q, rtol=1e-12)
    assert np.array_equal(o["nnz"], N)
    assert np.array_equal(o["kept_feature_ids"], kept)
    assert np.array_equal(o["group_rows"], np.bincount(labels[(rk == 1) & (labels < G)], minlength=G + 1)[:G])


def test_label_reduce_layouts_agree():
    rows, cols, G = 50, 40, 3
    ptr, idx, val = random_csr(rows, cols, 0.3, seed=5)
    bits = f16_bits(val)
    labels = (np.arange(rows) % G).astype(np.uint32)
    fk = (np.arange(cols) % 3 != 0).astype(np.uint8)
    a = O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, None, fk, G)
    s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, 8)
    b = O.label_reduce_sell(rows, cols, s["slice_rows"], s["offsets"], s["widths"], s["slot_offsets"], s["col_idx"], s["val"], F16,
                            labels, None, fk, G)
    bci, bv, ell, _ = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, 4)
    c = O.label_reduce_bell(rows, cols, 4, ell, bci, bv, F16, labels, None, fk, G)
    for k in ("sum", "sumsq", "nnz"):
        assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], c[k]), k


@pytest.fixture(scope="module")
def bell_gold():
    return np.load(GOLD / "bell_ref_vectors.npz")


def _cases(g):
    return sorted({k.split("/")[0] for k in g.files})


def test_blocked_ell_matches_reference_converter(bell_gold):
    g = bell_gold
    for name in _cases(g):
        rows, cols = (int(x) for x in g[f"{name}/meta"])
        ptr, idx, val = g[f"{name}/ptr"], g[f"{name}/idx"], g[f"{name}/val"]
        for key in [k for k in g.files if k.startswith(f"{name}/b") and k.endswith("/bci")]:
            bs = int(key.split("/")[1][1:])
            bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, val, bs)
            ref_ell, ref_nnz = (int(x) for x in g[f"{name}/b{bs}/shape"])
            assert ell == ref_ell and nnz == ref_nnz, (name, bs)
            assert np.array_equal(bci, g[f"{name}/b{bs}/bci"]), (name, bs)
            assert np.array_equal(v.view(np.int16), g[f"{name}/b{bs}/val"]), (name, bs)


def test_blocked_ell_tuner_matches_reference(bell_gold):
    g = bell_gold
    for name in _cases(g):
        rows, cols = (int(x) for x in g[f"{name}/meta"])
        for k in range(3):
            if f"{name}/choose{k}/cands" not in g.files:
                continue
            cands = g[f"{name}/choose{k}/cands"]
            bs, fill, pad = O.choose_blocked_ell_block_size(rows, cols, g[f"{name}/ptr"], g[f"{name}/idx"], cands, 2)
            rbs, rpad = (int(x) for x in g[f"{name}/choose{k}/result"])
            assert (bs, pad) == (rbs, rpad), (name, k)
            assert abs(fill - float(g[f"{name}/choose{k}/fill"][0])) < 1e-12


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU converter not built (no reference tree)")
def test_blocked_ell_matches_reference_converter_live():
    for seed in range(6):
        rows, cols = 20 + seed * 7, 30 + seed * 11
        ptr, idx, val = random_csr(rows, cols, 0.15, seed=100 + seed, sorted_rows=seed % 2 == 0, zeros=0.05,
                                   duplicates=0.05 * (seed % 2))
        bits = f16_bits(val)
        for bs in (1, 2, 5, 8, 16):
            a = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
            r = O.ref_blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
            assert a[2:] == r[2:]
            assert np.array_equal(a[0], r[0]) and np.array_equal(a[1].view(np.uint16), r[1])


def test_blocked_ell_dense_round_trip():
    rows, cols, bs = 30, 40, 4
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=8, zeros=0.1)
    bits = f16_bits(val)
    bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    w = ell // bs
    d = np.zeros((rows, cols))
    vals = v.view(np.float16).astype(np.float64).reshape(rows, ell)
    for r in range(rows):
        for s in range(w):
            b = bci[(r // bs) * w + s]
            if b == O.INVALID:
                continue
            for lane in range(bs):
                if b * bs + lane < cols:
                    d[r, b * bs + lane] = vals[r, s * bs + lane]
    ref = dense_of(rows, cols, ptr, idx, val)
    assert np.array_equal(d, ref)
    assert nnz == int((ref != 0).sum())


def test_sliced_ell_layout_closed_form_and_round_trip():
    rows, cols = 45, 30
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=9, empty_rows=0.2)
    bits = f16_bits(val)
    rn = np.diff(ptr)
    for S in (1, 4, 7, 32, 100):
        s = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, S)
        Se = min(S, rows)
        starts = np.arange(0, rows, Se)
        assert np.array_equal(s["offsets"], np.append(starts, rows))
        widths = np.array([rn[a:a + Se].max() for a in starts])
        assert np.array_equal(s["widths"], widths)
        heights = np.diff(np.append(starts, rows))
        assert np.array_equal(s["slot_offsets"], np.concatenate([[0], np.cumsum(heights * widths)[:-1]]))
        d = np.zeros((rows, cols))
        vv = s["val"].view(np.float16).astype(np.float64)
        for k, a in enumerate(starts):
            for r in range(a, min(a + Se, rows)):
                base = s["slot_offsets"][k] + (r - a) * widths[k]
                row_cols = s["col_idx"][base:base + widths[k]]
                n = rn[r]
                assert np.array_equal(row_cols[:n], idx[ptr[r]:ptr[r + 1]])
                assert np.all(row_cols[n:] == O.INVALID) and np.all(vv[base + n:base + widths[k]] == 0)
                d[r, row_cols[:n]] = vv[base:base + n]
        assert np.array_equal(d, dense_of(rows, cols, ptr, idx, val))


def test_sliced_ell_from_csc_equals_from_csr():
    rows, cols = 33, 28
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=10)
    bits = f16_bits(val)
    m = sp.csr_matrix((np.arange(1, len(idx) + 1), idx, ptr), shape=(rows, cols)).tocsc()
    cbits = bits[m.data - 1]
    a = O.sliced_ell_from_compressed(rows, cols, 0, ptr, idx, bits, 8)
    b = O.sliced_ell_from_compressed(rows, cols, 1, m.indptr, m.indices, cbits, 8)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_sliced_ell_from_blocked_ell_compacts_nonzeros():
    rows, cols, bs = 26, 31, 4
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=12, zeros=0.1)
    bits = f16_bits(val)
    bci, bv, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    s = O.sliced_ell_from_blocked_ell(rows, cols, bs, ell, bci, bv, 8)
    # equals the SELL of the CSR with explicit zeros removed
    keep = val != 0
    rows_of = np.repeat(np.arange(rows), np.diff(ptr))
    ptr2 = np.concatenate([[0], np.cumsum(np.bincount(rows_of[keep], minlength=rows))]).astype(np.uint32)
    t = O.sliced_ell_from_compressed(rows, cols, 0, ptr2, idx[keep], bits[keep], 8)
    for k in t:
        assert np.array_equal(np.asarray(s[k]), np.asarray(t[k])), k


def test_sliced_ell_tuner_closed_form():
    rn = np.array([5, 0, 3, 9, 1, 1, 2, 8, 4, 4], np.uint32)
    out = O.choose_sliced_ell_slice_rows(rn, len(rn), int(rn.sum()), [2, 4, 8])

    def score(S):
        pad = sum(int(rn[a:a + S].max()) * len(rn[a:a + S]) for a in range(0, len(rn), S))
        cnt = -(-len(rn) // S)
        return pad / rn.sum() + 0.25 * cnt / len(rn), pad
    best = min((score(S)[0], S) for S in (2, 4, 8))
    assert out["slice_rows"] == best[1]
    assert abs(out["score"] - best[0]) < 1e-12


def test_sliced_ell_tuner_ties():
    # S = 8 and S = 4 score the same (136/128 + 0.25 * 1/8 == 132/128 + 0.25 * 2/8):
    # the one with fewer padded bytes (S = 4: 812 < 828) wins, in either order
    rn = np.array([16, 16, 16, 16, 17, 17, 15, 15], np.uint32)
    for cands in ([8, 4], [4, 8]):
        out = O.choose_sliced_ell_slice_rows(rn, len(rn), int(rn.sum()), cands)
        assert out["slice_rows"] == 4 and out["padded_slots"] == 132
    # S = 8 and S = 2 tie on score (1.15625) and on padded bytes (892): the
    # larger slice height wins, in either order
    rn = np.array([7, 7, 7, 7, 7, 7, 1, 1] + [6] * 7 + [2] + [5] * 8, np.uint32)
    assert rn.sum() == 128
    for cands in ([2, 8], [8, 2]):
        out = O.choose_sliced_ell_slice_rows(rn, len(rn), 128, cands)
        assert out["slice_rows"] == 8 and out["padded_slots"] == 144


def test_label_reduce_chunked_equals_one_pass():
    rows, cols, G = 300, 90, 7
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=33, zeros=0.05, empty_rows=0.1)
    rng = np.random.default_rng(5)
    labels = rng.integers(0, G + 1, size=rows).astype(np.uint32)
    rk = (rng.random(rows) < 0.8).astype(np.uint8)
    fk = (rng.random(cols) < 0.5).astype(np.uint8)
    bits = f16_bits(val)
    a = O.label_reduce_csr(rows, cols, ptr, idx, bits, F16, labels, rk, fk, G)
    for threads in (1, 3):
        b = O.label_reduce_csr_chunked(rows, cols, ptr, idx, bits, F16, labels, rk, fk, G, threads=threads)
        for k in ("sum", "sumsq", "nnz", "group_rows", "kept_feature_ids"):
            assert np.array_equal(a[k], b[k]), k  # integer data: FP64 sums are exact in any order


def test_transpose_matches_scipy():
    for seed, (rows, cols) in enumerate([(40, 30), (1, 7), (25, 1), (60, 200)]):
        ptr, idx, val = random_csr(rows, cols, 0.2, seed=seed, empty_rows=0.1)
        bits = f16_bits(val)
        optr, oidx, oval = O.transpose(rows, cols, ptr, idx, bits)
        m = sp.csr_matrix((np.arange(1, len(idx) + 1), idx, ptr), shape=(rows, cols)).tocsc()
        m.sort_indices()
        assert np.array_equal(optr, m.indptr) and np.array_equal(oidx, m.indices)
        assert np.array_equal(oval, bits[m.data - 1])


def test_bucket_plan_matches_stable_argsort():
    rng = np.random.default_rng(7)
    for n, req in [(1000, 12), (17, 5), (5, 9), (0, 3), (1, 1)]:
        rn = rng.integers(0, 40, size=n).astype(np.uint32)
        ptr = np.concatenate([[0], np.cumsum(rn)]).astype(np.uint32)
        p = O.bucket_plan(ptr, n, req)
        order = np.argsort(rn, kind="stable")
        assert np.array_equal(p["major_order"], order) and np.array_equal(p["major_nnz_sorted"], rn[order])
        b = 1 if n == 0 else min(max(req, 1), n)
        assert p["bucket_count"] == b
        if n:
            sizes = np.diff(p["bucket_offsets"].astype(np.int64))
            assert sizes.sum() == n and sizes.max() - sizes.min() <= 1 and np.all(np.diff(sizes) <= 0)


def test_log2_buckets_closed_form():
    vals = np.array([0, 1, 31, 32, 33, 63, 64, 1000, 32767, 32768, 40000, 1 << 20], np.uint32)
    expect = [0, 0, 0, 1, 1, 1, 2, 5, 10, 11, 11, 11]
    assert [O.log2_bucket(v) for v in vals] == expect
    offs = O.log2_offsets(np.sort(vals))
    assert np.array_equal(np.diff(offs), np.bincount(expect, minlength=12))


def test_rebuild_invariants():
    ptr, idx, val = random_csr(30, 20, 0.3, seed=13, empty_rows=0.2)
    bits = f16_bits(val)
    order = np.argsort(np.diff(ptr), kind="stable").astype(np.uint32)
    dptr, didx, dval, inv = O.rebuild(ptr, idx, bits, order)
    assert np.array_equal(inv[order], np.arange(30))
    for d, s in enumerate(order):
        assert np.array_equal(didx[dptr[d]:dptr[d + 1]], idx[ptr[s]:ptr[s + 1]])
        assert np.array_equal(dval[dptr[d]:dptr[d + 1]], bits[ptr[s]:ptr[s + 1]])


def test_bell_bucket_keys_closed_form():
    rng = np.random.default_rng(3)
    rb, w = 9, 6
    bci = rng.integers(0, 50, size=rb * w).astype(np.uint32)
    bci[rng.random(rb * w) < 0.3] = O.INVALID
    keys = O.bell_bucket_keys(rb, w, bci)
    for r in range(rb):
        h, valid = 2166136261, 0
        for c in bci[r * w:(r + 1) * w]:
            if c == O.INVALID:
                continue
            valid += 1
            h = (h ^ ((int(c) + 0x9E3779B9 + (h << 6) + (h >> 2)) & 0xFFFFFFFF)) & 0xFFFFFFFF
            h = (h * 16777619) & 0xFFFFFFFF
        assert int(keys[r]) == (valid << 32) | h


def test_reoptimize_reference_fixture(mg):
    b = mg["blocked_ell"]
    r = mg["reoptimize"]
    cfg = r["config"]
    mp, mi, mv, nnz = O.reoptimize(mg["rows"], mg["cols"], b["block_size"], b["ell_cols"], b["block_col_idx"],
                                   f16_bits(b["values"]).view(np.uint16), mg["row_keep"], mg["feature_keep"], r["row_remap"],
                                   r["feature_remap"], cfg["output_rows"], cfg["output_cols"])
    assert nnz == r["expected"]["live_nnz"]


def test_reoptimize_export_fixture():
    e = json.loads((GOLD / "export_bell_fixture.json").read_text())
    mp, mi, mv, nnz = O.reoptimize(e["rows"], e["cols"], e["block_size"], e["ell_cols"], e["block_col_idx"],
                                   f16_bits(e["values"]).view(np.uint16), None, None, None, None, e["rows"], e["cols"])
    assert list(mp) == e["expected_csr"]["indptr"]
    assert list(mi) == e["expected_csr"]["indices"]
    assert list(mv.view(np.float16).astype(float)) == e["expected_csr"]["data"]


def test_filter_blocked_ell_reference_fixtures(mg):
    # the reference test's masked reoptimize case (same keep / remap rules)
    # and its Blocked-ELL -> CSR export case (all kept, identity maps)
    b = mg["blocked_ell"]
    r = mg["reoptimize"]
    cfg = r["config"]
    vals = f16_bits(b["values"]).view(np.uint16)
    fp, fi, fv, st = O.filter_blocked_ell(mg["rows"], mg["cols"], b["block_size"], b["ell_cols"], b["block_col_idx"], vals, O.VT_F16,
                                          cfg["output_rows"], cfg["output_cols"], mg["row_keep"], mg["feature_keep"], r["row_remap"],
                                          r["feature_remap"])
    assert st["live_nnz"] == r["expected"]["live_nnz"] and st["kept_rows"] == cfg["output_rows"]
    mp, mi, mv, nnz = O.reoptimize(mg["rows"], mg["cols"], b["block_size"], b["ell_cols"], b["block_col_idx"], vals, mg["row_keep"],
                                   mg["feature_keep"], r["row_remap"], r["feature_remap"], cfg["output_rows"], cfg["output_cols"])
    assert np.array_equal(fp, mp) and np.array_equal(fi, mi) and np.array_equal(fv, mv)
    e = json.loads((GOLD / "export_bell_fixture.json").read_text())
    fp, fi, fv, st = O.filter_blocked_ell(e["rows"], e["cols"], e["block_size"], e["ell_cols"], e["block_col_idx"],
                                          f16_bits(e["values"]).view(np.uint16), O.VT_F16, e["rows"], e["cols"])
    assert list(fp) == e["expected_csr"]["indptr"] and list(fi) == e["expected_csr"]["indices"]
    assert list(fv.view(np.float16).astype(float)) == e["expected_csr"]["data"]


@pytest.mark.parametrize("vt", [F16, F32, FP8])
def test_filter_blocked_ell_identity_is_csr_without_zeros(vt):
    # keep everything, identity maps: Blocked-ELL -> CSR returns the source
    # CSR minus its explicit zeros (any block size)
    rows, cols = 61, 83
    ptr, idx, val = random_csr(rows, cols, 0.1, seed=9, zeros=0.2, fp8_safe=(vt == FP8))
    bits = encode(val, vt)
    keep = val != 0
    for bs in (1, 4, 8):
        bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
        fp, fi, fv, st = O.filter_blocked_ell(rows, cols, bs, ell, bci, v, vt, rows, cols)
        cnt = np.array([keep[ptr[r]:ptr[r + 1]].sum() for r in range(rows)])
        assert np.array_equal(fp[1:], np.cumsum(cnt)) and np.array_equal(fi, np.asarray(idx)[keep])
        assert np.array_equal(np.ascontiguousarray(fv).view(np.uint8), np.ascontiguousarray(bits[keep]).view(np.uint8))
        assert st["live_nnz"] == int(keep.sum()) == nnz


def test_filter_blocked_ell_matches_reoptimize_restatement():
    # two restatements of one rule (numpy here, C in oracle.c) on random
    # masks and compacting remaps
    rng = np.random.default_rng(4)
    rows, cols, bs = 97, 150, 4
    ptr, idx, val = random_csr(rows, cols, 0.08, seed=12, zeros=0.1)
    bits = f16_bits(val)
    bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
    rk = (rng.random(rows) < 0.7).astype(np.uint8)
    fk = (rng.random(cols) < 0.5).astype(np.uint8)
    rr = np.full(rows, 0xFFFFFFFF, np.uint32)
    rr[rk != 0] = np.arange(int(rk.sum()))
    cr = np.full(cols, 0xFFFFFFFF, np.uint32)
    cr[fk != 0] = rng.permutation(int(fk.sum()))
    out_rows, out_cols = int(rk.sum()) - 3, int(fk.sum()) - 2  # some kept rows / columns map past the end
    fp, fi, fv, st = O.filter_blocked_ell(rows, cols, bs, ell, bci, v, O.VT_F16, out_rows, out_cols, rk, fk, rr, cr)
    mp, mi, mv, n2 = O.reoptimize(rows, cols, bs, ell, bci, v, rk, fk, rr, cr, out_rows, out_cols)
    assert np.array_equal(fp, mp) and np.array_equal(fi, mi) and np.array_equal(fv, mv) and st["live_nnz"] == n2


def _bell_dense(rows, cols, bs, ell_cols, bci, val):
    d = np.zeros((rows, cols), np.float64)
    W = ell_cols // bs if bs else 0
    v = np.asarray(val).view(np.float16).astype(np.float64).reshape(rows, ell_cols) if rows * ell_cols else None
    for r in range(rows):
        for s_ in range(W):
            c0 = int(bci[(r // bs) * W + s_])
            if c0 == O.INVALID:
                continue
            for l in range(bs):
                if c0 * bs + l < cols:
                    d[r, c0 * bs + l] = v[r, s_ * bs + l]
    return d


def test_repack_is_a_row_permutation_of_the_kept_matrix():
    # two parts of one shard, all rows and columns kept, identity maps: the
    # repacked Blocked-ELL's row k is the shard's row order[k], order = the
    # stable ascending-nnz bucket order; kept rows / slots by closed form
    parts, dense = [], []
    for k, (rows, bs) in enumerate(((37, 4), (29, 8))):
        ptr, idx, val = random_csr(rows, 90, 0.1, seed=50 + k, zeros=0.0)
        bits = f16_bits(val)
        bci, v, ell, nnz = O.blocked_ell_from_compressed(rows, 90, ptr, idx, bits, bs)
        parts.append({"rows": rows, "cols": 90, "block_size": bs, "ell_cols": ell, "bci": bci, "val": v})
        dense.append(_bell_dense(rows, 90, bs, ell, bci, v))
    D = np.vstack(dense)
    out = O.repack_blocked_ell(parts, O.VT_F16, 66, 90, 16, 5)
    order = out["plan"]["major_order"]
    R = _bell_dense(66, 90, 16, out["bell_ell_cols"], out["bell_bci"], out["bell_val"])
    assert np.array_equal(R, D[order])
    assert out["stats"]["kept_rows"] == 66 and out["stats"]["kept_row_slots"] == sum(p["rows"] * p["ell_cols"] for p in parts)
    assert np.array_equal(np.diff(out["ptr"].astype(np.int64))[order], np.sort(np.diff(out["ptr"].astype(np.int64)), kind="stable"))


def _fnv1a64(seed, data: bytes):
    for b in data:
        seed = ((seed ^ b) * 1099511628211) & ((1 << 64) - 1)
    return seed


def test_column_signature_mix_is_fnv1a():
    # mix_signature is FNV-1a 64 over the value's 8 little-endian bytes: the
    # published test vector for "a", then the 8-byte form
    assert _fnv1a64(0xCBF29CE484222325, b"a") == 0xAF63DC4C8601EC8C
    for seed, v in ((1469598103934665603, 1), (1099511628211, 1315423911 * 7), (12345, (1 << 64) - 1)):
        assert O._mix_signature(seed, v) == _fnv1a64(seed, v.to_bytes(8, "little"))


def test_shard_column_maps_known_answer():
    # 2 row blocks of 2 rows, B = 2, 7 columns.  Row block 0 touches columns
    # 0, 1, 4, 5 (column 1 only through row 1), row block 1 columns 2, 3, 4, 5;
    # column 6 never.  Columns with the same support set tie on every key but
    # the column: (0, 1), (2, 3), (4, 5) stay adjacent and in order; the
    # set {0} and the set {0, 1} share the first row block, so their hashes
    # decide; {1} comes after both; the unsupported column is last.
    part = {"rows": 4, "block_size": 2, "ell_cols": 4, "bci": np.array([0, 2, 1, 2], np.uint32),
            "val": np.array([1, 0, 1, 1, 0, 1, 0, 0, 1, 1, 1, 0, 0, 0, 0, 1], np.uint16) * 0x3C00}
    e2c, c2e = O.shard_column_maps([part], O.VT_F16, 7)
    init = 1469598103934665603
    h01 = O._mix_signature(init, 1)
    h45 = O._mix_signature(h01, 2)
    first = [0, 1, 4, 5] if h01 < h45 else [4, 5, 0, 1]
    assert list(e2c) == first + [2, 3, 6]
    assert np.array_equal(c2e[e2c], np.arange(7))


def test_barycenter_order_known_answer_and_float_sums():
    # columns touched by rows of ranks {0, 2} (mean 1), {3} (3), {1} (1),
    # none (FLT_MAX): order by mean, then column
    rows, cols = 4, 4
    ri = np.array([0, 2, 3, 1], np.uint32)
    ci = np.array([0, 0, 1, 2], np.uint32)
    rank = np.array([0, 1, 2, 3], np.uint32)
    e2c, c2e = O.barycenter_order(rows, cols, ri, ci, rank, by_col=True)
    assert list(e2c) == [0, 2, 1, 3]
    # with sums below 2^24 the reference's float accumulation gives the same order
    rng = np.random.default_rng(2)
    ri = rng.integers(0, 500, 4000).astype(np.uint32)
    ci = rng.integers(0, 300, 4000).astype(np.uint32)
    rank = rng.permutation(500).astype(np.uint32)
    rank[rng.random(500) < 0.1] = O.INVALID
    a = O.barycenter_order(500, 300, ri, ci, rank, by_col=True)
    b = O.barycenter_order(500, 300, ri, ci, rank, by_col=True, float_sums=True)
    assert np.array_equal(a[0], b[0])


def test_planted_mistake_is_caught():
    """A dropped term in the oracle (simulated by zeroing one entry) must fail
    the dense comparison used above."""
    rows, cols = 10, 10
    ptr, idx, val = random_csr(rows, cols, 0.4, seed=1)
    bits = f16_bits(val)
    o = O.group_reduce_csr(rows, cols, ptr, idx, bits, F16, None, None, None, 0, [0, 0, 1, 0])
    d = dense_of(rows, cols, ptr, idx, val)
    d[0, idx[ptr[0]]] = 0
    assert not np.allclose(o["row_totals"], d.sum(1))


@pytest.mark.parametrize("op", [0, 1, 2])
def test_serve_query_matches_dense(op):
    """Feature-subset grouped query restated on the CSR vs a dense numpy
    recomputation (explicit labels / mask; labels >= G drop the row)."""
    rows, cols, G = 120, 60, 7
    ptr, idx, val = random_csr(rows, cols, 0.2, seed=31, zeros=0.05)
    bits = encode(val, F16)
    rng = np.random.default_rng(8)
    labels = rng.integers(0, G + 2, size=rows).astype(np.uint32)
    keep = (rng.random(rows) < 0.7).astype(np.uint8)
    feats = np.sort(rng.choice(cols, 17, replace=False)).astype(np.uint32)
    got = O.serve_query(rows, cols, ptr, idx, bits, F16, feats, G, op, labels=labels, keep=keep)
    dense = dense_of(rows, cols, ptr, idx, decode(bits, F16).astype(np.float64))
    ref = np.zeros((len(feats), G))
    for r in range(rows):
        if keep[r] and labels[r] < G:
            x = dense[r, feats]
            ref[:, labels[r]] += (x != 0) if op == 2 else (x * x if op == 1 else x)
    if op == 2:
        assert got.dtype == np.uint32 and np.array_equal(got, ref.astype(np.uint32))
    else:
        np.testing.assert_allclose(got, ref, rtol=1e-12)


def test_serve_seeded_labels_and_mask_are_uniform():
    """Seeded labels cover [0, G) evenly and the mask keeps the requested
    fraction (a wrong shift or range reduction skews both)."""
    n = 40000
    lab = O.serve_label(77, n, 10)
    counts = np.bincount(lab, minlength=10)
    assert lab.max() < 10 and counts.min() > 0.9 * n / 10 and counts.max() < 1.1 * n / 10
    for frac in (0.05, 0.5, 1.0):
        k = O.serve_keep(5, n, int(frac * (1 << 24) + 0.5))
        assert abs(k.mean() - frac) < 0.01
    assert O.serve_keep(5, 100, 0).sum() == 0


def test_bell_execution_layout_known_answer():
    """18 rows, block 4: row blocks of valid widths [3, 1, 2, 1, 3], the last
    one partial (2 rows) and left in place; stable order [1, 3, 2, 0, 4]; two
    equal-count buckets [0, 2) and [2, 5) of widths 1 and 3; packed bytes
    96 + 300 = 396 < 516 for the original, so the split stands."""
    I = O.INVALID
    bci = np.array([0, 1, 2, 5, I, I, 1, 4, I, 2, I, I, 0, 3, 6], np.uint32)
    lay = O.bell_execution_layout(18, 4, 12, bci, 2)
    assert list(lay["order"]) == [1, 3, 2, 0, 4]
    assert lay["begin"] == [0, 2, 5] and lay["width"] == [1, 3] and lay["bytes"] == 396
    assert O.bell_packed_bytes(8, 4, 4) == 96 and O.bell_packed_bytes(10, 12, 4) == 300 and O.bell_packed_bytes(18, 12, 4) == 516
    # equal widths collapse to one segment; a split that does not shrink the bytes falls back
    same = np.array([0, 1, I, 2, 3, I, 4, 5, I, 6, 7, I], np.uint32)
    lay = O.bell_execution_layout(16, 4, 12, same, 3)
    assert lay["begin"] == [0, 4] and lay["width"] == [2]
    assert O.bell_choose_bucket_count(16, 4, 12, same)[0] == 1


def test_sell_min_split_cost_is_exhaustive_minimum():
    """The brute-force DP equals the minimum over every split of the sorted
    widths into exactly k non-empty segments."""
    from itertools import combinations
    rng = np.random.default_rng(3)
    for n, k in [(7, 3), (6, 1), (8, 4), (5, 5)]:
        w = np.sort(rng.integers(0, 40, n))
        best = min(sum(O.sell_packed_bytes(1, (b - a) * int(w[b - 1])) for a, b in zip((0,) + cuts, cuts + (n,)))
                   for cuts in combinations(range(1, n), k - 1))
        assert O.sell_min_split_cost(w, k) == best
