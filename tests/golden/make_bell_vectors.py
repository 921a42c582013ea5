"""Golden Blocked-ELL vectors produced by the reference's own CPU converter.

The converter (src/convert/blocked_ell_from_compressed.cuh, choose_blocked_ell_block_size
:157-190 and blocked_ell_from_compressed :258-347) is compiled from the
reference sources by oracle/build_ref.sh into oracle/_ref/libref_convert.so;
this script runs it on seeded CSR inputs and stores inputs + outputs in
bell_ref_vectors.npz, so the oracle and the HIP path can be checked against
the reference where the reference tree is absent.
Run: python tests/golden/make_bell_vectors.py  (needs oracle/_ref built)
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

from helpers import f16_bits, random_csr  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = [
    # name, rows, cols, density, seed, sorted, empty, zeros, duplicates
    ("sorted_small", 37, 53, 0.08, 1, True, 0.1, 0.0, 0.0),
    ("unsorted_dups_zeros", 41, 67, 0.1, 2, False, 0.05, 0.1, 0.1),
    ("wide", 64, 300, 0.03, 3, True, 0.0, 0.0, 0.0),
    ("ragged_last_block", 33, 17, 0.3, 4, True, 0.2, 0.05, 0.0),
    ("single_row", 1, 9, 0.5, 5, True, 0.0, 0.0, 0.0),
]
BLOCKS = [1, 2, 3, 4, 8, 16, 32]
CANDIDATES = [[4, 8, 16, 32], [8, 16, 32], [2, 3, 5]]


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref_convert.so missing; run oracle/build_ref.sh")
    out = {}
    for name, rows, cols, dens, seed, srt, empty, zeros, dups in CASES:
        ptr, idx, val = random_csr(rows, cols, dens, seed, sorted_rows=srt, empty_rows=empty, zeros=zeros, duplicates=dups)
        bits = f16_bits(val)
        out[f"{name}/meta"] = np.array([rows, cols], np.uint32)
        out[f"{name}/ptr"] = ptr
        out[f"{name}/idx"] = idx
        out[f"{name}/val"] = bits
        for bs in BLOCKS:
            bci, v, ell, nnz = O.ref_blocked_ell_from_compressed(rows, cols, ptr, idx, bits, bs)
            out[f"{name}/b{bs}/bci"] = bci
            out[f"{name}/b{bs}/val"] = v.view(np.int16)
            out[f"{name}/b{bs}/shape"] = np.array([ell, nnz], np.uint32)
        for k, cands in enumerate(CANDIDATES):
            bs, fill, pad = O.ref_choose_blocked_ell_block_size(rows, cols, ptr, idx, bits, cands)
            out[f"{name}/choose{k}/cands"] = np.array(cands, np.uint32)
            out[f"{name}/choose{k}/result"] = np.array([bs, pad], np.uint64)
            out[f"{name}/choose{k}/fill"] = np.array([fill], np.float64)
    np.savez_compressed(HERE / "bell_ref_vectors.npz", **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
