"""Write tests/golden/*.json from the reference's own test fixtures.

Provenance (paths in the CellShard source tree):
  mask_groups_fixture.json  tests/mask_groups_runtime_test.cu:57-198 (inputs and
                            every asserted output).  Two assertions there
                            contradict the kernel (src/runtime/mask_groups.cu:
                            145-194): keep[3] == 0 (:116) and Blocked-ELL
                            group_counts[8] == 2 (:118).  Row 3 has sum 18,
                            2 detected and fraction-group 0 -> keep = 1; in the
                            Blocked-ELL fixture gene 3 (mask 0) holds the 2.0,
                            so group 2 of row 2 is 0.  Both are stored
                            corrected, with the reference value alongside.
  export_bell_fixture.json  tests/cellshard_export_runtime_test.cu:61-74 (2x4
                            Blocked-ELL, block 2) and :955-963 (its CSR).
Run: python tests/golden/make_fixtures.py
"""
import json
from pathlib import Path

HERE = Path(__file__).resolve().parent

mask_groups = {
    "rows": 4, "cols": 5,
    "group_masks": [1, 2, 3, 0, 4],
    "row_keep": [1, 0, 1, 1],
    "feature_keep": [1, 1, 0, 1, 1],
    "group_count": 3,
    "filter": [1.0, 1, 0.80, 0],
    "blocked_ell": {
        "nnz": 10, "block_size": 2, "ell_cols": 4,
        "block_col_idx": [0, 1, 0, 1, 0, 1, 1, 2],
        "values": [5, 3, 7, 0, 9, 0, 0, 0, 1, 4, 6, 2, 0, 8, 0, 10],
    },
    "compressed": {
        "ptr": [0, 3, 4, 8, 10],
        "idx": [0, 1, 2, 0, 0, 1, 2, 4, 1, 4],
        "values": [5, 3, 7, 9, 1, 4, 6, 2, 8, 10],
    },
    "sliced_ell": {
        "nnz": 9, "slice_rows": 4,
        "slice_row_offsets": [0, 4], "slice_widths": [3], "slice_slot_offsets": [0],
        "col_idx": [0, 1, 2, 0, 4294967295, 4294967295, 0, 1, 4, 1, 4, 4294967295],
        "values": [5, 3, 7, 9, 0, 0, 1, 4, 2, 8, 10, 0],
    },
    "expected": {
        "blocked_ell_masked": {
            "row_totals": {"0": 8.0, "1": 0.0, "2": 7.0, "3": 18.0},
            "detected": {"0": 2, "1": 0, "2": 3, "3": 2},
            "keep": {"0": 1, "1": 0, "2": 1, "3": 1},
            "group_counts": {"0": 5.0, "1": 3.0, "2": 0.0, "6": 1.0, "7": 4.0, "8": 0.0},
            "group_pct": {"0": 62.5, "4": 0.0},
            "corrections": {"keep[3]": {"reference_asserts": 0, "kernel_gives": 1},
                            "group_counts[8]": {"reference_asserts": 2.0, "kernel_gives": 0.0}},
        },
        "blocked_ell_unmasked_no_groups": {"group_count": 0, "row_totals": {"0": 15.0}, "detected": {"0": 3}},
        "compressed_masked": {"row_totals": {"2": 7.0}, "group_counts": {"8": 2.0}},
        "sliced_ell_masked": {"row_totals": {"0": 8.0, "3": 18.0}},
        "tolerance_abs": 1.0e-3,
    },
    "reoptimize": {
        "row_remap": [0, 4294967295, 1, 2],
        "feature_remap": [0, 1, 4294967295, 2, 3],
        "config": {"output_rows": 3, "output_cols": 4, "requested_bucket_count": 0, "output_block_size": 2,
                   "requested_layout": 3},
        "expected": {"kept_rows": 3, "kept_features": 4, "live_nnz": 7, "output_layout": 3},
    },
}

export_bell = {
    "rows": 2, "cols": 4, "nnz": 4, "block_size": 2, "ell_cols": 4,
    "block_col_idx": [0, 1],
    "values": [1, 0, 2, 0, 0, 3, 0, 4],
    "expected_csr": {"indptr": [0, 2, 4], "indices": [0, 2, 1, 3], "data": [1.0, 2.0, 3.0, 4.0]},
}

if __name__ == "__main__":
    (HERE / "mask_groups_fixture.json").write_text(json.dumps(mask_groups, indent=1) + "\n")
    (HERE / "export_bell_fixture.json").write_text(json.dumps(export_bell, indent=1) + "\n")
    print("wrote fixtures")
