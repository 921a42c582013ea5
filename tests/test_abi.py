"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and the ctypes mirror matches the C compiler's layouts."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from cellshard_amd import abi as A
from cellshard_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "cellshard_amd" / "cellshard_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csamd_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes prototype and vice versa
    assert sorted(_lib.PROTOTYPES) == names


def test_no_gpu_calls_are_safe_on_host():
    lib = _lib.load()
    assert lib.csamd_abi_version() == 1
    assert lib.csamd_device_count() >= 0
    assert lib.csamd_clamp_bucket_count(0, 5) == 1
    assert lib.csamd_clamp_bucket_count(10, 0) == 1
    assert lib.csamd_clamp_bucket_count(10, 50) == 10
    assert lib.csamd_compressed_transpose_scratch_bytes(100, 50, 1000) > 0
    sz = C.c_size_t(0)
    assert lib.csamd_major_nnz_bucket_sort_scratch_bytes(1000, C.byref(sz)) == 1 and sz.value > 0


def test_error_paths_return_zero_and_set_message():
    lib = _lib.load()
    out = A.auto_reoptimize_prediction()
    assert lib.csamd_predict_masked_sparse_reoptimization(None, C.byref(out)) == 0
    assert "null" in _lib.last_error()
    inp = A.auto_reoptimize_prediction_input()
    assert lib.csamd_predict_masked_sparse_reoptimization(C.byref(inp), C.byref(out)) == 1
    assert out.should_reoptimize == 0
    ws = A.sparse_group_reduce_workspace()
    lib.csamd_group_reduce_init(C.byref(ws))
    assert ws.device == -1
    # group_count > 32 is rejected before any device work
    assert lib.csamd_group_reduce_reserve(C.byref(ws), 4, 5, 0, 33) == 0
    assert "group_count" in _lib.last_error()


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "cellshard_amd/cellshard_amd.h"
#define S(T) printf("%s size %zu\n", #T, sizeof(T));
#define F(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f));
int main(void) {
%BODY%
  return 0;
}
"""


def test_struct_layouts_match_c(tmp_path):
    body = []
    for cname, cls in A.ALL_STRUCTS.items():
        body.append(f"S({cname})")
        for fname, _ in cls._fields_:
            if not fname.startswith("_"):
                body.append(f"F({cname}, {fname})")
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C.replace("%BODY%", "\n".join(body)))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    got = {}
    for line in out:
        k, kind, v = line.split(" ")[0], line.split(" ")[1] if " size " in line else None, line.split(" ")[-1]
        got[line.rsplit(" ", 1)[0]] = int(v)
    for cname, cls in A.ALL_STRUCTS.items():
        assert got[f"{cname} size"] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            if not fname.startswith("_"):
                assert got[f"{cname}.{fname}"] == getattr(cls, fname).offset, (cname, fname)


def test_reference_layout_sizes():
    # reference structs are alignas(16) (runtime/device/sharded_device.cuh:67-98)
    assert C.sizeof(A.compressed_view) == 48
    assert C.sizeof(A.blocked_ell_view) == 48
    assert C.sizeof(A.sliced_ell_view) == 64
    assert C.sizeof(A.sparse_group_reduce_workspace) == 128


@pytest.mark.parametrize("path", ["include/cellshard_amd/cellshard.hh"])
def test_cpp_wrappers_compile(tmp_path, path):
    if not (ROOT / path).exists():
        pytest.skip(f"{path} not present")
    src = tmp_path / "use.cpp"
    src.write_text(f'#include "{path.split("include/")[1]}"\nint main() {{ return 0; }}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", str(ROOT / "include"), str(src)], check=True)


def test_cpp_program_host_part_runs_without_a_device():
    """The sharded<T> metadata part of the C++ drop-in program (compiled
    against the library, host code only) runs here."""
    exe = ROOT / "tests" / "cpp" / "bin" / "drop_in_boundary"
    if not exe.exists():
        from cellshard_amd import build
        build.build_cpp_tests()
    r = subprocess.run([str(exe), "--host-only"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0 and ", 0 failed" in r.stdout, r.stdout
