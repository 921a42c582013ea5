"""The drop-in C++ layer compiled into a running program: tests/cpp/*.cpp
include only <cellshard_amd/cellshard.hh>, are built against
libcellshard_amd.so by the package build, and run here on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent / "cpp" / "bin"


@pytest.mark.parametrize("name", ["drop_in_boundary"])
def test_cpp_program(gpu, name):
    exe = BIN / name
    assert exe.exists(), f"{exe} missing: run __graft_entry__.build()"
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert ", 0 failed" in r.stdout
