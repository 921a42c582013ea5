"""Build the HIP library (gfx950) and the CPU oracle libraries.

Outputs (in-tree, git-ignored):
  cellshard_amd/libcellshard_amd.so   HIP kernels + C ABI
  oracle/liboracle.so                 CPU restatement (test infrastructure)
  oracle/_ref/...                     reference CPU converters, when the
                                      reference sources are present
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libcellshard_amd.so"

HIP_SOURCES = [
    "runtime.hip",
    "mask_groups.hip",
    "label_reduce.hip",
    "convert.hip",
    "transpose.hip",
    "bucket.hip",
    "synth.hip",
    "staging.hip",
    "host_convert.hip",
    "serve.hip",
    "partition_tune.hip",
    "coo_convert.hip",
    "execution_layout.hip",
    "filter.hip",
    "column_order.hip",
    "cspack.hip",
    "sharded.hip",
    "group_reduce.hip",
    "bucket_ops.hip",
    "pipeline_convert.hip",
]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found")
    return p


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _jobs() -> int:
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    if n <= 0:
        n = min(16, os.cpu_count() or 4)
    return max(1, n)


def build_hip(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


CPP_TESTS = ["drop_in_boundary"]
CPP_TEST_DIR = ROOT / "tests" / "cpp"


def build_cpp_tests(verbose: bool = False) -> list:
    """C++ callers of the drop-in header (tests/cpp/*.cpp), linked against the
    HIP library by a relative rpath: tests/cpp/bin/<name>."""
    hipcc = _hipcc()
    out_dir = CPP_TEST_DIR / "bin"
    out_dir.mkdir(exist_ok=True)
    outs = []
    for name in CPP_TESTS:
        out = out_dir / name
        tmp = out.with_suffix(".tmp")
        _run([hipcc, "--offload-arch=gfx950", "-std=c++17", "-O1", "-Wall", "-I", str(ROOT / "include"),
              str(CPP_TEST_DIR / f"{name}.cpp"), "-o", str(tmp), "-L", str(PKG), "-lcellshard_amd",
              "-Wl,-rpath,$ORIGIN/../../../cellshard_amd"])
        os.replace(tmp, out)
        outs.append(out)
        if verbose:
            print(f"built {out}")
    return outs


def build_oracle(verbose: bool = False) -> Path:
    """Build the CPU oracle (plain C) and, if the reference tree exists, the
    reference CPU converter probe under oracle/_ref."""
    odir = ROOT / "oracle"
    out = odir / "liboracle.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-o", str(tmp), str(odir / "oracle.c"), "-lm"])
    os.replace(tmp, out)
    ref_script = odir / "build_ref.sh"
    if Path("/root/reference").exists() and ref_script.exists():
        _run(["bash", str(ref_script)])
    if verbose:
        print(f"built {out}")
    return out


def main() -> None:
    build_hip(verbose=True)
    build_cpp_tests(verbose=True)
    build_oracle(verbose=True)


if __name__ == "__main__":
    sys.exit(main())
