#!/usr/bin/env bash
# Compile the reference's CPU Blocked-ELL converter (header-only, CPU build)
# and its raw payload codec (src/io/pack/packfile.cu, plain C++) from the
# reference sources in place, through oracle/ref_driver.cpp and
# oracle/ref_pack_driver.cpp.
# Output: oracle/_ref/libref_convert.so, oracle/_ref/libref_pack.so
# (git-ignored).  Needs only g++.
set -euo pipefail
REF="${CELLSHARD_REFERENCE:-/root/reference}"
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
if [ ! -f "$REF/src/convert/blocked_ell_from_compressed.cuh" ]; then
  echo "reference sources not found at $REF; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=c++17 -O2 -fPIC -shared -w \
  -DCELLSHARD_ENABLE_CUDA=0 -DCELLSHARD_ENABLE_CELLERATOR_QUANTIZED=0 \
  -I"$REF" -I"$REF/include" -I"$REF/src" \
  "$HERE/ref_driver.cpp" -o "$HERE/_ref/libref_convert.so.tmp"
mv "$HERE/_ref/libref_convert.so.tmp" "$HERE/_ref/libref_convert.so"
g++ -std=c++17 -O2 -fPIC -shared -w \
  -DCELLSHARD_ENABLE_CUDA=0 -DCELLSHARD_ENABLE_CELLERATOR_QUANTIZED=0 \
  -I"$REF" -I"$REF/include" -I"$REF/src" \
  "$HERE/ref_pack_driver.cpp" -o "$HERE/_ref/libref_pack.so.tmp"
mv "$HERE/_ref/libref_pack.so.tmp" "$HERE/_ref/libref_pack.so"
