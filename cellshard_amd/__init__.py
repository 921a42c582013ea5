"""cellshard_amd: MI355X (gfx950) implementation of CellShard's sparse
execution hot path behind a C ABI (include/cellshard_amd/cellshard_amd.h).

Modules:
  abi      ctypes mirror of the C structs (reference layouts)
  runtime  mask_groups reference API (init/setup/reserve/compute_sparse_group_reduce/...)
  ops      conversion, transpose, bucketing, row-label reduction, generator
  device   device-resident CSR / Blocked-ELL / Sliced-ELL containers
"""
from . import abi
from ._lib import CellShardError, LibraryMissing, LIB_PATH, last_error, load

__all__ = ["abi", "load", "last_error", "CellShardError", "LibraryMissing", "LIB_PATH"]
