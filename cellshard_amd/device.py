"""Device-resident sparse containers (HBM via torch allocations).

The containers hold raw value bits (fp16/bf16 as int16, fp8 as uint8, fp32 as
float32) so every layout round-trips bit-exactly; `.view()` returns the
reference-layout ctypes view (runtime/device/sharded_device.cuh:67-98).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import abi as A

_TORCH_BITS = {1: torch.uint8, 2: torch.int16, 4: torch.float32}
_NUMPY_BITS = {1: np.uint8, 2: np.int16, 4: np.float32}


def value_bytes(value_type: int) -> int:
    vb = A.VALUE_BYTES.get(value_type, 0)
    if vb not in (1, 2, 4):
        raise ValueError(f"unsupported value type {value_type}")
    return vb


def torch_bits_dtype(value_type: int) -> torch.dtype:
    return _TORCH_BITS[value_bytes(value_type)]


def numpy_bits_dtype(value_type: int):
    return _NUMPY_BITS[value_bytes(value_type)]


def default_device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


def to_dev_u32(a, device=None) -> torch.Tensor:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint32))
    return torch.from_numpy(a.view(np.int32)).to(device or default_device())


def to_dev_u8(a, device=None) -> torch.Tensor:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint8))
    return torch.from_numpy(a).to(device or default_device())


def to_dev_f32(a, device=None) -> torch.Tensor:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return torch.from_numpy(a).to(device or default_device())


def to_dev_bits(a, value_type: int, device=None) -> torch.Tensor:
    dt = numpy_bits_dtype(value_type)
    a = np.ascontiguousarray(np.asarray(a))
    if a.dtype.itemsize != np.dtype(dt).itemsize:
        raise ValueError(f"value array itemsize {a.dtype.itemsize} does not match value type {value_type}")
    return torch.from_numpy(a.view(dt)).to(device or default_device())


def u32_np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy().view(np.uint32)


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr() if t.numel() > 0 else None


@dataclass
class DeviceCSR:
    """Compressed (CSR when axis == 0, CSC when axis == 1)."""
    rows: int
    cols: int
    major_ptr: torch.Tensor   # int32 bits of u32 [major + 1]
    minor_idx: torch.Tensor   # int32 bits of u32 [nnz]
    val: torch.Tensor         # raw value bits [nnz]
    value_type: int = A.VALUE_F16
    axis: int = A.COMPRESSED_BY_ROW

    @property
    def nnz(self) -> int:
        return int(self.minor_idx.numel())

    @property
    def major_dim(self) -> int:
        return self.cols if self.axis == A.COMPRESSED_BY_COL else self.rows

    def view(self) -> A.compressed_view:
        v = A.compressed_view()
        v.rows, v.cols, v.nnz, v.axis = self.rows, self.cols, self.nnz, self.axis
        v.majorPtr, v.minorIdx, v.val = ptr(self.major_ptr), ptr(self.minor_idx), ptr(self.val)
        return v

    @classmethod
    def from_numpy(cls, rows, cols, indptr, indices, val_bits, value_type=A.VALUE_F16, axis=A.COMPRESSED_BY_ROW, device=None):
        return cls(int(rows), int(cols), to_dev_u32(indptr, device), to_dev_u32(indices, device),
                   to_dev_bits(val_bits, value_type, device), value_type, axis)

    def to_numpy(self):
        return (u32_np(self.major_ptr), u32_np(self.minor_idx), self.val.detach().cpu().numpy())

    def format_bytes(self) -> int:
        return (self.major_dim + 1) * 4 + self.nnz * (4 + value_bytes(self.value_type))


@dataclass
class DeviceBlockedELL:
    rows: int
    cols: int
    nnz: int
    block_size: int
    ell_cols: int
    block_col_idx: torch.Tensor  # int32 bits [row_blocks * width]
    val: torch.Tensor            # raw bits [rows * ell_cols]
    value_type: int = A.VALUE_F16

    @property
    def row_blocks(self) -> int:
        return (self.rows + self.block_size - 1) // self.block_size if self.block_size else 0

    @property
    def width(self) -> int:
        return self.ell_cols // self.block_size if self.block_size else 0

    def view(self) -> A.blocked_ell_view:
        v = A.blocked_ell_view()
        v.rows, v.cols, v.nnz, v.block_size, v.ell_cols = self.rows, self.cols, self.nnz, self.block_size, self.ell_cols
        v.blockColIdx, v.val = ptr(self.block_col_idx), ptr(self.val)
        return v

    @classmethod
    def from_numpy(cls, rows, cols, nnz, block_size, ell_cols, block_col_idx, val_bits, value_type=A.VALUE_F16, device=None):
        return cls(int(rows), int(cols), int(nnz), int(block_size), int(ell_cols), to_dev_u32(block_col_idx, device),
                   to_dev_bits(val_bits, value_type, device), value_type)

    def to_numpy(self):
        return u32_np(self.block_col_idx), self.val.detach().cpu().numpy()

    def format_bytes(self) -> int:
        return self.row_blocks * self.width * 4 + self.rows * self.ell_cols * value_bytes(self.value_type)


@dataclass
class DeviceSlicedELL:
    rows: int
    cols: int
    nnz: int
    slice_count: int
    slice_rows: int                 # uniform slice height (0 if not uniform)
    slice_row_offsets: torch.Tensor  # [S + 1]
    slice_widths: torch.Tensor       # [S]
    slice_slot_offsets: torch.Tensor # [S]
    col_idx: torch.Tensor            # [total_slots]
    val: torch.Tensor                # [total_slots]
    value_type: int = A.VALUE_F16

    @property
    def total_slots(self) -> int:
        return int(self.col_idx.numel())

    def view(self) -> A.sliced_ell_view:
        v = A.sliced_ell_view()
        v.rows, v.cols, v.nnz, v.slice_count, v.slice_rows = self.rows, self.cols, self.nnz, self.slice_count, self.slice_rows
        v.slice_row_offsets, v.slice_widths = ptr(self.slice_row_offsets), ptr(self.slice_widths)
        v.slice_slot_offsets, v.col_idx, v.val = ptr(self.slice_slot_offsets), ptr(self.col_idx), ptr(self.val)
        return v

    @classmethod
    def from_numpy(cls, rows, cols, nnz, slice_rows, offsets, widths, slot_offsets, col_idx, val_bits, value_type=A.VALUE_F16,
                   device=None):
        offsets = np.asarray(offsets, dtype=np.uint32)
        return cls(int(rows), int(cols), int(nnz), int(len(offsets) - 1), int(slice_rows), to_dev_u32(offsets, device),
                   to_dev_u32(widths, device), to_dev_u32(slot_offsets, device), to_dev_u32(col_idx, device),
                   to_dev_bits(val_bits, value_type, device), value_type)

    def to_numpy(self):
        return (u32_np(self.slice_row_offsets), u32_np(self.slice_widths), u32_np(self.slice_slot_offsets),
                u32_np(self.col_idx), self.val.detach().cpu().numpy())

    def format_bytes(self) -> int:
        return (2 * self.slice_count + 1) * 4 + self.total_slots * (4 + value_bytes(self.value_type))
