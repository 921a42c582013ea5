"""Serving loop front end: resident CSC shards, feature-subset grouped
queries admitted into output slots under a device-memory budget, continuous
batching (csamd_serve_* in the C ABI).  torch is used for device memory only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi as A
from ._lib import check, load
from .device import DeviceCSR, ptr

SUM, SUMSQ, NNZ = A.SERVE_SUM, A.SERVE_SUMSQ, A.SERVE_NNZ


def keep_threshold(fraction: float) -> int:
    """Seeded row-mask threshold: a row is kept iff its 24-bit hash < this."""
    return int(min(max(fraction, 0.0), 1.0) * (1 << 24) + 0.5)


def make_query(shard: int, op: int, group_count: int, features: torch.Tensor, row_labels: torch.Tensor | None = None,
               row_keep: torch.Tensor | None = None, label_seed: int = 0, mask_seed: int = 0,
               keep_fraction: float = 1.0) -> A.serve_query:
    """A query struct; `features` (device u32, ascending) and the optional
    label / mask arrays must stay alive until the query finishes."""
    q = A.serve_query()
    q.shard, q.op, q.group_count = shard, op, group_count
    q.feature_count = int(features.numel())
    q.features = ptr(features)
    q.row_labels = ptr(row_labels)
    q.row_keep = ptr(row_keep)
    q.label_seed, q.mask_seed = label_seed & 0xFFFFFFFF, mask_seed & 0xFFFFFFFF
    q.keep_thr24 = keep_threshold(keep_fraction)
    return q


class Server:
    def __init__(self, device: int = 0, max_slots: int = 64, max_batch_queries: int = 64, max_batches_in_flight: int = 2,
                 slot_bytes: int = 5000 * 512 * 8, budget_bytes: int = 200 << 30):
        lib = load()
        self.ws = A.serve_workspace()
        lib.csamd_serve_init(C.byref(self.ws))
        cfg = A.serve_config()
        cfg.device, cfg.max_slots, cfg.max_batch_queries = device, max_slots, max_batch_queries
        cfg.max_batches_in_flight, cfg.slot_bytes, cfg.budget_bytes = max_batches_in_flight, slot_bytes, budget_bytes
        check(lib.csamd_serve_setup(C.byref(self.ws), C.byref(cfg)), "serve setup")
        self.shards: list[DeviceCSR] = []  # keeps the device arrays alive

    def close(self):
        if self.ws is not None:
            load().csamd_serve_clear(C.byref(self.ws))
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_shard(self, csc: DeviceCSR) -> int:
        if csc.axis != A.COMPRESSED_BY_COL:
            raise ValueError("serving shards are column-major (CSC)")
        sid = C.c_uint32(0)
        v = csc.view()
        check(load().csamd_serve_add_shard(C.byref(self.ws), C.byref(v), csc.value_type, C.byref(sid)), "serve add_shard")
        self.shards.append(csc)
        return int(sid.value)

    def submit(self, q: A.serve_query) -> int:
        t = C.c_uint64(0)
        check(load().csamd_serve_submit(C.byref(self.ws), C.byref(q), C.byref(t)), "serve submit")
        return int(t.value)

    def pump(self) -> int:
        n = C.c_uint32(0)
        check(load().csamd_serve_pump(C.byref(self.ws), C.byref(n)), "serve pump")
        return int(n.value)

    def drain(self) -> None:
        check(load().csamd_serve_drain(C.byref(self.ws)), "serve drain")

    def result(self, ticket: int, fetch: bool = True) -> dict:
        r = A.serve_result()
        check(load().csamd_serve_get_result(C.byref(self.ws), ticket, C.byref(r)), "serve result")
        out = {"status": int(r.status), "latency_us": float(r.latency_us), "op": int(r.op)}
        if fetch and r.status == 2:
            from .ops import device_to_numpy
            k, G = int(r.feature_count), int(r.group_count)
            dt = np.uint32 if r.op == NNZ else np.float64
            out["out"] = device_to_numpy(r.out, k * G, dt).reshape(k, G)
        return out

    def release(self, ticket: int) -> None:
        check(load().csamd_serve_release(C.byref(self.ws), ticket), "serve release")

    def replay(self, queries: list[A.serve_query], arrival_us: np.ndarray, hold: bool = False):
        """Submit query i at arrival_us[i] (microseconds from the start), pump
        until all finish.  Returns (tickets or None, latency_us, wall_us)."""
        n = len(queries)
        arr = (A.serve_query * n)(*queries)
        arrival = np.ascontiguousarray(arrival_us, dtype=np.float64)
        lat = np.zeros(n, np.float64)
        wall = C.c_double(0.0)
        tickets = (C.c_uint64 * n)() if hold else None
        check(load().csamd_serve_replay(C.byref(self.ws), arr, arrival.ctypes.data_as(C.POINTER(C.c_double)), n,
                                        tickets, lat.ctypes.data_as(C.POINTER(C.c_double)), C.byref(wall)), "serve replay")
        return (list(tickets) if hold else None), lat, float(wall.value)


def materialize(rows: int, group_count: int, label_seed: int, mask_seed: int, keep_fraction: float,
                device: torch.device | None = None):
    """Seeded labels (u32) and row mask (u8) of a query as device tensors."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    labels = torch.empty(max(rows, 1), dtype=torch.int32, device=dev)
    keep = torch.empty(max(rows, 1), dtype=torch.uint8, device=dev)
    check(load().csamd_serve_materialize(rows, group_count, label_seed & 0xFFFFFFFF, mask_seed & 0xFFFFFFFF,
                                         keep_threshold(keep_fraction), ptr(labels), ptr(keep),
                                         torch.cuda.current_stream().cuda_stream), "serve materialize")
    return labels, keep
