"""Batched and served SpGEMM over the C ABI (spws_batch_* and spws_serve).

``Engine`` wraps one batch engine (device pools + captured hipGraph).
``DeviceRequest`` holds one request's device-resident operands, the upper
bounds its admission is decided on, and optionally a device output buffer
sized by that bound.  ``serve`` runs the native request loop in plain or
continuously batched mode and returns per-request finish times and output
nonzero counts; ``DeviceRequest.result`` then downloads a served product.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .device import Context, DeviceArray, DeviceCSR
from .gen import CSR, row_products


def default_caps(max_requests=512, max_rows=1 << 21, max_nnz=1 << 27, max_products=(1 << 29),
                 max_row_products=1 << 21, max_cols=1 << 16) -> _lib.BatchCaps:
    """Engine capacities; max_cols bounds the request width (0 = any width)."""
    return _lib.BatchCaps(max_requests, max_rows, max_rows, max_cols, max_nnz, max_nnz, max_products,
                          max_row_products)


class Engine:
    """One batch engine: pools sized by ``caps``, one stream, one graph."""

    def __init__(self, ctx: Context, caps: _lib.BatchCaps | None = None):
        self.ctx = ctx
        self.caps = caps if caps is not None else default_caps()
        h = C.c_void_p()
        check(lib().spws_batch_create(ctx.h, C.byref(self.caps), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().spws_batch_destroy(self.h)
            self.h = None

    def reset(self):
        check(lib().spws_batch_reset(self.h))

    def add(self, r: "DeviceRequest") -> int:
        """Slot of the request; raises SpwsError when the batch cannot hold it."""
        slot = C.c_int32()
        check(lib().spws_batch_add(self.h, r.m, r.k, r.n, r.A.pos.ptr, r.A.crd.ptr, r.A.vals.ptr, r.A.nnz,
                                   r.B.pos.ptr, r.B.crd.ptr, r.B.vals.ptr, r.B.nnz, r.products, r.max_row_products,
                                   C.byref(slot)))
        return slot.value

    def launch(self):
        check(lib().spws_batch_launch(self.h))

    def wait(self):
        check(lib().spws_batch_wait(self.h))

    def poll(self) -> bool:
        d = C.c_int()
        check(lib().spws_batch_poll(self.h, C.byref(d)))
        return bool(d.value)

    def nnz(self) -> int:
        v = C.c_int64()
        check(lib().spws_batch_nnz(self.h, C.byref(v)))
        return v.value

    def set_output(self, slot: int, r: "DeviceRequest"):
        """Have the batch copy request `slot`'s result into r's output buffer."""
        check(lib().spws_batch_set_output(self.h, int(slot), r.C_pos.ptr, r.C_crd.ptr, r.C_vals.ptr, r.C_cap))

    def result_nnz(self, slot: int) -> int:
        v = C.c_int64()
        check(lib().spws_batch_result_nnz(self.h, int(slot), C.byref(v)))
        return v.value

    def result(self, slot: int, m: int, n: int) -> CSR:
        """Host copy of one request's output (sorted CSR)."""
        nnz = C.c_int64()
        pos = np.zeros(m + 1, dtype=np.int32)
        pc, pv = C.c_void_p(), C.c_void_p()
        check(lib().spws_batch_result(self.h, int(slot), C.byref(nnz), pos.ctypes.data, C.byref(pc), C.byref(pv)))
        k = nnz.value
        crd = DeviceArray.view(self.ctx, pc.value, k, np.int32).to_host()
        val = DeviceArray.view(self.ctx, pv.value, k, np.float64).to_host()
        return CSR(m, n, pos, crd, val)


@dataclass
class DeviceRequest:
    m: int
    k: int
    n: int
    A: DeviceCSR
    B: DeviceCSR
    products: int
    max_row_products: int
    C_pos: DeviceArray | None = None
    C_crd: DeviceArray | None = None
    C_vals: DeviceArray | None = None
    C_cap: int = 0

    @classmethod
    def upload(cls, ctx: Context, A: CSR, B: CSR, output: bool = False) -> "DeviceRequest":
        total, mx = row_products(A, B)
        r = cls(A.nrows, A.ncols, B.ncols, DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B), total, mx)
        if output:
            r.alloc_output(ctx)
        return r

    def alloc_output(self, ctx: Context, capacity: int | None = None):
        """Device output buffer: m + 1 row pointers and `capacity` (default:
        the upper bound of output entries) columns and values."""
        cap = self.products if capacity is None else int(capacity)
        self.C_pos = DeviceArray(ctx, self.m + 1, np.int32)
        self.C_crd = DeviceArray(ctx, max(cap, 1), np.int32)
        self.C_vals = DeviceArray(ctx, max(cap, 1), np.float64)
        self.C_cap = cap

    def result(self, nnz: int) -> CSR:
        """Host copy of the served product (nnz as returned by serve)."""
        if self.C_pos is None or nnz < 0:
            raise ValueError("no result in this request's output buffer")
        ctx = self.C_pos.ctx
        crd = DeviceArray.view(ctx, self.C_crd.ptr, nnz, np.int32).to_host()
        vals = DeviceArray.view(ctx, self.C_vals.ptr, nnz, np.float64).to_host()
        return CSR(self.m, self.n, self.C_pos.to_host(), crd, vals)

    def c_struct(self, arrival: float) -> _lib.Request:
        out = (self.C_pos.ptr, self.C_crd.ptr, self.C_vals.ptr, self.C_cap) if self.C_pos is not None \
            else (None, None, None, 0)
        return _lib.Request(self.m, self.k, self.n, 0, self.A.pos.ptr, self.A.crd.ptr, self.A.vals.ptr, self.A.nnz,
                            self.B.pos.ptr, self.B.crd.ptr, self.B.vals.ptr, self.B.nnz, self.products,
                            self.max_row_products, float(arrival), *out)


def serve(ctx: Context, engines, reqs, arrivals, mode: str, max_batch_requests: int = 0,
          max_batch_products: int = 0, large_products: int = 0, large_engines: int = 0,
          with_nnz: bool = False):
    """Run the native request loop.  Returns (finish seconds per request,
    batch id per request (-1 = plain), stats dict), and with ``with_nnz``
    also the output nnz per request (-1: it exceeded the request's output
    capacity; requests without an output buffer report their nnz too)."""
    order = np.argsort(np.asarray(arrivals), kind="stable")
    if not np.all(order == np.arange(len(reqs))):
        raise ValueError("requests must be given in arrival order")
    arr = (_lib.Request * len(reqs))(*[r.c_struct(a) for r, a in zip(reqs, arrivals)])
    handles = (C.c_void_p * max(len(engines), 1))(*[e.h for e in engines])
    opt = _lib.ServeOptions(0 if mode == "plain" else 1, int(max_batch_requests), int(max_batch_products),
                            int(large_products), int(large_engines), 0)
    finish = np.zeros(len(reqs), dtype=np.float64)
    batch_of = np.zeros(len(reqs), dtype=np.int32)
    nnz = np.zeros(len(reqs), dtype=np.int64)
    st = _lib.ServeStats()
    check(lib().spws_serve(ctx.h, handles, len(engines), len(reqs), arr, C.byref(opt), finish.ctypes.data,
                           batch_of.ctypes.data, nnz.ctypes.data, C.byref(st)))
    stats = {"wall_seconds": st.wall_seconds, "batches": st.batches, "fallback_requests": st.fallback_requests}
    if with_nnz:
        return finish, batch_of, stats, nnz
    return finish, batch_of, stats


def latency_summary(arrivals, finish) -> dict:
    lat = (np.asarray(finish) - np.asarray(arrivals)) * 1e3
    return {"p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
            "mean_ms": round(float(lat.mean()), 4), "max_ms": round(float(lat.max()), 4)}
