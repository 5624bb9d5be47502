"""Batched and served SpGEMM over the C ABI (spws_batch_* and spws_serve).

``Engine`` wraps one batch engine (device pools + captured hipGraph).
``DeviceRequest`` holds one request's device-resident operands and the
upper bounds its admission is decided on.  ``serve`` runs the native request
loop in plain or continuously batched mode and returns per-request finish
times.
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

    @classmethod
    def upload(cls, ctx: Context, A: CSR, B: CSR) -> "DeviceRequest":
        total, mx = row_products(A, B)
        return cls(A.nrows, A.ncols, B.ncols, DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B), total, mx)

    def c_struct(self, arrival: float) -> _lib.Request:
        return _lib.Request(self.m, self.k, self.n, 0, self.A.pos.ptr, self.A.crd.ptr, self.A.vals.ptr, self.A.nnz,
                            self.B.pos.ptr, self.B.crd.ptr, self.B.vals.ptr, self.B.nnz, self.products,
                            self.max_row_products, float(arrival))


def serve(ctx: Context, engines, reqs, arrivals, mode: str, max_batch_requests: int = 0,
          max_batch_products: int = 0):
    """Run the native request loop.  Returns (finish seconds per request,
    batch id per request (-1 = plain), stats dict)."""
    order = np.argsort(np.asarray(arrivals), kind="stable")
    if not np.all(order == np.arange(len(reqs))):
        raise ValueError("requests must be given in arrival order")
    arr = (_lib.Request * len(reqs))(*[r.c_struct(a) for r, a in zip(reqs, arrivals)])
    handles = (C.c_void_p * max(len(engines), 1))(*[e.h for e in engines])
    opt = _lib.ServeOptions(0 if mode == "plain" else 1, int(max_batch_requests), int(max_batch_products))
    finish = np.zeros(len(reqs), dtype=np.float64)
    batch_of = np.zeros(len(reqs), dtype=np.int32)
    st = _lib.ServeStats()
    check(lib().spws_serve(ctx.h, handles, len(engines), len(reqs), arr, C.byref(opt), finish.ctypes.data,
                           batch_of.ctypes.data, C.byref(st)))
    return finish, batch_of, {"wall_seconds": st.wall_seconds, "batches": st.batches,
                              "fallback_requests": st.fallback_requests}


def latency_summary(arrivals, finish) -> dict:
    lat = (np.asarray(finish) - np.asarray(arrivals)) * 1e3
    return {"p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
            "mean_ms": round(float(lat.mean()), 4), "max_ms": round(float(lat.max()), 4)}
