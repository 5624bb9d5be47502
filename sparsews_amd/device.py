"""Device-resident host API over the C ABI: contexts, device arrays, and the
kernels' device entry points.  All device memory here is allocated through
libspws (hipMalloc); no other GPU runtime is required."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .formats import CSF3, DCSR
from .gen import CSR


class DeviceArray:
    """A device allocation holding a 1-D numpy-typed array."""

    def __init__(self, ctx: "Context", count: int, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(self.count * self.dtype.itemsize, 1)
        self.ptr = ctx.alloc(self.nbytes)
        self.owned = True

    @classmethod
    def view(cls, ctx, ptr: int, count: int, dtype):
        """Wrap memory owned elsewhere (e.g. a context-owned output)."""
        self = cls.__new__(cls)
        self.ctx, self.dtype, self.count = ctx, np.dtype(dtype), int(count)
        self.nbytes = self.count * self.dtype.itemsize
        self.ptr, self.owned = int(ptr or 0), False
        return self

    @classmethod
    def from_host(cls, ctx, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        self = cls(ctx, arr.size, arr.dtype)
        if arr.size:
            ctx.h2d(self.ptr, arr)
        return self

    def to_host(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        if self.count:
            self.ctx.d2h(out, self.ptr)
        return out

    def free(self):
        if self.owned and self.ptr:
            self.ctx.free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One spws_ctx: a device, a stream and the context-owned workspaces."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = C.c_void_p()
        check(lib().spws_ctx_create(int(device), C.c_void_p(stream) if stream else None, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().spws_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- memory
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(lib().spws_dev_alloc(self.h, int(nbytes), C.byref(p)))
        return int(p.value)

    def free(self, p: int):
        check(lib().spws_dev_free(self.h, C.c_void_p(p)))

    def h2d(self, dst: int, arr: np.ndarray):
        check(lib().spws_memcpy_h2d(self.h, C.c_void_p(dst), C.c_void_p(arr.ctypes.data), arr.nbytes))

    def d2h(self, arr: np.ndarray, src: int):
        check(lib().spws_memcpy_d2h(self.h, C.c_void_p(arr.ctypes.data), C.c_void_p(src), arr.nbytes))

    def sync(self):
        check(lib().spws_sync(self.h))

    def copy(self, dst: int, src: int, nbytes: int):
        check(lib().spws_device_copy(self.h, C.c_void_p(dst), C.c_void_p(src), int(nbytes)))

    @property
    def stream(self) -> int:
        return int(lib().spws_ctx_get_stream(self.h) or 0)

    def stats(self) -> dict:
        s = _lib.Stats()
        check(lib().spws_ctx_get_stats(self.h, C.byref(s)))
        return {"products": s.products, "nnz": s.nnz, "rows_per_path": list(s.rows_per_path),
                "chunks": s.chunks, "sort_bits": s.sort_bits, "outer_method": s.outer_method,
                "window_bits": s.window_bits}

    # ---- kernel timing (HIP events around every launch on the context's stream)
    def set_profiling(self, on: bool):
        check(lib().spws_ctx_set_profiling(self.h, 1 if on else 0))

    def set_profiling_only(self, name=None):
        """Time only the launches called `name` (None: every launch)."""
        check(lib().spws_ctx_set_profiling_only(self.h, name.encode() if name else None))

    def reset_kernel_times(self):
        check(lib().spws_ctx_reset_kernel_times(self.h))

    def kernel_times(self) -> dict:
        """{kernel name: (total ms, launches)} since the last reset."""
        out, i = {}, 0
        buf = C.create_string_buffer(128)
        ms, n = C.c_double(), C.c_int64()
        while lib().spws_ctx_kernel_times(self.h, i, buf, 128, C.byref(ms), C.byref(n)) == 0:
            out[buf.value.decode()] = (ms.value, n.value)
            i += 1
        return out

    # ---- tuning
    def rowwise_tuning(self) -> dict:
        t = _lib.RowwiseTuning()
        check(lib().spws_ctx_get_rowwise_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def set_rowwise_tuning(self, **kw):
        cur = self.rowwise_tuning()
        cur.update(kw)
        t = _lib.RowwiseTuning(**cur)
        check(lib().spws_ctx_set_rowwise_tuning(self.h, C.byref(t)))


@dataclass
class DeviceCSR:
    nrows: int
    ncols: int
    nnz: int
    pos: DeviceArray
    crd: DeviceArray
    vals: DeviceArray | None

    @classmethod
    def upload(cls, ctx: Context, a: CSR, with_vals: bool = True) -> "DeviceCSR":
        return cls(a.nrows, a.ncols, a.nnz,
                   DeviceArray.from_host(ctx, np.asarray(a.pos, dtype=np.int32)),
                   DeviceArray.from_host(ctx, np.asarray(a.crd, dtype=np.int32)),
                   DeviceArray.from_host(ctx, np.asarray(a.vals, dtype=np.float64)) if with_vals else None)

    def download(self) -> CSR:
        return CSR(self.nrows, self.ncols, self.pos.to_host(), self.crd.to_host(),
                   self.vals.to_host() if self.vals is not None else np.zeros(0))


def spgemm_rowwise(ctx: Context, A: DeviceCSR, B: DeviceCSR) -> DeviceCSR:
    """C = A*B on the device; the result views context-owned buffers that stay
    valid until the next row-wise call on ``ctx``."""
    if A.ncols != B.nrows:
        raise ValueError("inner dimensions differ")
    nnz = C.c_int64()
    pp, pc, pv = C.c_void_p(), C.c_void_p(), C.c_void_p()
    check(lib().spws_spgemm_rowwise(ctx.h, A.nrows, A.ncols, B.ncols, A.pos.ptr, A.crd.ptr, A.vals.ptr,
                                    B.pos.ptr, B.crd.ptr, B.vals.ptr, C.byref(nnz), C.byref(pp),
                                    C.byref(pc), C.byref(pv)))
    n = nnz.value
    return DeviceCSR(A.nrows, B.ncols, n, DeviceArray.view(ctx, pp.value, A.nrows + 1, np.int32),
                     DeviceArray.view(ctx, pc.value, n, np.int32), DeviceArray.view(ctx, pv.value, n, np.float64))


def spgemm_rowwise_symbolic(ctx: Context, A: DeviceCSR, B: DeviceCSR) -> np.ndarray:
    """Exact C row pointers (the assemble phase), returned to the host."""
    nnz = C.c_int64()
    pp = C.c_void_p()
    check(lib().spws_spgemm_rowwise_symbolic(ctx.h, A.nrows, A.ncols, B.ncols, A.pos.ptr, A.crd.ptr,
                                             B.pos.ptr, B.crd.ptr, C.byref(nnz), C.byref(pp)))
    return DeviceArray.view(ctx, pp.value, A.nrows + 1, np.int32).to_host()


@dataclass
class DeviceDCSR:
    """DCSR (or DCSC) on the device: stored outer ids, pos, crd, vals.
    ``l1`` is None for a dense outer level (plain CSR/CSC)."""
    nouter: int
    l1: DeviceArray | None
    pos: DeviceArray
    crd: DeviceArray
    vals: DeviceArray

    @classmethod
    def upload(cls, ctx: Context, d: DCSR, dense_outer: bool = False) -> "DeviceDCSR":
        if dense_outer:
            n = len(d.l2_pos) - 1
            return cls(n, None, DeviceArray.from_host(ctx, d.l2_pos), DeviceArray.from_host(ctx, d.l2_crd),
                       DeviceArray.from_host(ctx, d.vals))
        return cls(len(d.l1_crd), DeviceArray.from_host(ctx, d.l1_crd), DeviceArray.from_host(ctx, d.l2_pos),
                   DeviceArray.from_host(ctx, d.l2_crd), DeviceArray.from_host(ctx, d.vals))

    @classmethod
    def from_csr(cls, ctx: Context, a: CSR) -> "DeviceDCSR":
        """A plain CSR as a dense-outer DCSR (every row index stored implicitly)."""
        return cls(a.nrows, None, DeviceArray.from_host(ctx, np.asarray(a.pos, dtype=np.int32)),
                   DeviceArray.from_host(ctx, a.crd), DeviceArray.from_host(ctx, a.vals))


@dataclass
class DeviceCSR64:
    """Context-owned CSR output with int64 row pointers."""
    nrows: int
    ncols: int
    nnz: int
    pos: DeviceArray
    crd: DeviceArray
    vals: DeviceArray

    def download(self) -> CSR:
        return CSR(self.nrows, self.ncols, self.pos.to_host(), self.crd.to_host(), self.vals.to_host())

    def rows(self, rows) -> tuple:
        """(pos slice, crd, vals) of selected rows, copied to the host."""
        pos = self.pos.to_host()
        out = []
        for r in rows:
            b, e = int(pos[r]), int(pos[r + 1])
            crd = DeviceArray.view(self.crd.ctx, self.crd.ptr + 4 * b, e - b, np.int32).to_host()
            val = DeviceArray.view(self.vals.ctx, self.vals.ptr + 8 * b, e - b, np.float64).to_host()
            out.append((crd, val))
        return pos, out


def set_outer_chunking(ctx: Context, max_products: int = 0, max_rows: int = 0):
    check(lib().spws_ctx_set_outer_chunking(ctx.h, int(max_products), int(max_rows)))


def set_outer_method(ctx: Context, method: int = 0, window_log2: int = 0):
    """0 automatic, 1 LSD radix sort, 2 row/window buckets, 3 fused row
    windows, 4 fused row windows with atomic (unordered) adds (see spws.h)."""
    check(lib().spws_ctx_set_outer_method(ctx.h, int(method), int(window_log2)))


def spgemm_outer(ctx: Context, A: DeviceDCSR, B: DeviceDCSR, m: int, k: int, n: int) -> DeviceCSR64:
    """C = A*B with A (m x k) in DCSC and B (k x n) in DCSR through the
    coordinate workspace: expand, sort on (i, k) (row/column-window buckets
    or an LSD radix sort of packed keys, see set_outer_method), segmented
    reduce; int64 row pointers."""
    nnz = C.c_int64()
    pp, pc, pv = C.c_void_p(), C.c_void_p(), C.c_void_p()
    check(lib().spws_spgemm_outer(ctx.h, int(m), int(k), int(n), A.nouter, A.l1.ptr if A.l1 else None,
                                  A.pos.ptr, A.crd.ptr, A.vals.ptr, B.nouter, B.l1.ptr if B.l1 else None,
                                  B.pos.ptr, B.crd.ptr, B.vals.ptr, C.byref(nnz), C.byref(pp), C.byref(pc),
                                  C.byref(pv)))
    n_ = nnz.value
    return DeviceCSR64(m, n, n_, DeviceArray.view(ctx, pp.value, m + 1, np.int64),
                       DeviceArray.view(ctx, pc.value, n_, np.int32), DeviceArray.view(ctx, pv.value, n_, np.float64))


def csr_transpose(ctx: Context, A: DeviceCSR) -> DeviceCSR:
    """A^T as a new device CSR (arrays owned by the result), each row's
    columns in increasing order (the device form of the reference's
    transpose(), nonfuse/CSC_CSR_T_hash.h:352)."""
    pos = DeviceArray(ctx, A.ncols + 1, np.int32)
    crd = DeviceArray(ctx, A.nnz, np.int32)
    vals = DeviceArray(ctx, A.nnz, np.float64)
    check(lib().spws_csr_transpose_dev(ctx.h, int(A.nrows), int(A.ncols), int(A.nnz), A.pos.ptr, A.crd.ptr,
                                       A.vals.ptr, pos.ptr, crd.ptr, vals.ptr))
    ctx.sync()
    return DeviceCSR(A.ncols, A.nrows, A.nnz, pos, crd, vals)


def csr64_transpose(ctx: Context, A: DeviceCSR64) -> DeviceCSR:
    """The transpose of an outer-product result (int64 row pointers) as a
    new device CSR with int32 row pointers: the second half of the non-fused
    transposed product, its first half's output left on the device."""
    pos = DeviceArray(ctx, A.ncols + 1, np.int32)
    crd = DeviceArray(ctx, A.nnz, np.int32)
    vals = DeviceArray(ctx, A.nnz, np.float64)
    check(lib().spws_csr64_transpose_dev(ctx.h, int(A.nrows), int(A.ncols), int(A.nnz), A.pos.ptr, A.crd.ptr,
                                         A.vals.ptr, pos.ptr, crd.ptr, vals.ptr))
    ctx.sync()
    return DeviceCSR(A.ncols, A.nrows, A.nnz, pos, crd, vals)


@dataclass
class DeviceCSF:
    """A 3-level CSF on the device (see formats.CSF3)."""
    nroots: int
    nfibers: int
    nnz: int
    l1: DeviceArray
    l2_pos: DeviceArray
    l2_crd: DeviceArray
    l3_pos: DeviceArray
    l3_crd: DeviceArray
    vals: DeviceArray

    @classmethod
    def upload(cls, ctx: Context, x: CSF3) -> "DeviceCSF":
        up = lambda a, t: DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=t))
        return cls(len(x.l1_crd), len(x.l2_crd), x.nnz, up(x.l1_crd, np.int32), up(x.l2_pos, np.int32),
                   up(x.l2_crd, np.int32), up(x.l3_pos, np.int32), up(x.l3_crd, np.int32), up(x.vals, np.float64))


def mttkrp(ctx: Context, X: DeviceCSF, I: int, J: int, K: int, R: int, Bf: DeviceArray, Cf: DeviceArray):
    """MTTKRP over a CSF with levels (j, k, i); returns (rows with a nonzero,
    dense M view I x R) -- both context-owned device arrays."""
    nrows = C.c_int32()
    pr, pm = C.c_void_p(), C.c_void_p()
    check(lib().spws_mttkrp(ctx.h, int(I), int(J), int(K), int(R), X.nroots, X.l1.ptr, X.l2_pos.ptr, X.l2_crd.ptr,
                            X.nfibers, X.l3_pos.ptr, X.l3_crd.ptr, X.vals.ptr, X.nnz, Bf.ptr, Cf.ptr,
                            C.byref(nrows), C.byref(pr), C.byref(pm)))
    return (DeviceArray.view(ctx, pr.value, nrows.value, np.int32),
            DeviceArray.view(ctx, pm.value, int(I) * int(R), np.float64))


def set_ttm_method(ctx: Context, method: int = 0):
    """0 automatic, 1 one radix sort of the whole tensor, 2 roots and
    column windows (see spws.h); identical results."""
    check(lib().spws_ctx_set_ttm_method(ctx.h, int(method)))


def ttm(ctx: Context, X: DeviceCSF, I: int, J: int, K: int, R: int, U: DeviceArray):
    """TTM over a CSF with levels (i, k, j); returns (fib_i, fib_j, Y) views,
    fibers sorted by (i, j), Y nfib x R."""
    nf = C.c_int64()
    pi, pj, py = C.c_void_p(), C.c_void_p(), C.c_void_p()
    check(lib().spws_ttm(ctx.h, int(I), int(J), int(K), int(R), X.nroots, X.l1.ptr, X.l2_pos.ptr, X.l2_crd.ptr,
                         X.nfibers, X.l3_pos.ptr, X.l3_crd.ptr, X.vals.ptr, X.nnz, U.ptr, C.byref(nf),
                         C.byref(pi), C.byref(pj), C.byref(py)))
    F = nf.value
    return (DeviceArray.view(ctx, pi.value, F, np.int32), DeviceArray.view(ctx, pj.value, F, np.int32),
            DeviceArray.view(ctx, py.value, F * int(R), np.float64))
