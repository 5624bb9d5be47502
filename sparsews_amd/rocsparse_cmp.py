"""rocSPARSE SpGEMM (the ROCm library installed with the toolchain) as a
comparator for the row-wise product: the same C = A * B, FP64, CSR with
int32 indices, on the context's stream and device memory.  It gives the
GFLOP/s of the vendor library beside bench.py's row-wise line and a
structural cross-check of our output on inputs too large for the CPU oracle.

Only the generic API is used (rocsparse_spgemm with the buffer-size, nnz and
compute stages; rocsparse/internal/generic/rocsparse_spgemm.h).  C = 1 * A * B
+ 0 * D with D an empty matrix of C's shape.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .device import Context, DeviceCSR

_OP_NONE = 111
_INDEX_I32 = 2
_BASE_ZERO = 0
_F64 = 152
_ALG_DEFAULT = 0
_STAGE_BUFFER, _STAGE_NNZ, _STAGE_COMPUTE = 1, 2, 3
_POINTER_HOST = 0

_lib = None


def available() -> bool:
    try:
        _load()
        return True
    except OSError:
        return False


def _load():
    global _lib
    if _lib is None:
        root = os.environ.get("ROCM_PATH", "/opt/rocm")
        lib = C.CDLL(os.path.join(root, "lib", "librocsparse.so"))
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
        lib.rocsparse_create_handle.argtypes = [C.POINTER(vp)]
        lib.rocsparse_destroy_handle.argtypes = [vp]
        lib.rocsparse_set_stream.argtypes = [vp, vp]
        lib.rocsparse_set_pointer_mode.argtypes = [vp, i32]
        lib.rocsparse_create_csr_descr.argtypes = [C.POINTER(vp), i64, i64, i64, vp, vp, vp, i32, i32, i32, i32]
        lib.rocsparse_destroy_spmat_descr.argtypes = [vp]
        lib.rocsparse_csr_set_pointers.argtypes = [vp, vp, vp, vp]
        lib.rocsparse_spmat_get_size.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
        lib.rocsparse_spgemm.argtypes = [vp, i32, i32, vp, vp, vp, vp, vp, vp, i32, i32, i32, C.POINTER(C.c_size_t),
                                         vp]
        for name in ("rocsparse_create_handle", "rocsparse_destroy_handle", "rocsparse_set_stream",
                     "rocsparse_set_pointer_mode", "rocsparse_create_csr_descr", "rocsparse_destroy_spmat_descr",
                     "rocsparse_csr_set_pointers", "rocsparse_spmat_get_size", "rocsparse_spgemm"):
            getattr(lib, name).restype = C.c_int
        _lib = lib
    return _lib


def _check(status: int, what: str):
    if status != 0:
        raise RuntimeError(f"{what} failed with rocsparse status {status}")


class RocsparseSpgemm:
    """C = A * B with rocSPARSE for fixed device operands; ``run()`` repeats
    the nnz and compute stages (the product proper), reusing the buffers."""

    def __init__(self, ctx: Context, A: DeviceCSR, B: DeviceCSR):
        lib = _load()
        self.ctx, self.A, self.B = ctx, A, B
        self.m, self.n = A.nrows, B.ncols
        self.h = C.c_void_p()
        _check(lib.rocsparse_create_handle(C.byref(self.h)), "rocsparse_create_handle")
        _check(lib.rocsparse_set_stream(self.h, C.c_void_p(ctx.stream)), "rocsparse_set_stream")
        _check(lib.rocsparse_set_pointer_mode(self.h, _POINTER_HOST), "rocsparse_set_pointer_mode")
        self._owned = []
        self.dA = self._descr(A.nrows, A.ncols, A.nnz, A.pos.ptr, A.crd.ptr, A.vals.ptr)
        self.dB = self._descr(B.nrows, B.ncols, B.nnz, B.pos.ptr, B.crd.ptr, B.vals.ptr)
        self.d_pos = self._alloc((self.m + 1) * 4)
        zero_pos = self._alloc((self.m + 1) * 4)
        ctx.h2d(zero_pos, np.zeros(self.m + 1, dtype=np.int32))
        self.dD = self._descr(self.m, self.n, 0, zero_pos, None, None)
        self.dC = self._descr(self.m, self.n, 0, self.d_pos, None, None)
        self.alpha, self.beta = C.c_double(1.0), C.c_double(0.0)
        size = C.c_size_t(0)
        self._stage(_STAGE_BUFFER, C.byref(size), None)
        self.buf = self._alloc(max(int(size.value), 256))
        self.cap = 0
        self.crd = self.vals = None
        self.nnz = 0

    def _alloc(self, nbytes: int) -> int:
        p = self.ctx.alloc(nbytes)
        self._owned.append(p)
        return p

    def _descr(self, rows, cols, nnz, pos, crd, vals):
        d = C.c_void_p()
        _check(_load().rocsparse_create_csr_descr(C.byref(d), rows, cols, nnz, C.c_void_p(pos), C.c_void_p(crd),
                                                  C.c_void_p(vals), _INDEX_I32, _INDEX_I32, _BASE_ZERO, _F64),
               "rocsparse_create_csr_descr")
        return d

    def _stage(self, stage, size_ref, buf):
        size = size_ref if size_ref is not None else C.byref(C.c_size_t(0))
        _check(_load().rocsparse_spgemm(self.h, _OP_NONE, _OP_NONE, C.byref(self.alpha), self.dA, self.dB,
                                        C.byref(self.beta), self.dD, self.dC, _F64, _ALG_DEFAULT, stage, size,
                                        C.c_void_p(buf)), f"rocsparse_spgemm stage {stage}")

    def run(self) -> int:
        """One product; returns nnz(C).  The host reads C's size between the
        nnz and compute stages (as rocSPARSE's own usage does)."""
        lib = _load()
        self._stage(_STAGE_NNZ, None, self.buf)
        r, c, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib.rocsparse_spmat_get_size(self.dC, C.byref(r), C.byref(c), C.byref(nnz)), "rocsparse_spmat_get_size")
        self.nnz = int(nnz.value)
        if self.nnz > self.cap:
            self.cap = self.nnz
            self.crd = self._alloc(max(self.cap, 1) * 4)
            self.vals = self._alloc(max(self.cap, 1) * 8)
        _check(lib.rocsparse_csr_set_pointers(self.dC, C.c_void_p(self.d_pos), C.c_void_p(self.crd),
                                              C.c_void_p(self.vals)), "rocsparse_csr_set_pointers")
        self._stage(_STAGE_COMPUTE, None, self.buf)
        return self.nnz

    def result(self):
        """(pos, crd, vals) of the last product, on the host."""
        self.ctx.sync()
        pos = np.empty(self.m + 1, dtype=np.int32)
        crd = np.empty(self.nnz, dtype=np.int32)
        vals = np.empty(self.nnz, dtype=np.float64)
        self.ctx.d2h(pos, self.d_pos)
        if self.nnz:
            self.ctx.d2h(crd, self.crd)
            self.ctx.d2h(vals, self.vals)
        return pos, crd, vals

    def close(self):
        lib = _load()
        self.ctx.sync()
        for d in (self.dA, self.dB, self.dD, self.dC):
            lib.rocsparse_destroy_spmat_descr(d)
        lib.rocsparse_destroy_handle(self.h)
        for p in self._owned:
            self.ctx.free(p)
        self._owned = []
