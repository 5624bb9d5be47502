"""CPU oracle for the sparse-workspace kernels -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker / CPU baseline.  The product
package ``sparsews_amd`` never imports it.

The C restatement lives in spws_oracle.c (built by oracle/Makefile into
oracle/libspws_oracle.so); each function there cites the reference code it
follows.  This module wraps it with numpy and adds the tolerance metric.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libspws_oracle.so")
_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
PI32 = C.POINTER(C.c_int32)
PF64 = C.POINTER(C.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_spgemm_hash.restype = C.c_int64
        L.orc_spgemm_hash.argtypes = [C.c_int32, C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p,
                                      C.c_int32, i32p, C.POINTER(PI32), C.POINTER(PF64)]
        L.orc_spgemm_hash_parallel.restype = C.c_int64
        L.orc_spgemm_hash_parallel.argtypes = [C.c_int32, C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p,
                                               C.c_int32, C.c_int32, i32p, C.POINTER(PI32), C.POINTER(PF64),
                                               C.POINTER(C.c_int64)]
        L.orc_spgemm_symbolic.restype = C.c_int64
        L.orc_spgemm_symbolic.argtypes = [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, C.c_int]
        L.orc_spgemm_row_counts.restype = C.c_int64
        L.orc_spgemm_row_counts.argtypes = [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, C.c_int]
        L.orc_spgemm_gustavson.restype = C.c_int64
        L.orc_spgemm_gustavson.argtypes = [C.c_int32, C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p,
                                           i32p, C.POINTER(PI32), C.POINTER(PF64), C.c_int]
        L.orc_spgemm_row_ub.restype = None
        L.orc_spgemm_row_ub.argtypes = [C.c_int32, i32p, i32p, i32p, i64p]
        L.orc_spgemm_refsum.restype = None
        L.orc_spgemm_refsum.argtypes = [C.c_int32, C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p,
                                        i32p, i32p, f64p, f64p, C.c_int]
        L.orc_outer_coord.restype = C.c_int64
        L.orc_outer_coord.argtypes = [C.c_int32, C.c_int32,
                                      C.c_int32, i32p, i32p, i32p, f64p,
                                      C.c_int32, i32p, i32p, i32p, f64p,
                                      C.c_int32, i64p, C.POINTER(PI32), C.POINTER(PF64)]
        L.orc_outer_products.restype = C.c_int64
        L.orc_outer_products.argtypes = [C.c_int32, i32p, i32p, C.c_int32, i32p, i32p]
        L.orc_mttkrp_hash.restype = C.c_int64
        L.orc_mttkrp_hash.argtypes = [C.c_int32, C.c_int32,
                                      C.c_int32, i32p, i32p, i32p, i32p, i32p, f64p,
                                      C.c_int32, i32p, i32p, i32p, f64p,
                                      C.c_int32, i32p, i32p, i32p, f64p,
                                      C.c_int32, C.POINTER(PI32), C.POINTER(PI32), C.POINTER(PF64)]
        L.orc_ttm_hash_il.restype = C.c_int64
        L.orc_ttm_hash_il.argtypes = [C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32, i32p, i32p, i32p, i32p, i32p, f64p,
                                      C.c_int32, i32p, i32p, i32p, f64p,
                                      C.POINTER(PI32), C.POINTER(PI32), C.POINTER(PI32), C.POINTER(PF64)]
        L.orc_wspace_sort.restype = C.c_int64
        L.orc_wspace_sort.argtypes = [C.c_void_p, C.c_int64, C.c_int]
        L.orc_wspace_merge.restype = C.c_int64
        L.orc_wspace_merge.argtypes = [i32p, i32p, f64p, C.c_int64, C.c_void_p, C.c_int64, C.c_int]
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_csr_transpose.restype = None
        L.orc_csr_transpose.argtypes = [C.c_int32, C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p]
        _lib = L
    return _lib


def _take(ptr, n, dtype):
    """Copy a malloc'd C array into numpy and free it."""
    if n > 0:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    else:
        out = np.zeros(0, dtype=dtype)
    lib().orc_free(C.cast(ptr, C.c_void_p))
    return out


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cf(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------- SpGEMM

def spgemm_hash(A, B, w_cap: int):
    """Row-wise bounded hash workspace (CSR_CSR_hash_sequential.h:232-353).
    Returns (pos int32[m+1], crd int32, vals f64)."""
    from sparsews_amd.gen import CSR
    pos = np.zeros(A.nrows + 1, dtype=np.int32)
    ci, cx = PI32(), PF64()
    nnz = lib().orc_spgemm_hash(A.nrows, B.ncols, _c32(A.pos), _c32(A.crd), _cf(A.vals),
                                _c32(B.pos), _c32(B.crd), _cf(B.vals), int(w_cap), pos,
                                C.byref(ci), C.byref(cx))
    return CSR(A.nrows, B.ncols, pos, _take(ci, nnz, np.int32), _take(cx, nnz, np.float64))


def spgemm_hash_parallel(A, B, w_cap: int, cap: int):
    """CAP-parallel hash workspace (CSR_CSR_hash_parallel.h:235-408): `cap`
    workers over rows i = t mod cap and one merger.  Returns (CSR, number of
    mid-run flushes)."""
    from sparsews_amd.gen import CSR
    pos = np.zeros(A.nrows + 1, dtype=np.int32)
    ci, cx = PI32(), PF64()
    fl = C.c_int64()
    nnz = lib().orc_spgemm_hash_parallel(A.nrows, B.ncols, _c32(A.pos), _c32(A.crd), _cf(A.vals),
                                         _c32(B.pos), _c32(B.crd), _cf(B.vals), int(w_cap), int(cap), pos,
                                         C.byref(ci), C.byref(cx), C.byref(fl))
    return CSR(A.nrows, B.ncols, pos, _take(ci, nnz, np.int32), _take(cx, nnz, np.float64)), int(fl.value)


def spgemm_gustavson(A, B, nthreads: int = 0):
    """OpenMP TACO Gustavson (CSR_CSR_taco_parallel.h:4-113)."""
    from sparsews_amd.gen import CSR
    pos = np.zeros(A.nrows + 1, dtype=np.int32)
    ci, cx = PI32(), PF64()
    nnz = lib().orc_spgemm_gustavson(A.nrows, B.ncols, _c32(A.pos), _c32(A.crd), _cf(A.vals),
                                     _c32(B.pos), _c32(B.crd), _cf(B.vals), pos,
                                     C.byref(ci), C.byref(cx), int(nthreads))
    if nnz < 0:
        raise OverflowError("product has more than 2^31 - 1 entries (int32 CSR)")
    return CSR(A.nrows, B.ncols, pos, _take(ci, nnz, np.int32), _take(cx, nnz, np.float64))


def spgemm_symbolic(A, B, nthreads: int = 0):
    pos = np.zeros(A.nrows + 1, dtype=np.int32)
    nnz = lib().orc_spgemm_symbolic(A.nrows, B.ncols, _c32(A.pos), _c32(A.crd), _c32(B.pos),
                                    _c32(B.crd), pos, int(nthreads))
    if nnz < 0:
        raise OverflowError("product has more than 2^31 - 1 entries (int32 CSR)")
    return pos


def spgemm_row_counts(A, B, nthreads: int = 0):
    """Entries per output row of A*B (int32 each) and their int64 total; no
    int32 limit on the total."""
    cnt = np.zeros(A.nrows, dtype=np.int32)
    total = lib().orc_spgemm_row_counts(A.nrows, B.ncols, _c32(A.pos), _c32(A.crd), _c32(B.pos), _c32(B.crd),
                                        cnt, int(nthreads))
    return cnt, int(total)


def spgemm_row_ub(A, B):
    ub = np.zeros(A.nrows, dtype=np.int64)
    lib().orc_spgemm_row_ub(A.nrows, _c32(A.pos), _c32(A.crd), _c32(B.pos), ub)
    return ub


def csr_transpose(A):
    """A^T as a CSR, each row's columns increasing (transpose() of
    benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:352-400)."""
    from sparsews_amd.gen import CSR
    nnz = int(A.pos[-1])
    pos = np.zeros(A.ncols + 1, dtype=np.int32)
    crd = np.zeros(nnz, dtype=np.int32)
    vals = np.zeros(nnz, dtype=np.float64)
    lib().orc_csr_transpose(A.nrows, A.ncols, _c32(A.pos), _c32(A.crd), _cf(A.vals), pos, crd, vals)
    return CSR(A.ncols, A.nrows, pos, crd, vals)


def spgemm_refsum(A, B, Cpos, Ccrd, nthreads: int = 0):
    """Long-double sums and sum|terms| for every entry of the pattern (Cpos, Ccrd)."""
    n = len(Ccrd)
    exact = np.zeros(n, dtype=np.float64)
    absum = np.zeros(n, dtype=np.float64)
    lib().orc_spgemm_refsum(A.nrows, B.ncols, _c32(A.pos), _c32(A.crd), _cf(A.vals),
                            _c32(B.pos), _c32(B.crd), _cf(B.vals), _c32(Cpos), _c32(Ccrd),
                            exact, absum, int(nthreads))
    return exact, absum


def outer_coord(Adcsc, Bdcsr, m: int, n: int, w_cap: int):
    """Outer-product coordinate workspace (DCSC_DCSR_coord_index_chase.h:194-311).
    Returns (pos int64[m+1], crd int32, vals f64)."""
    pos = np.zeros(m + 1, dtype=np.int64)
    ci, cx = PI32(), PF64()
    nnz = lib().orc_outer_coord(int(m), int(n),
                                len(Adcsc.l1_crd), _c32(Adcsc.l1_crd), _c32(Adcsc.l2_pos),
                                _c32(Adcsc.l2_crd), _cf(Adcsc.vals),
                                len(Bdcsr.l1_crd), _c32(Bdcsr.l1_crd), _c32(Bdcsr.l2_pos),
                                _c32(Bdcsr.l2_crd), _cf(Bdcsr.vals),
                                int(w_cap), pos, C.byref(ci), C.byref(cx))
    return pos, _take(ci, nnz, np.int32), _take(cx, nnz, np.float64)


def outer_products(Adcsc, Bdcsr) -> int:
    return int(lib().orc_outer_products(len(Adcsc.l1_crd), _c32(Adcsc.l1_crd), _c32(Adcsc.l2_pos),
                                        len(Bdcsr.l1_crd), _c32(Bdcsr.l1_crd), _c32(Bdcsr.l2_pos)))


def mttkrp_hash(Bcsf, Cd, Dd, I: int, R: int, w_cap: int):
    """MTTKRP over a CSF with levels (k, l, i) (COO_CSF_DCSR_DCSR_hash_free.h:233-366).
    Returns COO sorted by (i, r): (i int32, r int32, vals f64)."""
    oi, orr, ov = PI32(), PI32(), PF64()
    nnz = lib().orc_mttkrp_hash(int(I), int(R),
                                len(Bcsf.l1_crd), _c32(Bcsf.l1_crd), _c32(Bcsf.l2_pos),
                                _c32(Bcsf.l2_crd), _c32(Bcsf.l3_pos), _c32(Bcsf.l3_crd),
                                _cf(Bcsf.vals),
                                len(Cd.l1_crd), _c32(Cd.l1_crd), _c32(Cd.l2_pos), _c32(Cd.l2_crd),
                                _cf(Cd.vals),
                                len(Dd.l1_crd), _c32(Dd.l1_crd), _c32(Dd.l2_pos), _c32(Dd.l2_crd),
                                _cf(Dd.vals),
                                int(w_cap), C.byref(oi), C.byref(orr), C.byref(ov))
    return _take(oi, nnz, np.int32), _take(orr, nnz, np.int32), _take(ov, nnz, np.float64)


def ttm_hash_il(Bcsf, Cd, K: int, J: int, L: int):
    """TTM over a CSF with levels (k, i, j), i contracted (COO_CSF_DCSR_hash_il.h:247-377).
    Returns COO sorted by (k, j, l)."""
    ok, oj, ol, ov = PI32(), PI32(), PI32(), PF64()
    nnz = lib().orc_ttm_hash_il(int(K), int(J), int(L),
                                len(Bcsf.l1_crd), _c32(Bcsf.l1_crd), _c32(Bcsf.l2_pos),
                                _c32(Bcsf.l2_crd), _c32(Bcsf.l3_pos), _c32(Bcsf.l3_crd),
                                _cf(Bcsf.vals),
                                len(Cd.l1_crd), _c32(Cd.l1_crd), _c32(Cd.l2_pos), _c32(Cd.l2_crd),
                                _cf(Cd.vals),
                                C.byref(ok), C.byref(oj), C.byref(ol), C.byref(ov))
    return (_take(ok, nnz, np.int32), _take(oj, nnz, np.int32), _take(ol, nnz, np.int32),
            _take(ov, nnz, np.float64))


# --------------------------------------------------------------------------- tolerance

REL_TOL = 1e-12


# --------------------------------------------------------------------------- workspace API

WSPACE_DTYPE = np.dtype([("crd", np.int32, (2,)), ("val", np.float64)])


def wspace_sort(records, rev: bool):
    """Sort of utils/wspace.h:206-217 on a copy; returns (sorted, count)."""
    w = np.array(records, dtype=WSPACE_DTYPE, copy=True)
    n = lib().orc_wspace_sort(w.ctypes.data_as(C.c_void_p), w.size, int(bool(rev)))
    return w, int(n)


def wspace_merge(coo_crd0, coo_crd1, coo_vals, records, rev: bool = False):
    """Merge of utils/wspace.h:62-204; returns the merged (crd0, crd1, vals)."""
    acc = np.ascontiguousarray(records, dtype=WSPACE_DTYPE)
    n0 = len(coo_crd0)
    cap = max(n0 + acc.size, 1)
    c0 = np.zeros(cap, np.int32)
    c1 = np.zeros(cap, np.int32)
    v = np.zeros(cap, np.float64)
    c0[:n0], c1[:n0], v[:n0] = coo_crd0, coo_crd1, coo_vals
    n = lib().orc_wspace_merge(c0, c1, v, n0, acc.ctypes.data_as(C.c_void_p), acc.size, int(bool(rev)))
    return c0[:n].copy(), c1[:n].copy(), v[:n].copy()


def value_error(x, y, abssum):
    """Per-entry error |x - y| / max(|y|, sum|terms|); the bar is <= 1e-12."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    denom = np.maximum(np.abs(y), np.asarray(abssum, dtype=np.float64))
    denom = np.where(denom == 0, 1.0, denom)
    return np.abs(x - y) / denom


def max_value_error(x, y, abssum) -> float:
    if len(x) == 0:
        return 0.0
    return float(value_error(x, y, abssum).max())
