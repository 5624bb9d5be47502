"""Host mirror of the reference's TACO interface, over the C ABI.

* ``DC_to_taco_tensor`` / ``init_taco_tensor_DC`` build taco_tensor_t
  structs the way utils/dataloader.h:374-442 does (numpy arrays are kept
  alive by the returned handle).
* ``compute`` / ``assemble`` call the TACO-shaped entry points of libspws:
  the same argument order and meaning as the reference's generated code
  (output first), the same ownership (outputs malloc'd by the callee,
  released with ``free_outputs``), and a nonzero status raises.
* ``CSR_CSR_hash`` is the reference's timing wrapper
  (benchmark/Parallel/CSR_CSR_hash_sequential.h:355-382), wall-clock timed.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import TacoTensor, check, lib
from .gen import CSR

_P32 = C.POINTER(C.c_int32)


class Taco:
    """A taco_tensor_t plus the Python objects that own its arrays."""

    def __init__(self, order: int, dims):
        self.t = TacoTensor()
        self._keep = []
        self.t.order = order
        self._dims = np.ascontiguousarray(dims, dtype=np.int32)
        self._ord = np.arange(order, dtype=np.int32)
        self.t.dimensions = self._dims.ctypes.data_as(_P32)
        self.t.mode_ordering = self._ord.ctypes.data_as(_P32)
        self._levels = (C.POINTER(_P32) * (order + 1))()
        self._level_arrays = [(_P32 * 2)() for _ in range(order + 1)]
        for l in range(order + 1):
            self._levels[l] = C.cast(self._level_arrays[l], C.POINTER(_P32))
        self.t.indices = C.cast(self._levels, C.POINTER(C.POINTER(_P32)))
        self.owned_by_lib = False

    def set_index(self, level: int, which: int, arr: np.ndarray):
        arr = np.ascontiguousarray(arr, dtype=np.int32)
        self._keep.append(arr)
        self._level_arrays[level][which] = arr.ctypes.data_as(_P32)

    def set_vals(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._keep.append(arr)
        self.t.vals = arr.ctypes.data_as(C.POINTER(C.c_double))
        self.t.vals_size = len(arr)

    def index(self, level: int, which: int, count: int) -> np.ndarray:
        p = self._level_arrays[level][which]
        if count == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(count,)).copy()

    def vals(self, count: int) -> np.ndarray:
        if count == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(self.t.vals, shape=(count,)).copy()

    @property
    def ref(self):
        return C.byref(self.t)


def DC_to_taco_tensor(a: CSR) -> Taco:
    """CSR ("DC") taco tensor: indices[0][0] = {nrows}, indices[1] = (pos, crd)."""
    t = Taco(2, [a.nrows, a.ncols])
    t.set_index(0, 0, np.array([a.nrows], dtype=np.int32))
    t.set_index(1, 0, a.pos)
    t.set_index(1, 1, a.crd)
    t.set_vals(a.vals)
    return t


def DC_to_taco_tensor_csc(a: CSR) -> Taco:
    """CSC taco tensor of the matrix whose transpose is ``a`` (``a`` holds the
    columns as rows): dimensions {a.ncols, a.nrows}, mode_ordering {1, 0},
    indices[1] = (column pointers, row indices)."""
    t = Taco(2, [a.ncols, a.nrows])
    t._ord[:] = [1, 0]
    t.set_index(0, 0, np.array([a.nrows], dtype=np.int32))
    t.set_index(1, 0, a.pos)
    t.set_index(1, 1, a.crd)
    t.set_vals(a.vals)
    return t


def init_taco_tensor_DC(nrow: int, ncol: int) -> Taco:
    t = Taco(2, [nrow, ncol])
    t.set_index(0, 0, np.array([nrow], dtype=np.int32))
    return t


def taco_to_csr(t: Taco) -> CSR:
    m = int(t._dims[0])
    pos = t.index(1, 0, m + 1)
    nnz = int(pos[-1]) if m >= 0 else 0
    return CSR(m, int(t._dims[1]), pos, t.index(1, 1, nnz), t.vals(nnz))


def free_outputs(t: Taco):
    """free() the arrays a compute()/assemble() call malloc'd into ``t``."""
    L = lib()
    for which in (0, 1):
        p = t._level_arrays[1][which]
        if p:
            L.spws_host_free(C.cast(p, C.c_void_p))
            t._level_arrays[1][which] = _P32()
    if t.t.vals:
        L.spws_host_free(C.cast(t.t.vals, C.c_void_p))
        t.t.vals = C.POINTER(C.c_double)()


def compute(Ct: Taco, At: Taco, Bt: Taco, w_cap: int) -> int:
    """int compute(C, A, B, w_cap): row-wise hash SpGEMM (CSR_CSR_hash_sequential.h:232)."""
    check(lib().spws_csr_csr_hash_compute(Ct.ref, At.ref, Bt.ref, int(w_cap)))
    return 0


def assemble(Ct: Taco, At: Taco, Bt: Taco) -> int:
    """int assemble(C, A, B): symbolic phase (seperate/CSR_CSR_1.h:69-167)."""
    check(lib().spws_csr_csr_assemble(Ct.ref, At.ref, Bt.ref))
    return 0


def compute_split(Ct: Taco, At: Taco, Bt: Taco) -> int:
    """int compute(C, A, B): numeric phase after assemble (seperate/CSR_CSR_1.h:4-67)."""
    check(lib().spws_csr_csr_compute(Ct.ref, At.ref, Bt.ref))
    return 0


def CSR_CSR_hash(At: Taco, Bt: Taco, Ct: Taco, w_cap: int, warmup: int, repeat: int,
                 bench: bool = False, print_: bool = False) -> float:
    """Timing wrapper: warmup + repeat calls of compute(); seconds per call."""
    for _ in range(warmup):
        compute(Ct, At, Bt, w_cap)
        if bench:
            free_outputs(Ct)
    t0 = time.perf_counter()
    for i in range(repeat):
        compute(Ct, At, Bt, w_cap)
        if bench and i != repeat - 1:
            free_outputs(Ct)
    dt = (time.perf_counter() - t0) / max(repeat, 1)
    if print_:
        print(taco_to_csr(Ct))
    return dt


# --------------------------------------------------------------------------- CC, CSF and COO tensors

def CC_to_taco_tensor(d, nrows: int, ncols: int) -> Taco:
    """DCSR / DCSC taco tensor (CC_to_taco_tensor, utils/dataloader.h:504-569)."""
    t = Taco(2, [nrows, ncols])
    t.set_index(0, 0, np.array([0, len(d.l1_crd)], dtype=np.int32))
    t.set_index(0, 1, d.l1_crd)
    t.set_index(1, 0, d.l2_pos)
    t.set_index(1, 1, d.l2_crd)
    t.set_vals(d.vals)
    return t


def csf_to_taco_tensor(x) -> Taco:
    """3-level CSF taco tensor: level l holds (pos, crd); the root pos is {0, n}."""
    t = Taco(3, [x.dims[o] for o in x.order])
    t.set_index(0, 0, np.array([0, len(x.l1_crd)], dtype=np.int32))
    t.set_index(0, 1, x.l1_crd)
    t.set_index(1, 0, x.l2_pos)
    t.set_index(1, 1, x.l2_crd)
    t.set_index(2, 0, x.l3_pos)
    t.set_index(2, 1, x.l3_crd)
    t.set_vals(x.vals)
    return t


def init_taco_tensor_COO(dims) -> Taco:
    """COO output tensor (init_taco_tensor_COO, utils/dataloader.h:964-982)."""
    t = Taco(len(dims), dims)
    t.set_index(0, 0, np.array([dims[0]], dtype=np.int32))
    return t


def taco_coo(t: Taco) -> tuple:
    """(coordinate arrays..., vals) of a COO output; frees the callee's arrays."""
    n = int(t.t.vals_size)
    L = lib()
    out = []
    for lvl in range(1, int(t.t.order) + 1):
        out.append(t.index(lvl, 0, n))
        p = t._level_arrays[lvl][0]
        if p:
            L.spws_host_free(C.cast(p, C.c_void_p))
            t._level_arrays[lvl][0] = _P32()
    out.append(t.vals(n))
    if t.t.vals:
        L.spws_host_free(C.cast(t.t.vals, C.c_void_p))
        t.t.vals = C.POINTER(C.c_double)()
    return tuple(out)


def dcsc_dcsr_compute(Ct: Taco, At: Taco, Bt: Taco, w_cap: int) -> int:
    """Outer-product coordinate workspace (DCSC_DCSR_coord_index_chase.h:194)."""
    check(lib().spws_dcsc_dcsr_coord_compute(Ct.ref, At.ref, Bt.ref, int(w_cap)))
    return 0


def csc_csr_t_compute(Ct: Taco, At: Taco, Bt: Taco, w_cap: int) -> int:
    """Transposed-output product C = (A B)^T, A CSC, B CSR
    (CSC_CSR_T_hash.h:232)."""
    check(lib().spws_csc_csr_t_hash_compute(Ct.ref, At.ref, Bt.ref, int(w_cap)))
    return 0


def csc_csr_nofuse_compute(Ct: Taco, At: Taco, Bt: Taco, w_cap: int) -> int:
    """C = A B without the transposition, A CSC, B CSR
    (nonfuse/CSC_CSR_T_hash.h:232, compute_nofuse)."""
    check(lib().spws_csc_csr_nofuse_compute(Ct.ref, At.ref, Bt.ref, int(w_cap)))
    return 0


def transpose(At: Taco, Bt: Taco) -> int:
    """A = B^T for CSR B (nonfuse/CSC_CSR_T_hash.h:352); A receives malloc'd
    arrays (free_outputs)."""
    check(lib().spws_csr_transpose(At.ref, Bt.ref))
    return 0


def mttkrp_compute(At: Taco, Bt: Taco, Ct: Taco, Dt: Taco, w_cap: int) -> int:
    """MTTKRP hash workspace (COO_CSF_DCSR_DCSR_hash_free.h:233)."""
    check(lib().spws_mttkrp_hash_compute(At.ref, Bt.ref, Ct.ref, Dt.ref, int(w_cap)))
    return 0


def ttm_compute(At: Taco, Bt: Taco, Ct: Taco, w_cap: int) -> int:
    """TTM hash_il workspace (COO_CSF_DCSR_hash_il.h:247)."""
    check(lib().spws_ttm_hash_il_compute(At.ref, Bt.ref, Ct.ref, int(w_cap)))
    return 0
