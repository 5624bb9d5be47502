"""ctypes binding of libspws.so (the C ABI declared in include/spws/spws.h).

The library is loaded from this package directory.  There is no fallback:
if libspws.so is missing, ``lib()`` raises, so nothing here can silently run
on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspws.so")

SPWS_OK = 0
STATUS_NAMES = {0: "SPWS_OK", 1: "SPWS_ERR_INVALID", 2: "SPWS_ERR_OOM", 3: "SPWS_ERR_DEVICE",
                4: "SPWS_ERR_UNSUPPORTED"}


class SpwsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class TacoTensor(C.Structure):
    """taco_tensor_t (include/spws/taco_tensor.h; utils/lib.h:47-56 with FP64 vals)."""
    _fields_ = [
        ("order", C.c_int32),
        ("dimensions", C.POINTER(C.c_int32)),
        ("csize", C.c_int32),
        ("mode_ordering", C.POINTER(C.c_int32)),
        ("mode_types", C.POINTER(C.c_int)),
        ("indices", C.POINTER(C.POINTER(C.POINTER(C.c_int32)))),
        ("vals", C.POINTER(C.c_double)),
        ("vals_size", C.c_int32),
    ]


class RowwiseTuning(C.Structure):
    _fields_ = [("wave_ub", C.c_int32), ("wave2_ub", C.c_int32), ("block_ub", C.c_int32),
                ("block2_ub", C.c_int32), ("dense_ratio", C.c_int32), ("force_path", C.c_int32)]


class BatchCaps(C.Structure):
    _fields_ = [("max_requests", C.c_int32), ("max_rows", C.c_int32), ("max_brows", C.c_int32),
                ("max_cols", C.c_int32), ("max_annz", C.c_int64), ("max_bnnz", C.c_int64),
                ("max_products", C.c_int64), ("max_row_products", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("m", C.c_int32), ("k", C.c_int32), ("n", C.c_int32), ("reserved", C.c_int32),
                ("A_pos", C.c_void_p), ("A_crd", C.c_void_p), ("A_vals", C.c_void_p), ("annz", C.c_int64),
                ("B_pos", C.c_void_p), ("B_crd", C.c_void_p), ("B_vals", C.c_void_p), ("bnnz", C.c_int64),
                ("products", C.c_int64), ("max_row_products", C.c_int64), ("arrival", C.c_double),
                ("C_pos", C.c_void_p), ("C_crd", C.c_void_p), ("C_vals", C.c_void_p), ("C_cap", C.c_int64)]


class ServeOptions(C.Structure):
    _fields_ = [("mode", C.c_int32), ("max_batch_requests", C.c_int32), ("max_batch_products", C.c_int64),
                ("large_products", C.c_int64), ("large_engines", C.c_int32), ("reserved", C.c_int32)]


class ServeStats(C.Structure):
    _fields_ = [("wall_seconds", C.c_double), ("batches", C.c_int64), ("fallback_requests", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("products", C.c_int64), ("nnz", C.c_int64), ("rows_per_path", C.c_int64 * 8),
                ("chunks", C.c_int64), ("sort_bits", C.c_int64), ("outer_method", C.c_int64),
                ("window_bits", C.c_int64)]


VP = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
PVP = C.POINTER(C.c_void_p)
PI64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every entry is an extern "C" symbol of spws.h
SIGNATURES = {
    "spws_last_error": (C.c_char_p, []),
    "spws_ctx_create": (C.c_int, [C.c_int, VP, C.POINTER(VP)]),
    "spws_ctx_destroy": (C.c_int, [VP]),
    "spws_ctx_set_stream": (C.c_int, [VP, VP]),
    "spws_ctx_get_stream": (VP, [VP]),
    "spws_ctx_get_stats": (C.c_int, [VP, C.POINTER(Stats)]),
    "spws_ctx_set_rowwise_tuning": (C.c_int, [VP, C.POINTER(RowwiseTuning)]),
    "spws_ctx_get_rowwise_tuning": (C.c_int, [VP, C.POINTER(RowwiseTuning)]),
    "spws_ctx_set_profiling": (C.c_int, [VP, C.c_int]),
    "spws_ctx_set_profiling_only": (C.c_int, [VP, C.c_char_p]),
    "spws_ctx_reset_kernel_times": (C.c_int, [VP]),
    "spws_ctx_kernel_times": (C.c_int, [VP, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_double), PI64]),
    "spws_dev_alloc": (C.c_int, [VP, C.c_size_t, C.POINTER(VP)]),
    "spws_dev_free": (C.c_int, [VP, VP]),
    "spws_memcpy_h2d": (C.c_int, [VP, VP, VP, C.c_size_t]),
    "spws_memcpy_d2h": (C.c_int, [VP, VP, VP, C.c_size_t]),
    "spws_sync": (C.c_int, [VP]),
    "spws_default_ctx": (C.c_int, [C.POINTER(VP)]),
    "spws_host_free": (None, [VP]),
    "spws_device_copy": (C.c_int, [VP, VP, VP, C.c_size_t]),
    "spws_spgemm_rowwise": (C.c_int, [VP, I32, I32, I32, VP, VP, VP, VP, VP, VP,
                                      PI64, PVP, PVP, PVP]),
    "spws_spgemm_rowwise_symbolic": (C.c_int, [VP, I32, I32, I32, VP, VP, VP, VP, PI64, PVP]),
    "spws_spgemm_outer": (C.c_int, [VP, I32, I32, I32, I32, VP, VP, VP, VP, I32, VP, VP, VP, VP,
                                    PI64, PVP, PVP, PVP]),
    "spws_ctx_set_outer_chunking": (C.c_int, [VP, I64, I32]),
    "spws_ctx_set_outer_method": (C.c_int, [VP, I32, I32]),
    "spws_csr_transpose_dev": (C.c_int, [VP, I32, I32, I64, VP, VP, VP, VP, VP, VP]),
    "spws_csr_transpose_dev_noalloc": (C.c_int, [VP, I32, I32, I64, VP, VP, VP, VP, VP, VP]),
    "spws_csr64_transpose_dev": (C.c_int, [VP, I32, I32, I64, VP, VP, VP, VP, VP, VP]),
    "spws_ctx_set_ttm_method": (C.c_int, [VP, I32]),
    "spws_mttkrp": (C.c_int, [VP, I32, I32, I32, I32, I32, VP, VP, VP, I32, VP, VP, VP, I64, VP, VP,
                              C.POINTER(I32), PVP, PVP]),
    "spws_ttm": (C.c_int, [VP, I32, I32, I32, I32, I32, VP, VP, VP, I32, VP, VP, VP, I64, VP,
                           PI64, PVP, PVP, PVP]),
    "spws_mttkrp_masked": (C.c_int, [VP, I32, I32, I32, I32, I32, VP, VP, VP, I32, VP, VP, VP, I64, VP, VP,
                                     VP, VP, C.POINTER(I32), PVP, PVP, PVP]),
    "spws_ttm_masked": (C.c_int, [VP, I32, I32, I32, I32, I32, VP, VP, VP, I32, VP, VP, VP, I64, VP, VP,
                                  PI64, PVP, PVP, PVP, PVP]),
    "spws_wspace_sort": (C.c_int, [VP, VP, I64, C.c_int, PI64]),
    "spws_wspace_merge": (C.c_int, [VP, VP, VP, VP, I64, VP, I64, PI64]),
    "spws_batch_create": (C.c_int, [VP, C.POINTER(BatchCaps), C.POINTER(VP)]),
    "spws_batch_destroy": (C.c_int, [VP]),
    "spws_batch_reset": (C.c_int, [VP]),
    "spws_batch_add": (C.c_int, [VP, I32, I32, I32, VP, VP, VP, I64, VP, VP, VP, I64, I64, I64,
                                 C.POINTER(I32)]),
    "spws_batch_launch": (C.c_int, [VP]),
    "spws_batch_poll": (C.c_int, [VP, C.POINTER(C.c_int)]),
    "spws_batch_wait": (C.c_int, [VP]),
    "spws_batch_nnz": (C.c_int, [VP, PI64]),
    "spws_batch_result": (C.c_int, [VP, I32, PI64, VP, PVP, PVP]),
    "spws_batch_set_output": (C.c_int, [VP, I32, VP, VP, VP, I64]),
    "spws_batch_result_nnz": (C.c_int, [VP, I32, PI64]),
    "spws_batch_stream": (VP, [VP]),
    "spws_serve": (C.c_int, [VP, VP, I32, I32, C.POINTER(Request), C.POINTER(ServeOptions), VP, VP, VP,
                             C.POINTER(ServeStats)]),
    "spws_csr_csr_hash_compute": (C.c_int, [C.POINTER(TacoTensor)] * 3 + [I32]),
    "spws_csr_csr_assemble": (C.c_int, [C.POINTER(TacoTensor)] * 3),
    "spws_csr_csr_compute": (C.c_int, [C.POINTER(TacoTensor)] * 3),
    "spws_dcsc_dcsr_coord_compute": (C.c_int, [C.POINTER(TacoTensor)] * 3 + [I32]),
    "spws_csc_csr_t_hash_compute": (C.c_int, [C.POINTER(TacoTensor)] * 3 + [I32]),
    "spws_csc_csr_nofuse_compute": (C.c_int, [C.POINTER(TacoTensor)] * 3 + [I32]),
    "spws_csr_transpose": (C.c_int, [C.POINTER(TacoTensor)] * 2),
    "spws_mttkrp_hash_compute": (C.c_int, [C.POINTER(TacoTensor)] * 4 + [I32]),
    "spws_ttm_hash_il_compute": (C.c_int, [C.POINTER(TacoTensor)] * 3 + [I32]),
}

_lib = None


def header_symbols() -> list[str]:
    """Function names declared in include/spws/*.h (for the export check)."""
    import re
    inc = os.path.join(os.path.dirname(HERE), "include", "spws")
    names = []
    for f in sorted(os.listdir(inc)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(spws_\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m sparsews_amd.build` "
                               "(the HIP library is required, there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != SPWS_OK:
        msg = lib().spws_last_error()
        raise SpwsError(rc, msg.decode() if msg else "")


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data
