"""The coordinate-workspace API of utils/wspace.h over numpy records, run
on the GPU through libspws.

A workspace is a numpy array of ``WSPACE_DTYPE`` records -- the reference's
``struct wspace {int32_t crd[2]; float val;}`` (utils/wspace.h:10-13) with
an FP64 value.  ``init_wspace`` marks slots empty (crd[0] = -1), ``Sort``
sorts records by (crd[0], crd[1]) and ``Merge`` merges a sorted workspace
into a sorted COO, adding the values of equal coordinates.  Both run on the
device (packed-key radix sort, run-head scan, segmented sum); the arrays are
host memory and are staged over PCIe by the call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

WSPACE_DTYPE = np.dtype([("crd", np.int32, (2,)), ("val", np.float64)])
assert WSPACE_DTYPE.itemsize == 16


def _handle(ctx):
    if ctx is not None:
        return ctx.h
    h = C.c_void_p()
    check(lib().spws_default_ctx(C.byref(h)))
    return h


def empty_wspace(capacity: int) -> np.ndarray:
    """A workspace of ``capacity`` empty slots."""
    w = np.zeros(int(capacity), dtype=WSPACE_DTYPE)
    init_wspace(w, 0, w.size)
    return w


def init_wspace(w: np.ndarray, begin: int, end: int) -> bool:
    """Mark slots [begin, end) empty (utils/wspace.h:15-23)."""
    w["crd"][begin:end, 0] = -1
    return True


def Sort(array: np.ndarray, size: int, rev: bool, ctx=None) -> int:  # noqa: N802 (reference name)
    """Sort the first ``size`` records in place (utils/wspace.h:206-217).

    Ascending (rev False) returns ``size``; descending (rev True) puts the
    empty slots last and returns the number before them."""
    if array.dtype != WSPACE_DTYPE or not array.flags.c_contiguous:
        raise TypeError("Sort needs a C-contiguous WSPACE_DTYPE array")
    size = int(size)
    if size < 0 or size > array.size:
        raise ValueError("size out of range")
    out = C.c_int64()
    check(lib().spws_wspace_sort(_handle(ctx), array.ctypes.data_as(C.c_void_p), size, int(bool(rev)),
                                 C.byref(out)))
    return int(out.value)


def Merge(COO_crd, COO_vals: np.ndarray, COO_size: int, accumulator: np.ndarray,  # noqa: N802,N803
          accumulator_size: int, rev: bool = False, ctx=None) -> int:
    """Merge a sorted workspace into a sorted COO in place
    (utils/wspace.h:62-204).  ``COO_crd`` is a pair of int32 arrays and
    ``COO_vals`` an f64 array, each with capacity for COO_size +
    accumulator_size entries; returns the merged size."""
    c0, c1 = COO_crd
    for a, dt in ((c0, np.int32), (c1, np.int32), (COO_vals, np.float64)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError("COO arrays must be C-contiguous int32 / float64")
        if a.size < int(COO_size) + int(accumulator_size):
            raise ValueError("COO capacity below COO_size + accumulator_size")
    if accumulator.dtype != WSPACE_DTYPE or not accumulator.flags.c_contiguous:
        raise TypeError("accumulator must be a C-contiguous WSPACE_DTYPE array")
    if int(accumulator_size) > accumulator.size:
        raise ValueError("accumulator_size out of range")
    del rev  # the accumulator is re-sorted on the device in either direction
    out = C.c_int64()
    check(lib().spws_wspace_merge(_handle(ctx), c0.ctypes.data_as(C.c_void_p), c1.ctypes.data_as(C.c_void_p),
                                  COO_vals.ctypes.data_as(C.c_void_p), int(COO_size),
                                  accumulator.ctypes.data_as(C.c_void_p), int(accumulator_size), C.byref(out)))
    return int(out.value)
