"""The C-ABI library loads and exports every symbol include/spws/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

from sparsews_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.header_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_binding_table_covers_the_header():
    declared = set(_lib.header_symbols())
    bound = set(_lib.SIGNATURES)
    assert bound == declared, (declared - bound, bound - declared)


def test_exports_are_extern_c():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for s in _lib.header_symbols():
        assert s in exported, s


def test_null_arguments_are_rejected_without_a_device():
    L = _lib.lib()
    # a null output pointer is an argument error before any device call
    rc = L.spws_ctx_create(0, None, None)
    assert rc == 1
    assert b"null" in L.spws_last_error()
    assert L.spws_ctx_destroy(None) == 0
    assert L.spws_ctx_get_stats(None, None) == 1


def test_taco_struct_layout_matches_header():
    # order, dimensions, csize, mode_ordering, mode_types, indices, vals, vals_size
    t = _lib.TacoTensor
    assert t.dimensions.offset == 8 and t.indices.offset == 40 and t.vals.offset == 48
    assert ctypes.sizeof(t) == 64
