"""The no-allocation CSR transpose (spws_csr_transpose_dev_noalloc) on the
GPU, through the C ABI: bit-identical to spws_csr_transpose_dev on two calls
in a row, the second one smaller (the shared scratch keeps the first call's
contents, so a stale column count or histogram would show), and inputs
beyond the fixed scratch refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import random_csr
from sparsews_amd._lib import SpwsError, check, lib
from sparsews_amd.device import DeviceArray, DeviceCSR, csr_transpose

pytestmark = pytest.mark.gpu


def noalloc_transpose(ctx, dA):
    m, n, nnz = dA.nrows, dA.ncols, dA.nnz
    out_pos = DeviceArray(ctx, n + 1, np.int32)
    out_crd = DeviceArray(ctx, max(nnz, 1), np.int32)
    out_val = DeviceArray(ctx, max(nnz, 1), np.float64)
    check(lib().spws_csr_transpose_dev_noalloc(C.c_void_p(ctx.stream), m, n, nnz, dA.pos.ptr, dA.crd.ptr,
                                               dA.vals.ptr, out_pos.ptr, out_crd.ptr, out_val.ptr))
    ctx.sync()
    return out_pos.to_host(), out_crd.to_host()[:nnz], out_val.to_host()[:nnz]


@pytest.mark.parametrize("sizes", [((3000, 5000, 40), (700, 900, 9)), ((20_000, 30_000, 30), (50, 64, 3))])
def test_noalloc_transpose_matches_the_allocating_entry(ctx, sizes):
    for call, (m, n, r) in enumerate(sizes):
        print(f"case: no-alloc transpose call {call}: {m} x {n}, {r} per row", flush=True)
        A = random_csr(m, n, np.random.default_rng(m).integers(0, r + 1, m), 40 + call, 0)
        dA = DeviceCSR.upload(ctx, A)
        pos, crd, val = noalloc_transpose(ctx, dA)
        T = csr_transpose(ctx, dA).download()
        np.testing.assert_array_equal(pos, T.pos)
        np.testing.assert_array_equal(crd, T.crd)
        np.testing.assert_array_equal(val, T.vals)
        # and the transpose itself: scipy's, bit for bit (pure data movement)
        S = A.to_scipy().T.tocsr()
        S.sort_indices()
        np.testing.assert_array_equal(pos, S.indptr)
        np.testing.assert_array_equal(crd, S.indices)
        np.testing.assert_array_equal(val, S.data)


def test_noalloc_transpose_refuses_inputs_beyond_the_scratch(ctx):
    print("case: no-alloc transpose beyond its scratch", flush=True)
    nnz = 4_000_000  # 33 B per entry is more than 48 MiB; rejected on the host
    dummy = DeviceArray(ctx, 16, np.int32)
    with pytest.raises(SpwsError) as ex:
        check(lib().spws_csr_transpose_dev_noalloc(C.c_void_p(ctx.stream), 10, 10, nnz, dummy.ptr, dummy.ptr,
                                                   dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr))
    assert ex.value.code == 4
    assert "scratch" in str(ex.value)
