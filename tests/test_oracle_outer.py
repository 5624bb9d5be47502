"""Pins the oracle's outer-product coordinate workspace (no GPU)."""
import numpy as np
import pytest

import oracle
from helpers import assert_same_structure, pattern_product, powerlaw_rows, uniform_rows
from sparsews_amd import formats, gen
from sparsews_amd.gen import CSR


def dcsc_dcsr(A, B):
    return formats.dcsc_from_csr(A), formats.dcsr_from_csr(B)


@pytest.mark.parametrize("w_cap", [7, 100, 1 << 20])
def test_outer_coord_matches_rowwise(w_cap):
    A, B = powerlaw_rows(300, 250, 1.8, 120, 50, 0), powerlaw_rows(250, 280, 1.8, 120, 50, 1)
    Ac, Bd = dcsc_dcsr(A, B)
    pos, crd, vals = oracle.outer_coord(Ac, Bd, A.nrows, B.ncols, w_cap)
    G = oracle.spgemm_gustavson(A, B)
    assert_same_structure(CSR(A.nrows, B.ncols, pos, crd, vals), G.pos, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(vals, exact, absum) <= 1e-14
    if w_cap == 1 << 20:
        # one workspace, stable sort: every entry summed in increasing j
        np.testing.assert_array_equal(vals, G.vals)


def test_outer_products_count():
    A, B = uniform_rows(100, 90, 0, 9, 51, 0), uniform_rows(90, 120, 0, 9, 51, 1)
    Ac, Bd = dcsc_dcsr(A, B)
    _, _, cnt = pattern_product(A, B)
    assert oracle.outer_products(Ac, Bd) == int(cnt.sum())


def test_outer_cc_gaps():
    """Stored columns of A without a stored row of B (and vice versa) are skipped
    by the co-iteration, exactly as a zero row would be."""
    A = uniform_rows(40, 60, 0, 6, 52, 0)
    B = uniform_rows(60, 50, 0, 6, 52, 1)
    # empty every third row of B and every fifth column of A
    keepB = np.ones(B.nnz, bool)
    for j in range(0, 60, 3):
        keepB[B.pos[j]:B.pos[j + 1]] = False
    lensB = np.array([keepB[B.pos[j]:B.pos[j + 1]].sum() for j in range(60)])
    B2 = CSR(60, 50, np.concatenate([[0], np.cumsum(lensB)]).astype(np.int32), B.crd[keepB], B.vals[keepB])
    keepA = A.crd % 5 != 0
    lensA = np.array([keepA[A.pos[i]:A.pos[i + 1]].sum() for i in range(40)])
    A2 = CSR(40, 60, np.concatenate([[0], np.cumsum(lensA)]).astype(np.int32), A.crd[keepA], A.vals[keepA])
    Ac, Bd = dcsc_dcsr(A2, B2)
    pos, crd, vals = oracle.outer_coord(Ac, Bd, 40, 50, 16)
    G = oracle.spgemm_gustavson(A2, B2)
    assert_same_structure(CSR(40, 50, pos, crd, vals), G.pos, G.crd)


def test_outer_workload_scaled():
    """The outer-product generator at N = 6,000, cap 600 (A = B^T)."""
    N = 6_000
    B = gen.zipf_csr(N, 1.8, 600, seed=1, stream=2)
    A = B.transpose()
    Ac = formats.dcsr_from_csr(B)  # CSC(A) shares B's arrays
    Ac_check = formats.dcsc_from_csr(A)
    np.testing.assert_array_equal(Ac.l2_pos, Ac_check.l2_pos)
    np.testing.assert_array_equal(Ac.l2_crd, Ac_check.l2_crd)
    Bd = formats.dcsr_from_csr(B)
    L = B.row_lengths()
    assert oracle.outer_products(Ac, Bd) == int((L * L).sum())
    pos, crd, vals = oracle.outer_coord(Ac, Bd, N, N, 1 << 20)
    G = oracle.spgemm_gustavson(A, B)
    assert_same_structure(CSR(N, N, pos, crd, vals), G.pos, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(vals, exact, absum) <= 1e-14
