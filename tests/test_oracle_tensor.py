"""Pins the oracle's MTTKRP and TTM restatements against numpy.einsum (no GPU).

MTTKRP with dense factors is the textbook M = X_(1) (Cf kr Bf), i.e.
einsum('ijk,jr,kr->ir'); TTM along mode 3 is einsum('ijk,kr->ijr').
"""
import numpy as np
import pytest

import oracle
from sparsews_amd import formats, gen


def small_tensor(seed, dims=(40, 30, 20), nnz=900, exps=(1.5, 1.5, 1.5), lo=-1.0, hi=1.0):
    return gen.zipf_coo3(dims, nnz, exps, seed, 5, lo, hi)


def dense(X):
    D = np.zeros(X.dims)
    D[X.idx[:, 0], X.idx[:, 1], X.idx[:, 2]] = X.vals
    return D


@pytest.mark.parametrize("w_cap", [3, 64, 1 << 16])
def test_mttkrp_matches_einsum(w_cap):
    X = small_tensor(1, lo=0.0, hi=1.0)
    R = 8
    rng = np.random.default_rng(1)
    Bf, Cf = rng.random((X.dims[1], R)), rng.random((X.dims[2], R))
    csf = formats.csf3_from_coo(X, (1, 2, 0))  # levels (j, k, i): output mode is the leaf
    oi, orr, ov = oracle.mttkrp_hash(csf, formats.dense_as_dcsr(Bf), formats.dense_as_dcsr(Cf), X.dims[0], R, w_cap)
    M = np.einsum("ijk,jr,kr->ir", dense(X), Bf, Cf)
    rows = np.unique(X.idx[:, 0])
    np.testing.assert_array_equal(oi, np.repeat(rows, R))
    np.testing.assert_array_equal(orr, np.tile(np.arange(R), len(rows)))
    np.testing.assert_allclose(ov, M[rows].ravel(), rtol=1e-13, atol=1e-13)


def test_mttkrp_sparse_factor_rows():
    """A factor row missing from the DCSR drops the products of that slice
    (co-iteration of the root level with the factor's rows)."""
    X = small_tensor(2)
    R = 4
    rng = np.random.default_rng(2)
    Bf, Cf = rng.random((X.dims[1], R)), rng.random((X.dims[2], R))
    Bd = formats.dense_as_dcsr(Bf)
    keep = np.arange(X.dims[1]) % 2 == 0
    Bd_sparse = formats.DCSR(Bd.nrows, R, Bd.l1_crd[keep],
                             (np.arange(keep.sum() + 1) * R).astype(np.int32),
                             np.tile(np.arange(R, dtype=np.int32), keep.sum()),
                             Bf[keep].ravel())
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    oi, orr, ov = oracle.mttkrp_hash(csf, Bd_sparse, formats.dense_as_dcsr(Cf), X.dims[0], R, 1 << 16)
    Bz = Bf.copy()
    Bz[~keep] = 0
    M = np.einsum("ijk,jr,kr->ir", dense(X), Bz, Cf)
    rows = np.unique(X.idx[keep[X.idx[:, 1]], 0])
    np.testing.assert_array_equal(np.unique(oi), rows)
    np.testing.assert_allclose(ov, M[rows].ravel(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("col_density", [0.4, 0.8])
def test_mttkrp_partial_factor_rows(col_density):
    """Factors with missing rows and rows storing some columns: an entry
    (i, r) exists when a nonzero X(i, j, k) meets stored Bf(j, r) and
    Cf(k, r) (the rank co-iteration of hash_free.h:286-300), its value the
    einsum over the stored entries (zero elsewhere)."""
    X = small_tensor(4, lo=0.0, hi=1.0)
    R = 6
    rng = np.random.default_rng(4)
    Bf, Cf = rng.random((X.dims[1], R)), rng.random((X.dims[2], R))
    Bd = formats.partial_dcsr(Bf, rng.random(X.dims[1]) < 0.7, col_density, 41)
    Cd = formats.partial_dcsr(Cf, rng.random(X.dims[2]) < 0.8, col_density, 42)
    Bz, Bs = formats.dcsr_dense_and_mask(Bd)
    Cz, Cs = formats.dcsr_dense_and_mask(Cd)
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    oi, orr, ov = oracle.mttkrp_hash(csf, Bd, Cd, X.dims[0], R, 1 << 16)
    present = np.zeros((X.dims[0], R), dtype=bool)
    for i, j, k in X.idx:
        present[i] |= Bs[j] & Cs[k]
    ei, er = np.nonzero(present)
    np.testing.assert_array_equal(oi, ei)
    np.testing.assert_array_equal(orr, er)
    M = np.einsum("ijk,jr,kr->ir", dense(X), Bz, Cz)
    np.testing.assert_allclose(ov, M[ei, er], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("col_density", [0.4, 1.0])
def test_ttm_partial_factor_rows(col_density):
    """U with missing rows and partial rows: fiber entry (i, j, r) exists
    when a nonzero X(i, j, k) meets a stored U(k, r) (hash_il.h:294-303 and
    its rank loop), its value the einsum over the stored entries."""
    X = small_tensor(5, dims=(30, 25, 20), exps=(1.5, 1.5, None))
    R = 5
    rng = np.random.default_rng(5)
    U = rng.uniform(-1, 1, (X.dims[2], R))
    Ud = formats.partial_dcsr(U, rng.random(X.dims[2]) < 0.6, col_density, 51)
    Uz, Us = formats.dcsr_dense_and_mask(Ud)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, Ud, X.dims[0], X.dims[1], R)
    present = np.zeros((X.dims[0], X.dims[1], R), dtype=bool)
    for i, j, k in X.idx:
        present[i, j] |= Us[k]
    ei, ej, er = np.nonzero(present)
    np.testing.assert_array_equal(ok, ei)
    np.testing.assert_array_equal(oj, ej)
    np.testing.assert_array_equal(ol, er)
    Y = np.einsum("ijk,kr->ijr", dense(X), Uz)
    np.testing.assert_allclose(ov, Y[ei, ej, er], rtol=1e-13, atol=1e-13)


def test_ttm_matches_einsum():
    X = small_tensor(3, dims=(30, 25, 20), exps=(1.5, 1.5, None))
    R = 6
    U = np.random.default_rng(3).uniform(-1, 1, (X.dims[2], R))
    csf = formats.csf3_from_coo(X, (0, 2, 1))  # levels (i, k, j): contracted mode in the middle
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), X.dims[0], X.dims[1], R)
    Y = np.einsum("ijk,kr->ijr", dense(X), U)
    fib = np.unique(X.idx[:, 0].astype(np.int64) * X.dims[1] + X.idx[:, 1])
    fi, fj = fib // X.dims[1], fib % X.dims[1]
    np.testing.assert_array_equal(ok, np.repeat(fi, R))
    np.testing.assert_array_equal(oj, np.repeat(fj, R))
    np.testing.assert_array_equal(ol, np.tile(np.arange(R), len(fib)))
    np.testing.assert_allclose(ov, Y[fi, fj].ravel(), rtol=1e-13, atol=1e-13)


def test_generator_exact_count_sorted_and_thread_independent():
    a = gen.zipf_coo3((300, 200, 100), 5000, (1.5, 1.5, None), 3, 4, 0, 1, threads=1)
    b = gen.zipf_coo3((300, 200, 100), 5000, (1.5, 1.5, None), 3, 4, 0, 1, threads=7)
    np.testing.assert_array_equal(a.idx, b.idx)
    np.testing.assert_array_equal(a.vals, b.vals)
    key = (a.idx[:, 0].astype(np.int64) * 200 + a.idx[:, 1]) * 100 + a.idx[:, 2]
    assert len(key) == 5000 and np.all(np.diff(key) > 0)
    assert a.idx[:, 0].max() < 300 and a.idx[:, 1].max() < 200 and a.idx[:, 2].max() < 100
    # Zipf marginals: the most frequent index of mode 0 is far above uniform
    assert np.bincount(a.idx[:, 0]).max() > 5 * 5000 / 300
    with pytest.raises(ValueError):
        gen.zipf_coo3((2, 2, 2), 9, (None, None, None), 0, 0, 0, 1)
