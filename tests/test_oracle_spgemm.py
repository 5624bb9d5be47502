"""Pins the CPU oracle's SpGEMM functions (no GPU).

The reference holds no golden vectors for this path and cannot be compiled
here (its headers need Eigen/Boost/SPLATT), so the oracle is pinned by:
scipy (a library) and integer pattern products for structure, dense numpy
products for values, exact special cases (identity, permutation, diagonal
scaling), and a hand-derived known answer for the hash workspace's flush
grouping.
"""
import numpy as np
import pytest

import oracle
from helpers import (assert_same_structure, dense, pattern_product, powerlaw_rows, random_csr,
                     scipy_product, uniform_rows)
from sparsews_amd.gen import CSR, rowwise_pair


CASES = [
    ("uniform", lambda: (uniform_rows(300, 200, 0, 12, 1, 0), uniform_rows(200, 250, 0, 12, 1, 1))),
    ("powerlaw", lambda: (powerlaw_rows(400, 400, 1.8, 200, 2, 0), powerlaw_rows(400, 400, 1.8, 200, 2, 1))),
    ("rect", lambda: (uniform_rows(50, 500, 1, 40, 3, 0), uniform_rows(500, 30, 1, 30, 3, 1))),
]


@pytest.mark.parametrize("name,make", CASES)
def test_gustavson_matches_scipy_and_pattern(name, make):
    A, B = make()
    C = oracle.spgemm_gustavson(A, B)
    pos, crd, cnt = pattern_product(A, B)
    assert_same_structure(C, pos, crd)
    S = scipy_product(A, B)
    # scipy prunes exact zeros; U[-1,1] products never cancel exactly here
    assert S.nnz == C.nnz
    exact, absum = oracle.spgemm_refsum(A, B, C.pos, C.crd)
    assert oracle.max_value_error(C.vals, S.data, absum) <= 1e-15
    assert oracle.max_value_error(C.vals, exact, absum) <= 1e-15


@pytest.mark.parametrize("name,make", CASES)
def test_values_match_dense_product(name, make):
    A, B = make()
    C = oracle.spgemm_gustavson(A, B)
    D = dense(A) @ dense(B)
    rows = np.repeat(np.arange(C.nrows), np.diff(C.pos))
    np.testing.assert_allclose(C.vals, D[rows, C.crd], rtol=1e-12, atol=1e-12)


def test_symbolic_and_upper_bound():
    A, B = uniform_rows(300, 300, 0, 20, 7, 0), powerlaw_rows(300, 300, 1.8, 100, 7, 1)
    pos, crd, cnt = pattern_product(A, B)
    np.testing.assert_array_equal(oracle.spgemm_symbolic(A, B), pos)
    ub = oracle.spgemm_row_ub(A, B)
    # ub counts products: sum of the per-entry product counts of each row
    rows = np.repeat(np.arange(A.nrows), np.diff(pos))
    np.testing.assert_array_equal(ub, np.bincount(rows, weights=cnt, minlength=A.nrows).astype(np.int64))
    assert np.all(ub >= np.diff(pos))


@pytest.mark.parametrize("name,make", CASES)
def test_row_counts_match_the_pattern_product(name, make):
    """Per-row entry counts (int64 total, no int32 prefix) equal the row
    lengths of the integer pattern product."""
    A, B = make()
    pos, _, _ = pattern_product(A, B)
    cnt, total = oracle.spgemm_row_counts(A, B)
    np.testing.assert_array_equal(cnt, np.diff(pos))
    assert total == int(pos[-1])


def test_identity_gives_b_exactly():
    B = uniform_rows(120, 90, 0, 15, 4, 1)
    n = B.nrows
    I = CSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    for C in (oracle.spgemm_gustavson(I, B), oracle.spgemm_hash(I, B, 4)):
        np.testing.assert_array_equal(C.pos, B.pos)
        np.testing.assert_array_equal(C.crd, B.crd)
        np.testing.assert_array_equal(C.vals, B.vals)


def test_permutation_and_diagonal_scaling_exact():
    B = uniform_rows(64, 80, 1, 10, 5, 1)
    n = B.nrows
    perm = np.random.default_rng(5).permutation(n)
    d = np.random.default_rng(6).uniform(-2, 2, n)
    P = CSR(n, n, np.arange(n + 1, dtype=np.int32), perm.astype(np.int32), d)
    C = oracle.spgemm_gustavson(P, B)
    L = np.diff(B.pos)
    for i in range(n):
        j = perm[i]
        sl = slice(B.pos[j], B.pos[j + 1])
        np.testing.assert_array_equal(C.crd[C.pos[i]:C.pos[i + 1]], B.crd[sl])
        np.testing.assert_array_equal(C.vals[C.pos[i]:C.pos[i + 1]], d[i] * B.vals[sl])
    assert C.nnz == L.sum()


def test_empty_operands():
    A = CSR(5, 4, np.zeros(6, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    B = uniform_rows(4, 7, 1, 3, 8, 1)
    for C in (oracle.spgemm_gustavson(A, B), oracle.spgemm_hash(A, B, 2)):
        assert C.nnz == 0 and np.all(C.pos == 0)


@pytest.mark.parametrize("w_cap", [1, 3, 16, 1000])
def test_hash_flush_schedule_structure(w_cap):
    A, B = powerlaw_rows(200, 200, 1.8, 80, 9, 0), uniform_rows(200, 200, 0, 10, 9, 1)
    G = oracle.spgemm_gustavson(A, B)
    H = oracle.spgemm_hash(A, B, w_cap)
    assert_same_structure(H, G.pos, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(H.vals, exact, absum) <= 1e-14


def test_hash_without_flush_is_left_to_right():
    """With a workspace larger than the output no flush happens, so every
    entry is the left-to-right sum in j order -- Gustavson's order."""
    A, B = powerlaw_rows(200, 200, 1.8, 80, 10, 0), powerlaw_rows(200, 200, 1.8, 80, 10, 1)
    G = oracle.spgemm_gustavson(A, B)
    H = oracle.spgemm_hash(A, B, 1 << 20)
    np.testing.assert_array_equal(H.vals, G.vals)


@pytest.mark.parametrize("cap,w_cap", [(4, 1 << 20), (7, 1 << 20), (3, 64), (5, 16)])
def test_hash_parallel_is_each_workers_sequential_hash(cap, w_cap):
    """The CAP-parallel workspace (CSR_CSR_hash_parallel.h): worker t owns
    rows t, t + cap, ..., so each of its rows must equal, bit for bit, the
    sequential hash path run on A restricted to worker t's rows (the same
    flush schedule over the same inserts); without flushes every row is the
    left-to-right (Gustavson) sum."""
    A, B = powerlaw_rows(300, 250, 1.8, 120, 11, 0), powerlaw_rows(250, 280, 1.8, 120, 11, 1)
    P, flushes = oracle.spgemm_hash_parallel(A, B, w_cap, cap)
    G = oracle.spgemm_gustavson(A, B)
    assert_same_structure(P, G.pos, G.crd)
    if w_cap >= 1 << 20:
        assert flushes == 0
        np.testing.assert_array_equal(P.vals, G.vals)
    else:
        assert flushes > 0
    for t in range(cap):
        lens = A.row_lengths().copy()
        lens[np.arange(A.nrows) % cap != t] = 0
        keep = np.repeat(np.arange(A.nrows) % cap == t, A.row_lengths())
        At = CSR(A.nrows, A.ncols, np.concatenate([[0], np.cumsum(lens)]).astype(np.int32), A.crd[keep],
                 A.vals[keep])
        H = oracle.spgemm_hash(At, B, w_cap)
        for i in range(t, A.nrows, cap):
            np.testing.assert_array_equal(P.vals[P.pos[i]:P.pos[i + 1]], H.vals[H.pos[i]:H.pos[i + 1]])


def test_hash_flush_grouping_known_answer():
    """Hand-derived from CSR_CSR_hash_sequential.h:151-190, 280-298.

    One row, A(0,:) = [1, 1, 1] at j = 0, 1, 2; B(0,:) = {0: 1e16, 1: 1},
    B(1,:) = {0: -1e16, 1: 1}, B(2,:) = {0: 1}.  With w_cap = 1 the buffer
    holds 2 entries: products (0,0,1e16), (0,1,1) fill it; the next product
    (0,0,-1e16) finds numel == buffer_capacity, so epoch 1 is flushed into
    the all-array, the capacity grows to 4 and the product opens epoch 2,
    which then receives (0,1,1) and (0,0,1).  Epoch 2's partial for (0,0) is
    -1e16 + 1 = -1e16 (ulp(1e16) = 2), and Merge_hash adds it to the stored
    1e16: C(0,0) = 0.  Without a flush the left-to-right sum is
    (1e16 - 1e16) + 1 = 1.  C(0,1) = 1 + 1 = 2 either way.
    """
    A = CSR(1, 3, np.array([0, 3], dtype=np.int32), np.array([0, 1, 2], dtype=np.int32), np.ones(3))
    B = CSR(3, 2, np.array([0, 2, 4, 5], dtype=np.int32), np.array([0, 1, 0, 1, 0], dtype=np.int32),
            np.array([1e16, 1.0, -1e16, 1.0, 1.0]))
    H1 = oracle.spgemm_hash(A, B, 1)
    np.testing.assert_array_equal(H1.vals, [0.0, 2.0])
    Hbig = oracle.spgemm_hash(A, B, 1024)
    np.testing.assert_array_equal(Hbig.vals, [1.0, 2.0])
    G = oracle.spgemm_gustavson(A, B)
    np.testing.assert_array_equal(G.vals, [1.0, 2.0])


def test_rowwise_workload_full_size():
    """The 10,000 x 10,000, 8-per-row workload: oracle vs scipy, whole matrix."""
    A, B = rowwise_pair()
    G = oracle.spgemm_gustavson(A, B)
    pos, crd, cnt = pattern_product(A, B)
    assert_same_structure(G, pos, crd)
    assert int(cnt.sum()) == 640_000
    H = oracle.spgemm_hash(A, B, 1 << 16)
    assert_same_structure(H, pos, crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(H.vals, exact, absum) <= 1e-15
    assert oracle.max_value_error(G.vals, exact, absum) <= 1e-15


def test_value_error_metric_catches_a_dropped_term():
    """A plausible kernel mistake (one product dropped) must fail the bar."""
    A, B = uniform_rows(100, 100, 2, 10, 11, 0), uniform_rows(100, 100, 2, 10, 11, 1)
    G = oracle.spgemm_gustavson(A, B)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    bad = G.vals.copy()
    # drop the first product of entry 0 of row 0
    j = A.crd[A.pos[0]]
    k = G.crd[0]
    q = np.nonzero(B.crd[B.pos[j]:B.pos[j + 1]] == k)[0]
    if len(q):
        bad[0] -= A.vals[A.pos[0]] * B.vals[B.pos[j] + q[0]]
        assert oracle.max_value_error(bad, exact, absum) > 1e-12


@pytest.mark.parametrize("m,n,r,seed", [(300, 500, 7, 1), (1, 40, 40, 2), (50, 1, 1, 3), (200, 200, 0, 4)])
def test_csr_transpose_oracle(m, n, r, seed):
    """The transpose() restatement against scipy's transpose (row pointers,
    increasing columns, values); transposing twice gives A back; empty rows
    and columns (a transposed operand or a dropped count fails these)."""
    import scipy.sparse as sp
    from sparsews_amd.gen import uniform_csr
    A = uniform_csr(m, n, r, seed, 0)
    T = oracle.csr_transpose(A)
    S = sp.csr_matrix((A.vals, A.crd, A.pos), shape=(m, n)).T.tocsr()
    S.sort_indices()
    assert (T.nrows, T.ncols) == (n, m)
    np.testing.assert_array_equal(T.pos, S.indptr)
    np.testing.assert_array_equal(T.crd, S.indices)
    np.testing.assert_array_equal(T.vals, S.data)
    TT = oracle.csr_transpose(T)
    np.testing.assert_array_equal(TT.pos, A.pos)
    np.testing.assert_array_equal(TT.crd, A.crd)
    np.testing.assert_array_equal(TT.vals, A.vals)

