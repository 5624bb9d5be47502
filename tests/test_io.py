"""Matrix Market / FROSTT readers (CPU).  Parsing is pinned against
scipy.io.mmread (an independent reader) on general, symmetric,
skew-symmetric, integer and pattern files; the tensor reader against the
nonzeros it was written from."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import oracle
from helpers import powerlaw_rows
from sparsews_amd import gen, io as spio


def as_scipy(A):
    return sp.csr_matrix((A.vals, A.crd, np.asarray(A.pos, np.int64)), shape=(A.nrows, A.ncols))


def check_against_mmread(path, A):
    R = scipy.io.mmread(str(path)).tocsr()
    R.sum_duplicates()
    R.sort_indices()
    assert (A.nrows, A.ncols) == R.shape
    np.testing.assert_array_equal(np.asarray(A.pos, np.int64), R.indptr)
    np.testing.assert_array_equal(A.crd, R.indices)
    np.testing.assert_array_equal(A.vals, R.data)


@pytest.mark.parametrize("field,symmetry", [("real", "general"), ("real", "symmetric"),
                                            ("real", "skew-symmetric"), ("integer", "general"),
                                            ("pattern", "general"), ("pattern", "symmetric")])
def test_mtx_matches_scipy(tmp_path, field, symmetry):
    rng = np.random.default_rng(7)
    n = 40
    M = sp.random(n, n, density=0.1, random_state=7, format="coo")
    if field == "integer":
        M.data = rng.integers(-9, 10, M.nnz).astype(float)
    if symmetry != "general":
        M = sp.tril(M, k=-1 if symmetry == "skew-symmetric" else 0).tocoo()
    path = tmp_path / f"m_{field}_{symmetry}.mtx"
    scipy.io.mmwrite(str(path), M, field=field, symmetry=symmetry)
    A = spio.read_mtx(path)
    check_against_mmread(path, A)


def test_mtx_duplicates_sum_or_keep(tmp_path):
    path = tmp_path / "dup.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real general\n% two entries at (2, 3)\n"
                    "3 4 4\n2 3 1.5\n1 1 2.0\n2 3 0.25\n3 4 -1\n")
    A = spio.read_mtx(path)
    np.testing.assert_array_equal(np.asarray(A.pos), [0, 1, 2, 3])
    np.testing.assert_array_equal(A.crd, [0, 2, 3])
    np.testing.assert_array_equal(A.vals, [2.0, 1.75, -1.0])
    K = spio.read_mtx(path, duplicates="keep")
    np.testing.assert_array_equal(np.asarray(K.pos), [0, 1, 3, 4])
    np.testing.assert_array_equal(K.crd, [0, 2, 2, 3])
    np.testing.assert_array_equal(K.vals, [2.0, 1.5, 0.25, -1.0])  # file order within the coordinate


def test_mtx_round_trip_and_product(tmp_path):
    """write_mtx / read_mtx keep values bit for bit, and the loaded matrix
    feeds the SpGEMM oracle like a generated one."""
    A = powerlaw_rows(120, 90, 1.8, 40, 3, 0)
    B = powerlaw_rows(90, 70, 1.8, 30, 3, 1)
    pa, pb = tmp_path / "a.mtx", tmp_path / "b.mtx"
    spio.write_mtx(pa, A, comment="test matrix")
    spio.write_mtx(pb, B)
    A2, B2 = spio.read_mtx(pa), spio.read_mtx(pb)
    np.testing.assert_array_equal(np.asarray(A2.pos), np.asarray(A.pos))
    np.testing.assert_array_equal(A2.crd, A.crd)
    np.testing.assert_array_equal(A2.vals, A.vals)
    G = oracle.spgemm_gustavson(A2, B2)
    ref = (as_scipy(A) @ as_scipy(B)).tocsr()
    ref.sort_indices()
    assert G.nnz >= ref.nnz  # the oracle keeps explicit zeros from cancellation
    dense = np.zeros((A.nrows, B.ncols))
    rows = np.repeat(np.arange(A.nrows), np.diff(np.asarray(G.pos, np.int64)))
    dense[rows, G.crd] = G.vals
    np.testing.assert_allclose(dense, ref.toarray(), rtol=1e-12, atol=1e-12)


def test_mtx_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(ValueError):
        spio.read_mtx(p)
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n3 1 1\n")
    with pytest.raises(ValueError):
        spio.read_mtx(p)
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 1 1\n")
    with pytest.raises(ValueError):
        spio.read_mtx(p)


def test_tns_round_trip(tmp_path):
    X = gen.zipf_coo3((50, 40, 30), 2000, (1.5, 1.5, None), 21, 5, -1.0, 1.0)
    p = tmp_path / "x.tns"
    spio.write_tns(p, X)
    Y = spio.read_tns(p, dims=X.dims)
    assert Y.dims == X.dims
    np.testing.assert_array_equal(Y.idx, X.idx)
    np.testing.assert_array_equal(Y.vals, X.vals)
    Z = spio.read_tns(p)  # dims from the largest index of each mode
    assert all(z <= x for z, x in zip(Z.dims, X.dims))


def test_tns_unsorted_with_comments(tmp_path):
    p = tmp_path / "y.tns"
    p.write_text("# a FROSTT tensor\n3 1 2 0.5\n1 2 1 -2.0\n1 1 3 1e-3\n")
    Y = spio.read_tns(p)
    assert Y.dims == (3, 2, 3)
    np.testing.assert_array_equal(Y.idx, [[0, 0, 2], [0, 1, 0], [2, 0, 1]])
    np.testing.assert_array_equal(Y.vals, [1e-3, -2.0, 0.5])
    p.write_text("1 1 1 1.0\n1 1 1 2.0\n")
    with pytest.raises(ValueError):
        spio.read_tns(p)
