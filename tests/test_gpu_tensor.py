"""MTTKRP and TTM on the GPU vs the CPU oracle, through the C ABI.

Structure (output rows / fibers) must be bit-exact; values within
|x - y| <= 1e-12 * max(|y|, sum|terms|).
"""
import numpy as np
import pytest

import oracle
from sparsews_amd import formats, gen, taco
from sparsews_amd.device import Context, DeviceArray, DeviceCSF, mttkrp, set_ttm_method, ttm

pytestmark = pytest.mark.gpu
TOL = 1e-12


def dense(X):
    D = np.zeros(X.dims)
    D[X.idx[:, 0], X.idx[:, 1], X.idx[:, 2]] = X.vals
    return D


def run_mttkrp(ctx, X, Bf, Cf):
    R = Bf.shape[1]
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    rows, M = mttkrp(ctx, DeviceCSF.upload(ctx, csf), X.dims[0], X.dims[1], X.dims[2], R,
                     DeviceArray.from_host(ctx, np.ascontiguousarray(Bf.ravel())),
                     DeviceArray.from_host(ctx, np.ascontiguousarray(Cf.ravel())))
    return rows.to_host(), M.to_host().reshape(X.dims[0], R)


def run_ttm(ctx, X, U, method=0):
    R = U.shape[1]
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    set_ttm_method(ctx, method)
    try:
        fi, fj, Y = ttm(ctx, DeviceCSF.upload(ctx, csf), X.dims[0], X.dims[1], X.dims[2], R,
                        DeviceArray.from_host(ctx, np.ascontiguousarray(U.ravel())))
        return fi.to_host(), fj.to_host(), Y.to_host().reshape(-1, R)
    finally:
        set_ttm_method(ctx, 0)


@pytest.mark.parametrize("R", [32, 8, 48])
def test_mttkrp_small(ctx, R):
    print(f"case: mttkrp small R={R}", flush=True)
    X = gen.zipf_coo3((400, 300, 200), 20000, (1.5, 1.5, 1.5), 11, 5, 0.0, 1.0)
    rng = np.random.default_rng(R)
    Bf, Cf = rng.random((300, R)), rng.random((200, R))
    rows, M = run_mttkrp(ctx, X, Bf, Cf)
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    oi, orr, ov = oracle.mttkrp_hash(csf, formats.dense_as_dcsr(Bf), formats.dense_as_dcsr(Cf), 400, R, 1 << 16)
    np.testing.assert_array_equal(np.repeat(rows, R), oi)
    absM = np.einsum("ijk,jr,kr->ir", np.abs(dense(X)), np.abs(Bf), np.abs(Cf))
    assert oracle.max_value_error(M[rows].ravel(), ov, absM[rows].ravel()) <= TOL
    # rows without a nonzero stay exactly zero
    empty = np.setdiff1d(np.arange(400), rows)
    assert np.all(M[empty] == 0)


def test_mttkrp_taco_entry():
    print("case: mttkrp taco entry", flush=True)
    X = gen.zipf_coo3((120, 90, 70), 3000, (1.5, 1.5, 1.5), 12, 5, 0.0, 1.0)
    R = 16
    rng = np.random.default_rng(12)
    Bf, Cf = rng.random((90, R)), rng.random((70, R))
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    At = taco.init_taco_tensor_COO([120, R])
    Bt = taco.csf_to_taco_tensor(csf)
    Ct = taco.CC_to_taco_tensor(formats.dense_as_dcsr(Bf), 90, R)
    Dt = taco.CC_to_taco_tensor(formats.dense_as_dcsr(Cf), 70, R)
    taco.mttkrp_compute(At, Bt, Ct, Dt, 1 << 10)
    gi, gr, gv = taco.taco_coo(At)
    oi, orr, ov = oracle.mttkrp_hash(csf, formats.dense_as_dcsr(Bf), formats.dense_as_dcsr(Cf), 120, R, 1 << 10)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gr, orr)
    absM = np.einsum("ijk,jr,kr->ir", np.abs(dense(X)), np.abs(Bf), np.abs(Cf))
    assert oracle.max_value_error(gv, ov, absM[oi, orr]) <= TOL


@pytest.mark.parametrize("R", [16, 5, 40])
def test_ttm_small(ctx, R):
    print(f"case: ttm small R={R}", flush=True)
    X = gen.zipf_coo3((500, 400, 60), 20000, (1.5, 1.5, None), 13, 5, -1.0, 1.0)
    U = np.random.default_rng(R).uniform(-1, 1, (60, R))
    fi, fj, Y = run_ttm(ctx, X, U)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), 500, 400, R)
    np.testing.assert_array_equal(np.repeat(fi, R), ok)
    np.testing.assert_array_equal(np.repeat(fj, R), oj)
    absY = np.einsum("ijk,kr->ijr", np.abs(dense(X)), np.abs(U))
    assert oracle.max_value_error(Y.ravel(), ov, absY[fi, fj].ravel()) <= TOL


TTM_ROOT_CASES = [
    # (dims, nnz, exponents, seed): Zipf roots -- a few heavy roots of many
    # 2,048-nonzero tiles and light ones; uniform roots -- light roots only;
    # Zipf roots over many roots -- both kinds, heavy ones of one tile
    ((60, 5000, 300), 200_000, (1.2, 1.5, None), 23),
    ((5000, 3000, 50), 100_000, (None, None, None), 24),
    ((20_000, 4000, 200), 300_000, (1.1, 1.3, None), 25),
]


@pytest.mark.parametrize("R", [16, 5, 32, 40, 64])
@pytest.mark.parametrize("case", range(len(TTM_ROOT_CASES)))
def test_ttm_roots_and_windows(ctx, R, case):
    """The roots-and-windows path (light roots sorted in registers, heavy
    roots bucketed by column window and added in dense LDS windows of
    1024 / R' columns): every fiber summed from 0.0 over k in increasing
    order, so structure and values equal the sequential hash_il oracle bit
    for bit; the sorting path gives the same structure and values within
    the tolerance (it splits fibers of more than 128 nonzeros into partial
    sums)."""
    dims, nnz, exps, seed = TTM_ROOT_CASES[case]
    print(f"case: ttm roots and windows R={R} dims={dims} nnz={nnz}", flush=True)
    X = gen.zipf_coo3(dims, nnz, exps, seed, 5, -1.0, 1.0)
    U = np.random.default_rng(50 + R).uniform(-1, 1, (dims[2], R))
    fi, fj, Y = run_ttm(ctx, X, U, method=2)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), dims[0], dims[1], R)
    np.testing.assert_array_equal(np.repeat(fi, R), ok)
    np.testing.assert_array_equal(np.repeat(fj, R), oj)
    np.testing.assert_array_equal(Y.ravel(), ov)
    si, sj, SY = run_ttm(ctx, X, U, method=1)
    np.testing.assert_array_equal(si, fi)
    np.testing.assert_array_equal(sj, fj)
    scale = np.abs(Y).max() + 1.0
    assert np.max(np.abs(SY - Y)) <= 1e-12 * scale * 16


def test_ttm_roots_declines_wide_output_mode(ctx):
    """More column windows than the per-root counters hold (J = 300,000 at
    R = 16): method 2 refuses with SPWS_ERR_UNSUPPORTED, the automatic method
    takes the sorting path."""
    print("case: ttm roots path declines wide J", flush=True)
    from sparsews_amd._lib import SpwsError
    X = gen.zipf_coo3((40, 300_000, 30), 5000, (None, None, None), 26, 5, -1.0, 1.0)
    U = np.random.default_rng(7).uniform(-1, 1, (30, 16))
    with pytest.raises(SpwsError):
        run_ttm(ctx, X, U, method=2)
    fi, fj, Y = run_ttm(ctx, X, U)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), 40, 300_000, 16)
    np.testing.assert_array_equal(np.repeat(fi, 16), ok)
    np.testing.assert_array_equal(np.repeat(fj, 16), oj)


@pytest.mark.parametrize("R", [16, 40])
def test_ttm_long_fibers(ctx, R):
    """Fibers of 139-200 nonzeros: two 128-nonzero work items per fiber
    (partial sums combined in item order), items longer than one 16-lane
    load, ranks in three 16-lane slices for R = 40."""
    print(f"case: ttm long fibers R={R}", flush=True)
    X = gen.zipf_coo3((20, 30, 400), 100000, (None, None, None), 21, 5, -1.0, 1.0)
    U = np.random.default_rng(30 + R).uniform(-1, 1, (400, R))
    fi, fj, Y = run_ttm(ctx, X, U)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), 20, 30, R)
    np.testing.assert_array_equal(np.repeat(fi, R), ok)
    np.testing.assert_array_equal(np.repeat(fj, R), oj)
    absY = np.einsum("ijk,kr->ijr", np.abs(dense(X)), np.abs(U))
    assert oracle.max_value_error(Y.ravel(), ov, absY[fi, fj].ravel()) <= TOL


def test_ttm_many_items_per_group(ctx):
    """~400k short fibers: each 16-lane group of the accumulation grid walks
    about two dozen items, with the next item's keys and record loaded
    ahead."""
    print("case: ttm many items per group", flush=True)
    X = gen.zipf_coo3((3000, 3000, 50), 400000, (None, None, None), 22, 5, -1.0, 1.0)
    R = 16
    U = np.random.default_rng(40).uniform(-1, 1, (50, R))
    fi, fj, Y = run_ttm(ctx, X, U)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), 3000, 3000, R)
    np.testing.assert_array_equal(np.repeat(fi, R), ok)
    np.testing.assert_array_equal(np.repeat(fj, R), oj)
    # |X| |U| per fiber from the sorted coordinates (no dense 3000 x 3000 x 50 copy)
    order = np.lexsort((X.idx[:, 2], X.idx[:, 1], X.idx[:, 0]))
    xi, xj, xk = (X.idx[order, m].astype(np.int64) for m in range(3))
    fid = np.searchsorted(fi.astype(np.int64) * 3000 + fj, xi * 3000 + xj)
    absY = np.zeros((len(fi), R))
    np.add.at(absY, fid, np.abs(X.vals[order])[:, None] * np.abs(U[xk]))
    assert oracle.max_value_error(Y.ravel(), ov, absY.ravel()) <= TOL


def test_ttm_wide_coordinates(ctx):
    """Dimensions whose packed (i, j, k) coordinates need more than 64 bits
    (2^24 x 2^24 x 2^17): the fibers are grouped through the hash table
    instead of one sort on (i, j, k).  Fibers and their sums are checked
    against a numpy grouping of the nonzeros, summed in increasing k."""
    print("case: ttm 65-bit coordinates", flush=True)
    dims = (1 << 24, 1 << 24, 1 << 17)
    X = gen.zipf_coo3(dims, 3000, (1.5, 1.5, None), 16, 5, -1.0, 1.0)
    R = 4
    U = np.random.default_rng(16).uniform(-1, 1, (dims[2], R))
    fi, fj, Y = run_ttm(ctx, X, U)
    i, j, k, v = X.idx[:, 0], X.idx[:, 1], X.idx[:, 2], X.vals
    order = np.lexsort((k, j, i))
    i, j, k, v = i[order], j[order], k[order], v[order]
    head = np.ones(len(i), dtype=bool)
    head[1:] = (i[1:] != i[:-1]) | (j[1:] != j[:-1])
    np.testing.assert_array_equal(fi, i[head])
    np.testing.assert_array_equal(fj, j[head])
    fib = np.cumsum(head) - 1
    ref = np.zeros((head.sum(), R))
    ab = np.zeros((head.sum(), R))
    for p in range(len(i)):  # in increasing k within each fiber
        ref[fib[p]] += v[p] * U[k[p]]
        ab[fib[p]] += np.abs(v[p] * U[k[p]])
    assert oracle.max_value_error(Y.ravel(), ref.ravel(), ab.ravel()) <= TOL


def test_ttm_frostt_file(ctx, tmp_path):
    """A FROSTT .tns file read with its values, then TTM on the device."""
    print("case: ttm on a FROSTT file", flush=True)
    from sparsews_amd import io as spio
    X0 = gen.zipf_coo3((300, 200, 40), 8000, (1.5, 1.5, None), 17, 5, -1.0, 1.0)
    path = tmp_path / "x.tns"
    spio.write_tns(path, X0)
    X = spio.read_tns(path, dims=X0.dims)
    R = 8
    U = np.random.default_rng(17).uniform(-1, 1, (40, R))
    fi, fj, Y = run_ttm(ctx, X, U)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), 300, 200, R)
    np.testing.assert_array_equal(np.repeat(fi, R), ok)
    np.testing.assert_array_equal(np.repeat(fj, R), oj)
    absY = np.einsum("ijk,kr->ijr", np.abs(dense(X)), np.abs(U))
    assert oracle.max_value_error(Y.ravel(), ov, absY[fi, fj].ravel()) <= TOL


def test_ttm_taco_entry():
    print("case: ttm taco entry", flush=True)
    X = gen.zipf_coo3((80, 70, 30), 2500, (1.5, 1.5, None), 14, 5, -1.0, 1.0)
    R = 16
    U = np.random.default_rng(14).uniform(-1, 1, (30, R))
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    At = taco.init_taco_tensor_COO([80, 70, R])
    taco.ttm_compute(At, taco.csf_to_taco_tensor(csf), taco.CC_to_taco_tensor(formats.dense_as_dcsr(U), 30, R), 1)
    gk, gj, gl, gv = taco.taco_coo(At)
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, formats.dense_as_dcsr(U), 80, 70, R)
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gj, oj)
    np.testing.assert_array_equal(gl, ol)
    absY = np.einsum("ijk,kr->ijr", np.abs(dense(X)), np.abs(U))
    assert oracle.max_value_error(gv, ov, absY[ok, oj, ol]) <= TOL


@pytest.mark.parametrize("col_density", [1.0, 0.5])
def test_mttkrp_taco_hypersparse_factors(col_density):
    """DCSR factors with missing rows (and, at density 0.5, rows storing only
    some columns) through the taco entry point: the co-iteration of
    hash_free.h:277-300 on the device (factor row masks); entries, their
    order and values against the oracle."""
    print(f"case: mttkrp hypersparse factors density {col_density}", flush=True)
    X = gen.zipf_coo3((120, 90, 70), 3000, (1.5, 1.5, 1.5), 16, 5, 0.0, 1.0)
    R = 16
    rng = np.random.default_rng(16)
    Bf, Cf = rng.random((90, R)), rng.random((70, R))
    Bd = formats.partial_dcsr(Bf, rng.random(90) < 0.6, col_density, 161)
    Cd = formats.partial_dcsr(Cf, rng.random(70) < 0.7, col_density, 162)
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    At = taco.init_taco_tensor_COO([120, R])
    taco.mttkrp_compute(At, taco.csf_to_taco_tensor(csf), taco.CC_to_taco_tensor(Bd, 90, R),
                        taco.CC_to_taco_tensor(Cd, 70, R), 1 << 10)
    gi, gr, gv = taco.taco_coo(At)
    oi, orr, ov = oracle.mttkrp_hash(csf, Bd, Cd, 120, R, 1 << 10)
    assert len(oi) < 120 * R
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gr, orr)
    Bz, _ = formats.dcsr_dense_and_mask(Bd)
    Cz, _ = formats.dcsr_dense_and_mask(Cd)
    absM = np.einsum("ijk,jr,kr->ir", np.abs(dense(X)), np.abs(Bz), np.abs(Cz))
    assert oracle.max_value_error(gv, ov, absM[oi, orr]) <= TOL


@pytest.mark.parametrize("col_density", [1.0, 0.5])
def test_ttm_taco_hypersparse_factor(col_density):
    """A DCSR U with missing rows (and partial rows at density 0.5) through
    the taco entry point: hash_il's co-iteration of the contracted level with
    U's rows on the device; fibers, entries and values against the oracle."""
    print(f"case: ttm hypersparse factor density {col_density}", flush=True)
    X = gen.zipf_coo3((80, 70, 30), 2500, (1.5, 1.5, None), 17, 5, -1.0, 1.0)
    R = 16
    rng = np.random.default_rng(17)
    U = rng.uniform(-1, 1, (30, R))
    Ud = formats.partial_dcsr(U, rng.random(30) < 0.5, col_density, 171)
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    At = taco.init_taco_tensor_COO([80, 70, R])
    taco.ttm_compute(At, taco.csf_to_taco_tensor(csf), taco.CC_to_taco_tensor(Ud, 30, R), 1)
    gk, gj, gl, gv = taco.taco_coo(At)
    ok, oj, ol, ov = oracle.ttm_hash_il(csf, Ud, 80, 70, R)
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gj, oj)
    np.testing.assert_array_equal(gl, ol)
    Uz, _ = formats.dcsr_dense_and_mask(Ud)
    absY = np.einsum("ijk,kr->ijr", np.abs(dense(X)), np.abs(Uz))
    assert oracle.max_value_error(gv, ov, absY[ok, oj, ol]) <= TOL


def test_tensor_edge_cases(ctx):
    print("case: tensor edge cases", flush=True)
    X = gen.zipf_coo3((10, 9, 8), 1, (None, None, None), 15, 5, 0.5, 1.0)
    Bf, Cf = np.ones((9, 4)), np.ones((8, 4))
    rows, M = run_mttkrp(ctx, X, Bf, Cf)
    np.testing.assert_array_equal(rows, [X.idx[0, 0]])
    np.testing.assert_array_equal(M[X.idx[0, 0]], np.full(4, X.vals[0]))
    fi, fj, Y = run_ttm(ctx, X, np.ones((8, 3)))
    np.testing.assert_array_equal(fi, [X.idx[0, 0]])
    np.testing.assert_array_equal(fj, [X.idx[0, 1]])


def sample_rows_reference(X, Bf, Cf, rows):
    """M(rows, :) from the nonzeros of the sampled rows, and sum|terms|."""
    sel = np.isin(X.idx[:, 0], rows)
    xi, xj, xk, xv = X.idx[sel, 0], X.idx[sel, 1], X.idx[sel, 2], X.vals[sel]
    where = np.searchsorted(rows, xi)
    R = Bf.shape[1]
    ref = np.zeros((len(rows), R))
    ab = np.zeros((len(rows), R))
    t = (xv[:, None] * Bf[xj]) * Cf[xk]
    np.add.at(ref, where, t)
    np.add.at(ab, where, np.abs(t))
    return ref, ab


def test_mttkrp_full_workload(ctx):
    """50,000 x 40,000 x 30,000, 20M nonzeros, R = 32: present rows exact and
    256 sampled rows against a float reference."""
    print("case: mttkrp full workload (generating)", flush=True)
    X, Bf, Cf = gen.mttkrp_tensor()
    print("case: mttkrp full workload (running)", flush=True)
    rows, M = run_mttkrp(ctx, X, Bf, Cf)
    np.testing.assert_array_equal(rows, np.unique(X.idx[:, 0]))
    samp = np.sort(np.random.default_rng(5).choice(rows, 256, replace=False))
    ref, ab = sample_rows_reference(X, Bf, Cf, samp)
    assert oracle.max_value_error(M[samp].ravel(), ref.ravel(), ab.ravel()) <= TOL


def test_ttm_full_workload(ctx):
    """100,000 x 80,000 x 2,000, 10M nonzeros, R = 16: fiber list exact and
    sampled fibers against a float reference."""
    print("case: ttm full workload (generating)", flush=True)
    X, U = gen.ttm_tensor()
    print("case: ttm full workload (running)", flush=True)
    fi, fj, Y = run_ttm(ctx, X, U)
    key = X.idx[:, 0].astype(np.int64) * X.dims[1] + X.idx[:, 1]
    fib = np.unique(key)
    np.testing.assert_array_equal(fi.astype(np.int64) * X.dims[1] + fj, fib)
    samp = np.sort(np.random.default_rng(6).choice(len(fib), 512, replace=False))
    sel = np.isin(key, fib[samp])
    where = np.searchsorted(fib[samp], key[sel])
    t = X.vals[sel, None] * U[X.idx[sel, 2]]
    ref = np.zeros((len(samp), U.shape[1]))
    ab = np.zeros_like(ref)
    np.add.at(ref, where, t)
    np.add.at(ab, where, np.abs(t))
    assert oracle.max_value_error(Y[samp].ravel(), ref.ravel(), ab.ravel()) <= TOL
