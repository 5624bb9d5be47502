"""The workspace API (Sort / Merge of utils/wspace.h) on the GPU vs the
oracle, through the C ABI, and the reference-style C++ drivers over the
drop-in headers (include/spws/compat/) vs the oracle.

Integer fields and record order are compared bit-exact; a merged value is
one addition of two FP64 numbers, so values are compared bit-exact too.
"""
import subprocess

import numpy as np
import pytest

import oracle
from helpers import (build_compat_driver, powerlaw_rows, random_records, read_frames, sorted_coo, uniform_rows,
                     write_frames)
from sparsews_amd import formats
from sparsews_amd.wspace import WSPACE_DTYPE, Merge, Sort, empty_wspace

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_record_layout_matches_oracle():
    assert WSPACE_DTYPE == oracle.WSPACE_DTYPE


@pytest.mark.parametrize("n,rev,empty", [(1, False, 0.0), (777, False, 0.0), (777, True, 0.25),
                                         (200_000, False, 0.1), (200_000, True, 0.4)])
def test_sort_matches_oracle(ctx, n, rev, empty):
    print(f"case: wspace Sort n={n} rev={rev}", flush=True)
    rng = np.random.default_rng(n + int(rev))
    w = random_records(rng, n, 1 << 12, 1 << 12, empty_frac=empty)
    exp, cnt = oracle.wspace_sort(w, rev)
    got = w.copy()
    assert Sort(got, got.size, rev, ctx=ctx) == cnt
    real = exp["crd"][:, 0] != -1
    # real coordinates are distinct: every real record is in a fixed place
    np.testing.assert_array_equal(got[real], exp[real])
    # empty slots (crd0 = -1) may repeat a key: compare them as a multiset
    np.testing.assert_array_equal(np.sort(got[~real]["crd"][:, 1]), np.sort(exp[~real]["crd"][:, 1]))
    np.testing.assert_array_equal(got["crd"][:, 0] == -1, ~real)


def test_sort_all_empty(ctx):
    w = empty_wspace(64)
    w["crd"][:, 1] = np.arange(64)[::-1]
    assert Sort(w, 64, True, ctx=ctx) == 0
    assert np.all(w["crd"][:, 0] == -1)


@pytest.mark.parametrize("rev", [False, True])
@pytest.mark.parametrize("ncoo,nacc", [(0, 500), (500, 0), (3000, 2500), (150_000, 120_000)])
def test_merge_matches_oracle(ctx, rev, ncoo, nacc):
    print(f"case: wspace Merge coo={ncoo} acc={nacc} rev={rev}", flush=True)
    rng = np.random.default_rng(ncoo + 7 * nacc + int(rev))
    c0, c1, v = sorted_coo(rng, ncoo, 1 << 10, 1 << 10)
    w = random_records(rng, nacc, 1 << 10, 1 << 10)
    s, n = oracle.wspace_sort(w, rev)
    e0, e1, ev = oracle.wspace_merge(c0, c1, v, s[:n], rev)
    cap = ncoo + nacc
    g0, g1, gv = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap)
    g0[:ncoo], g1[:ncoo], gv[:ncoo] = c0, c1, v
    acc = s.copy()
    m = Merge((g0, g1), gv, ncoo, acc, n, rev, ctx=ctx)
    assert m == len(e0)
    np.testing.assert_array_equal(g0[:m], e0)
    np.testing.assert_array_equal(g1[:m], e1)
    np.testing.assert_array_equal(gv[:m], ev)


def test_merge_rejects_short_capacity(ctx):
    c = np.zeros(3, np.int32)
    with pytest.raises(ValueError):
        Merge((c, c.copy()), np.zeros(3), 2, empty_wspace(4), 4, ctx=ctx)


# --------------------------------------------------------------------------- drop-in header drivers

@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    d = tmp_path_factory.mktemp("compat")
    return {name: build_compat_driver(name, d) for name in ("drv_csr", "drv_coord", "drv_cscT")}


def run_driver(args, timeout=120):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])


def check_csr(pos, crd, vals, A, B):
    G = oracle.spgemm_gustavson(A, B)
    np.testing.assert_array_equal(pos, G.pos)
    np.testing.assert_array_equal(crd, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(vals, exact, absum) <= TOL


@pytest.mark.parametrize("mode", ["hash", "split"])
def test_csr_driver(drivers, tmp_path, mode):
    print(f"case: drop-in driver {mode}", flush=True)
    A, B = powerlaw_rows(900, 700, 1.8, 300, 70, 0), powerlaw_rows(700, 800, 1.8, 300, 70, 1)
    write_frames(tmp_path / "in", [np.array([A.nrows, A.ncols, B.ncols], np.int32), A.pos, A.crd, A.vals,
                                   B.pos, B.crd, B.vals])
    run_driver([drivers["drv_csr"], mode, str(tmp_path / "in"), str(tmp_path / "out")])
    pos, crd, vals = read_frames(tmp_path / "out", [np.int32, np.int32, np.float64])
    check_csr(pos, crd, vals, A, B)


@pytest.mark.parametrize("rev", [0, 1])
def test_wspace_driver(drivers, tmp_path, rev):
    print(f"case: drop-in driver wspace rev={rev}", flush=True)
    rng = np.random.default_rng(80 + rev)
    w = random_records(rng, 4000, 300, 300, empty_frac=0.2 if rev else 0.0)
    c0, c1, v = sorted_coo(rng, 3000, 300, 300)
    write_frames(tmp_path / "in", [np.array([rev], np.int32), w, c0, c1, v])
    run_driver([drivers["drv_csr"], "wspace", str(tmp_path / "in"), str(tmp_path / "out")])
    sizes, srt, m0, m1, mv = read_frames(tmp_path / "out", [np.int32, WSPACE_DTYPE, np.int32, np.int32, np.float64])
    exp, cnt = oracle.wspace_sort(w, bool(rev))
    assert sizes[0] == cnt
    real = exp["crd"][:, 0] != -1
    np.testing.assert_array_equal(srt[real], exp[real])
    e0, e1, ev = oracle.wspace_merge(c0, c1, v, exp[:cnt], bool(rev))
    assert sizes[1] == len(e0)
    np.testing.assert_array_equal(m0, e0)
    np.testing.assert_array_equal(m1, e1)
    np.testing.assert_array_equal(mv, ev)


def test_coord_driver(drivers, tmp_path):
    print("case: drop-in driver outer coord", flush=True)
    A, B = uniform_rows(500, 400, 0, 20, 71, 0), powerlaw_rows(400, 600, 1.8, 200, 71, 1)
    Ac, Bd = formats.dcsc_from_csr(A), formats.dcsr_from_csr(B)
    write_frames(tmp_path / "in", [np.array([A.nrows, A.ncols, B.ncols], np.int32),
                                   Ac.l1_crd, Ac.l2_pos, Ac.l2_crd, Ac.vals, Bd.l1_crd, Bd.l2_pos, Bd.l2_crd, Bd.vals])
    run_driver([drivers["drv_coord"], str(tmp_path / "in"), str(tmp_path / "out")])
    pos, crd, vals = read_frames(tmp_path / "out", [np.int32, np.int32, np.float64])
    check_csr(pos, crd, vals, A, B)


def test_nonfused_transposed_driver(drivers, tmp_path):
    """The reference's non-fused transposed-output driver call
    (CSC_CSR_T_hash_nonfuse: compute_nofuse, then transpose) through the
    drop-in header: C = (A B)^T equals the oracle's Gustavson B^T A^T bit
    for bit."""
    print("case: drop-in driver non-fused transposed output", flush=True)
    A, B = powerlaw_rows(400, 300, 1.8, 150, 74, 0), uniform_rows(300, 500, 0, 20, 74, 1)
    Ac = A.transpose()  # the columns of A as rows: its CSC arrays
    write_frames(tmp_path / "in", [np.array([A.nrows, A.ncols, B.ncols], np.int32), Ac.pos, Ac.crd, Ac.vals,
                                   B.pos, B.crd, B.vals])
    run_driver([drivers["drv_cscT"], str(tmp_path / "in"), str(tmp_path / "out")])
    pos, crd, vals = read_frames(tmp_path / "out", [np.int32, np.int32, np.float64])
    G = oracle.spgemm_gustavson(B.transpose(), A.transpose())
    np.testing.assert_array_equal(np.asarray(pos, np.int64), np.asarray(G.pos, np.int64))
    np.testing.assert_array_equal(crd, G.crd)
    np.testing.assert_array_equal(vals, G.vals)
