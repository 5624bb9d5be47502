"""Outer-product SpGEMM (DCSC x DCSR, expand / radix sort / segmented reduce)
on the GPU vs the CPU oracle, through the C ABI, with both sort methods
(1: LSD radix sort of packed keys; 2: row/column-window buckets).

Structure must be bit-exact.  Values: the device sums each entry over j in
increasing order, the order of the oracle's Gustavson restatement, so the
values are compared bit-for-bit as well as against the 1e-12 bar.
"""
import contextlib

import numpy as np
import pytest

import oracle
from helpers import powerlaw_rows, random_csr, uniform_rows
from sparsews_amd import formats, gen
from sparsews_amd.device import DeviceArray, DeviceDCSR, set_outer_chunking, set_outer_method, spgemm_outer
from sparsews_amd.gen import CSR

pytestmark = pytest.mark.gpu
TOL = 1e-12
METHODS = [1, 2, 3, 4]
# methods 1-3 sum every entry in increasing j (bit-identical to the oracle's
# Gustavson); method 4 adds the products into the LDS window atomically, so
# its sums are reassociated (within the tolerance only)
UNORDERED = {4}
_current = {"method": 0}


@contextlib.contextmanager
def outer_method(ctx, method, window_log2=0):
    set_outer_method(ctx, method, window_log2)
    _current["method"] = method
    try:
        yield
    finally:
        set_outer_method(ctx, 0, 0)
        _current["method"] = 0


def run_outer(ctx, A, B, dense_outer=False):
    Ac, Bd = formats.dcsc_from_csr(A), formats.dcsr_from_csr(B)
    dA = DeviceDCSR.upload(ctx, Ac, dense_outer=False)
    dB = DeviceDCSR.upload(ctx, Bd, dense_outer=False)
    if dense_outer:
        dA = DeviceDCSR.from_csr(ctx, A.transpose())
        dB = DeviceDCSR.from_csr(ctx, B)
    return spgemm_outer(ctx, dA, dB, A.nrows, A.ncols, B.ncols).download()


def check(C, A, B, exact_order=None, nthreads=0):
    if exact_order is None:
        exact_order = _current["method"] not in UNORDERED
    G = oracle.spgemm_gustavson(A, B, nthreads)
    assert C.nnz == G.nnz
    np.testing.assert_array_equal(np.asarray(C.pos, np.int64), np.asarray(G.pos, np.int64))
    np.testing.assert_array_equal(C.crd, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd, nthreads)
    assert oracle.max_value_error(C.vals, exact, absum) <= TOL
    if exact_order:
        np.testing.assert_array_equal(C.vals, G.vals)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("seed", [60, 61])
def test_outer_powerlaw(ctx, seed, method):
    print(f"case: outer power-law seed {seed} method {method}", flush=True)
    A, B = powerlaw_rows(600, 500, 1.8, 300, seed, 0), powerlaw_rows(500, 700, 1.8, 300, seed, 1)
    with outer_method(ctx, method):
        C = run_outer(ctx, A, B)
        assert ctx.stats()["outer_method"] == method
    check(C, A, B)


@pytest.mark.parametrize("method", [3, 4])
def test_outer_fused_wide_offsets(ctx, method, monkeypatch):
    """The fused kernels address B with 32-bit byte offsets while B has
    fewer than 2^29 entries; SPWS_OUTER_WIDE_OFFSETS=1 makes them take the
    64-bit addresses that larger B needs.  Both variants against the
    oracle, and against each other bit for bit (method 3, in order)."""
    print(f"case: outer fused 64-bit offsets method {method}", flush=True)
    A, B = powerlaw_rows(600, 500, 1.8, 300, 65, 0), powerlaw_rows(500, 700, 1.8, 300, 65, 1)
    with outer_method(ctx, method):
        C32 = run_outer(ctx, A, B)
        monkeypatch.setenv("SPWS_OUTER_WIDE_OFFSETS", "1")
        C64 = run_outer(ctx, A, B)
        monkeypatch.delenv("SPWS_OUTER_WIDE_OFFSETS")
        assert ctx.stats()["outer_method"] == method
        check(C64, A, B)
        check(C32, A, B)
    if method == 3:
        np.testing.assert_array_equal(C32.vals, C64.vals)


@pytest.mark.parametrize("method", [3, 4])
def test_outer_fused_dense_windows(ctx, method):
    """Windows of 2^11 columns holding more entries than the emit lists at
    once (1,024): those windows are listed and written in two phases, lanes
    0-31's columns then lanes 32-63's.  Rows of every density, so that
    single-phase and two-phase windows (and phases with no entries) mix."""
    print(f"case: outer fused dense windows method {method}", flush=True)
    A = uniform_rows(300, 400, 0, 30, 66, 0)
    B = uniform_rows(400, 4096, 0, 400, 66, 1)
    G = oracle.spgemm_gustavson(A, B)
    per_window = np.bincount(np.repeat(np.arange(A.nrows), np.diff(G.pos)) * 2 + G.crd // 2048,
                             minlength=2 * A.nrows)
    assert (per_window > 1024).sum() > 50 and ((per_window > 0) & (per_window <= 1024)).sum() > 50
    with outer_method(ctx, method, 11):
        C = run_outer(ctx, A, B)
        assert ctx.stats()["outer_method"] == method
        assert ctx.stats()["window_bits"] == 11
        check(C, A, B)


@pytest.mark.parametrize("method", METHODS)
def test_outer_dense_outer_levels(ctx, method):
    print(f"case: outer CSC x CSR method {method}", flush=True)
    A, B = uniform_rows(300, 200, 0, 15, 62, 0), uniform_rows(200, 260, 0, 15, 62, 1)
    with outer_method(ctx, method):
        C = run_outer(ctx, A, B, dense_outer=True)
    check(C, A, B)


@pytest.mark.parametrize("method", METHODS)
def test_outer_many_chunks(ctx, method):
    print(f"case: outer many chunks method {method}", flush=True)
    A, B = powerlaw_rows(700, 600, 1.8, 300, 63, 0), powerlaw_rows(600, 650, 1.8, 300, 63, 1)
    try:
        set_outer_chunking(ctx, 3000, 7)
        with outer_method(ctx, method):
            C = run_outer(ctx, A, B)
            chunks = ctx.stats()["chunks"]
    finally:
        set_outer_chunking(ctx, 0, 0)
    # the fused methods (3, 4) need no product workspace, so they never chunk
    assert chunks == 1 if method in (3, 4) else chunks >= 100
    check(C, A, B)


@pytest.mark.parametrize("method", METHODS)
def test_outer_cc_gaps(ctx, method):
    print(f"case: outer co-iteration gaps method {method}", flush=True)
    A = uniform_rows(80, 120, 0, 8, 64, 0)
    B = uniform_rows(120, 90, 0, 8, 64, 1)
    lens = B.row_lengths().copy()
    lens[::3] = 0
    B2 = random_csr(120, 90, lens, 64, 5)
    with outer_method(ctx, method):
        C = run_outer(ctx, A, B2)
    check(C, A, B2)


@pytest.mark.parametrize("method", METHODS)
def test_outer_edge_cases(ctx, method):
    print(f"case: outer edge cases method {method}", flush=True)
    A = CSR(6, 5, np.zeros(7, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    B = uniform_rows(5, 9, 1, 4, 65, 1)
    with outer_method(ctx, method):
        C = run_outer(ctx, A, B)
    assert C.nnz == 0 and np.all(C.pos == 0) and len(C.pos) == 7
    n = 150
    I = CSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    Bm = uniform_rows(n, 130, 0, 20, 66, 1)
    with outer_method(ctx, method):
        CI = run_outer(ctx, I, Bm)
    np.testing.assert_array_equal(CI.pos, Bm.pos)
    np.testing.assert_array_equal(CI.crd, Bm.crd)
    np.testing.assert_array_equal(CI.vals, Bm.vals)


@pytest.mark.parametrize("method", [2, 3])
@pytest.mark.parametrize("wlog", [6, 8, 11])
def test_outer_bucket_windows(ctx, wlog, method):
    """Window widths from 64 to 2048 columns; the column count is not a
    multiple of the window, and some B rows run over many windows."""
    print(f"case: outer window 2^{wlog} method {method}", flush=True)
    A = powerlaw_rows(500, 400, 1.8, 250, 68, 0)
    B = powerlaw_rows(400, 3001, 1.8, 900, 68, 1)
    with outer_method(ctx, method, wlog):
        C = run_outer(ctx, A, B)
        st = ctx.stats()
    assert st["outer_method"] == method and st["window_bits"] == wlog
    check(C, A, B)


@pytest.mark.parametrize("method", [2, 3])
def test_outer_bucket_wide_window_table(ctx, method):
    """More than 1024 windows per row (64-column windows over 80,000
    columns): the bucket expand reads each product's bucket slot from the
    global window table instead of a per-entry copy in LDS; the fused walk
    crosses 1,250 windows per row, most of them empty."""
    print(f"case: outer 1250 windows per row method {method}", flush=True)
    A = powerlaw_rows(300, 400, 1.8, 200, 71, 0)
    B = powerlaw_rows(400, 80000, 1.8, 3000, 71, 1)
    with outer_method(ctx, method, 6):
        C = run_outer(ctx, A, B)
        st = ctx.stats()
    assert st["outer_method"] == method and st["window_bits"] == 6
    check(C, A, B)


@pytest.mark.parametrize("method", [2, 3])
def test_outer_bucket_long_column_runs(ctx, method):
    """Outputs where one entry gathers many products (dense rows of A against
    a B whose rows all share a few columns), so one 64-lane stride holds
    several products of the same column and of different j; rows of A have
    150 to 300 entries, so the fused walk takes each window's entries in
    chunks of 64."""
    print(f"case: outer repeated columns method {method}", flush=True)
    rng = np.random.default_rng(69)
    A = uniform_rows(90, 300, 150, 300, 69, 0)
    lens = rng.integers(1, 6, 300)
    cols = np.concatenate([np.sort(rng.choice(8, size=l, replace=False)) for l in lens]).astype(np.int32)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    B = CSR(300, 8, pos, cols, rng.uniform(-1, 1, len(cols)))
    with outer_method(ctx, method, 6):
        C = run_outer(ctx, A, B)
    check(C, A, B)


@pytest.mark.parametrize("wlog", [6, 11])
def test_outer_fused_rows_over_many_chunks(ctx, wlog):
    """Fused method: rows with 64 to 700 entries (one to eleven chunks of 64
    per window) next to rows with none, against the bucket method bit for
    bit and the oracle."""
    print(f"case: outer fused long rows window 2^{wlog}", flush=True)
    rng = np.random.default_rng(72)
    lens = rng.integers(0, 700, 120)
    lens[::5] = 0
    lens[1] = 64
    lens[2] = 65
    A = random_csr(120, 900, lens, 72, 0)
    B = powerlaw_rows(900, 5000, 1.8, 400, 72, 1)
    with outer_method(ctx, 3, wlog):
        C3 = run_outer(ctx, A, B)
    with outer_method(ctx, 2, wlog):
        C2 = run_outer(ctx, A, B)
    np.testing.assert_array_equal(C3.pos, C2.pos)
    np.testing.assert_array_equal(C3.crd, C2.crd)
    np.testing.assert_array_equal(C3.vals, C2.vals)
    check(C3, A, B)


def test_outer_auto_method(ctx):
    """Automatic choice: the fused row windows (atomic adds) for the
    power-law workload shape, the LSD sort for very wide, very sparse
    outputs."""
    print("case: outer automatic method", flush=True)
    B = gen.zipf_csr(4000, 1.8, 400, seed=3, stream=2)
    dB = DeviceDCSR.upload(ctx, formats.dcsr_from_csr(B))
    spgemm_outer(ctx, dB, dB, 4000, 4000, 4000)
    assert ctx.stats()["outer_method"] == 4
    A = uniform_rows(100, 80, 0, 3, 70, 0)
    Bw = random_csr(80, 1 << 22, np.full(80, 3), 70, 1)
    C = run_outer(ctx, A, Bw)
    assert ctx.stats()["outer_method"] == 1
    check(C, A, Bw)


def test_outer_wide_columns_64bit_keys(ctx):
    """ncols > 2^24 switches to 64-bit sort keys."""
    print("case: outer 64-bit keys", flush=True)
    n = (1 << 24) + 123
    A = uniform_rows(200, 150, 0, 10, 67, 0)
    B = random_csr(150, n, np.random.default_rng(67).integers(0, 12, 150), 67, 1)
    with outer_method(ctx, 1):
        C = run_outer(ctx, A, B)
        assert ctx.stats()["sort_bits"] > 32
    check(C, A, B, nthreads=1)


@pytest.mark.parametrize("method", METHODS)
def test_outer_workload_scaled(ctx, method):
    """The outer-product generator at N = 20,000, cap 2,000, A = B^T."""
    print(f"case: outer scaled workload method {method}", flush=True)
    B = gen.zipf_csr(20_000, 1.8, 2_000, seed=1, stream=2)
    A = B.transpose()
    dA = DeviceDCSR.upload(ctx, formats.dcsr_from_csr(B))  # CSC(A) = B's arrays
    dB = DeviceDCSR.upload(ctx, formats.dcsr_from_csr(B))
    with outer_method(ctx, method):
        C = spgemm_outer(ctx, dA, dB, 20_000, 20_000, 20_000).download()
    check(C, A, B)


def row_sums_chunked(C, chunk=1 << 27):
    """Row sums of a device CSR with int64 row pointers, its values copied
    to the host a chunk at a time (a row split by a chunk boundary sums its
    parts)."""
    pos = C.pos.to_host().astype(np.int64)
    out = np.zeros(C.nrows)
    for c0 in range(0, C.nnz, chunk):
        c1 = min(c0 + chunk, C.nnz)
        v = DeviceArray.view(C.vals.ctx, C.vals.ptr + 8 * c0, c1 - c0, np.float64).to_host()
        r_lo = int(np.searchsorted(pos, c0, "right")) - 1
        r_hi = int(np.searchsorted(pos, c1, "left"))
        rows = np.arange(r_lo, r_hi)
        beg = np.maximum(pos[rows], c0) - c0
        end = np.minimum(pos[rows + 1], c1) - c0
        keep = end > beg
        out[rows[keep]] += np.add.reduceat(v, beg[keep])
    return pos, out


def test_outer_workload_full(ctx):
    """The 200,000 x 200,000, Zipf 1.8 workload at full size (1.16e10
    products, 1.0e10 entries): every row's entry count against the oracle's
    int64 symbolic count (bit-exact), every row sum against A (B 1) (the sum
    of a row of C is the row of A times B's row sums; within the rounding
    bound n eps sum|terms| of reordering n terms, 1e-10 sum|terms| here), the
    total, and 64 sampled rows entry by entry against the oracle."""
    print("case: outer full workload", flush=True)
    B = gen.outer_operand()
    L = B.row_lengths()
    Bd = formats.dcsr_from_csr(B)
    dA, dB = DeviceDCSR.upload(ctx, Bd), DeviceDCSR.upload(ctx, Bd)
    N = B.nrows
    C = spgemm_outer(ctx, dA, dB, N, N, N)
    st = ctx.stats()
    assert st["products"] == int((L * L).sum())
    AT = B.transpose()
    cnt, total = oracle.spgemm_row_counts(AT, B)
    assert C.nnz == total
    pos, rs = row_sums_chunked(C)
    assert pos[0] == 0 and pos[-1] == C.nnz
    np.testing.assert_array_equal(np.diff(pos), cnt.astype(np.int64))
    Sa, Sb = AT.to_scipy(), B.to_scipy()
    ref = Sa @ (Sb @ np.ones(N))
    absum = abs(Sa) @ (abs(Sb) @ np.ones(N))
    assert np.all(np.abs(rs - ref) <= 1e-10 * absum)
    assert abs(rs.sum() - ref.sum()) <= 1e-10 * absum.sum()
    rng = np.random.default_rng(123)
    rows = np.sort(rng.choice(N, 64, replace=False))
    _, got = C.rows(rows)
    # oracle on the sampled rows of A = B^T
    lens = AT.row_lengths()[rows]
    sub_pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    sub_crd = np.concatenate([AT.crd[AT.pos[r]:AT.pos[r + 1]] for r in rows])
    sub_val = np.concatenate([AT.vals[AT.pos[r]:AT.pos[r + 1]] for r in rows])
    Asub = CSR(len(rows), N, sub_pos, sub_crd, sub_val)
    G = oracle.spgemm_gustavson(Asub, B)
    exact, absum = oracle.spgemm_refsum(Asub, B, G.pos, G.crd)
    for t, (crd, val) in enumerate(got):
        sl = slice(G.pos[t], G.pos[t + 1])
        np.testing.assert_array_equal(crd, G.crd[sl])
        if st["outer_method"] not in UNORDERED:  # in-order sums: the oracle's bit for bit
            np.testing.assert_array_equal(val, G.vals[sl])
        assert oracle.max_value_error(val, exact[sl], absum[sl]) <= TOL
    print(f"outer full: nnz(C) = {C.nnz}, products = {st['products']}, chunks = {st['chunks']}", flush=True)


@pytest.mark.parametrize("w_cap", [1 << 10, 1 << 16])
def test_csc_csr_transposed_output(w_cap):
    """C = (A B)^T through the reference's CSC x CSR transposed-output entry
    (CSC_CSR_T_hash.h:232): the entries of row k of C are summed over j in
    increasing order, so C equals the oracle's Gustavson product B^T A^T bit
    for bit; it also equals scipy's (A B)^T within the tolerance."""
    print(f"case: CSC x CSR transposed output w_cap={w_cap}", flush=True)
    import scipy.sparse as sp
    from sparsews_amd import taco
    A = powerlaw_rows(300, 250, 1.8, 120, 72, 0)   # m x k
    B = powerlaw_rows(250, 400, 1.8, 150, 72, 1)   # k x n
    At = taco.DC_to_taco_tensor_csc(A.transpose())
    Bt = taco.DC_to_taco_tensor(B)
    Ct = taco.init_taco_tensor_DC(B.ncols, A.nrows)
    taco.csc_csr_t_compute(Ct, At, Bt, w_cap)
    CT = taco.taco_to_csr(Ct)
    taco.free_outputs(Ct)
    G = oracle.spgemm_gustavson(B.transpose(), A.transpose())
    np.testing.assert_array_equal(np.asarray(CT.pos, np.int64), np.asarray(G.pos, np.int64))
    np.testing.assert_array_equal(CT.crd, G.crd)
    np.testing.assert_array_equal(CT.vals, G.vals)
    sa = sp.csr_matrix((A.vals, A.crd, A.pos), shape=(A.nrows, A.ncols))
    sb = sp.csr_matrix((B.vals, B.crd, B.pos), shape=(B.nrows, B.ncols))
    ref = (sa @ sb).T.toarray()
    rows = np.repeat(np.arange(CT.nrows), np.diff(np.asarray(CT.pos, np.int64)))
    dense = np.zeros((CT.nrows, CT.ncols))
    dense[rows, CT.crd] = CT.vals
    np.testing.assert_allclose(dense, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("w_cap", [1 << 10, 1 << 16])
def test_csc_csr_nonfused_transposed_output(w_cap):
    """The reference's non-fused variant (nonfuse/CSC_CSR_T_hash.h:232 and
    :352): C_noT = A B equals the oracle's Gustavson product bit for bit
    (entries summed over j in increasing order), and transpose(C_noT)
    equals the fused entry's C = (A B)^T bit for bit."""
    print(f"case: CSC x CSR non-fused transposed output w_cap={w_cap}", flush=True)
    from sparsews_amd import taco
    A = powerlaw_rows(300, 250, 1.8, 120, 73, 0)   # m x k
    B = powerlaw_rows(250, 400, 1.8, 150, 73, 1)   # k x n
    At = taco.DC_to_taco_tensor_csc(A.transpose())
    Bt = taco.DC_to_taco_tensor(B)
    Cn = taco.init_taco_tensor_DC(A.nrows, B.ncols)
    taco.csc_csr_nofuse_compute(Cn, At, Bt, w_cap)
    CN = taco.taco_to_csr(Cn)
    G = oracle.spgemm_gustavson(A, B)
    np.testing.assert_array_equal(np.asarray(CN.pos, np.int64), np.asarray(G.pos, np.int64))
    np.testing.assert_array_equal(CN.crd, G.crd)
    np.testing.assert_array_equal(CN.vals, G.vals)
    Ct = taco.init_taco_tensor_DC(B.ncols, A.nrows)
    taco.transpose(Ct, Cn)
    CT = taco.taco_to_csr(Ct)
    Cf = taco.init_taco_tensor_DC(B.ncols, A.nrows)
    taco.csc_csr_t_compute(Cf, At, Bt, w_cap)
    CF = taco.taco_to_csr(Cf)
    for t in (Cn, Ct, Cf):
        taco.free_outputs(t)
    np.testing.assert_array_equal(np.asarray(CT.pos, np.int64), np.asarray(CF.pos, np.int64))
    np.testing.assert_array_equal(CT.crd, CF.crd)
    np.testing.assert_array_equal(CT.vals, CF.vals)


@pytest.mark.parametrize("m,n,lo,hi,seed", [(500, 700, 0, 30, 1), (1, 1, 1, 1, 2), (300, 70_000, 0, 50, 3),
                                            (40, 5, 0, 5, 4), (200, 300, 0, 0, 5), (0, 10, 0, 0, 6)])
def test_device_transpose(ctx, m, n, lo, hi, seed):
    """The device CSR transpose equals scipy's (counting-sort) transpose bit
    for bit: row pointers, columns in increasing order, values; empty rows,
    columns and matrices included."""
    print(f"case: device transpose m={m} n={n} per row {lo}..{hi}", flush=True)
    import scipy.sparse as sp
    from sparsews_amd.device import DeviceCSR, csr_transpose
    A = uniform_rows(m, n, lo, hi, seed, 0) if m > 0 else CSR(0, n, np.zeros(1, np.int32), np.zeros(0, np.int32),
                                                                np.zeros(0))
    T = csr_transpose(ctx, DeviceCSR.upload(ctx, A)).download()
    ref = sp.csr_matrix((A.vals, A.crd, A.pos), shape=(m, n)).T.tocsr()
    ref.sort_indices()
    assert (T.nrows, T.ncols, T.nnz) == (n, m, A.nnz)
    np.testing.assert_array_equal(np.asarray(T.pos, np.int64), ref.indptr.astype(np.int64))
    np.testing.assert_array_equal(T.crd, ref.indices)
    np.testing.assert_array_equal(T.vals, ref.data)
    O = oracle.csr_transpose(A)  # the restatement of the reference's transpose()
    np.testing.assert_array_equal(T.pos, O.pos)
    np.testing.assert_array_equal(T.crd, O.crd)
    np.testing.assert_array_equal(T.vals, O.vals)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("m,k,n,seed", [(300, 250, 400, 81), (2000, 1500, 1000, 82), (50, 3000, 70_000, 83)])
def test_device_fused_vs_nonfused_transposed_output(ctx, method, m, k, n, seed):
    """The transposed-output product with every array on the device, both
    ways: fused (the outer product B^T A^T, which leaves (A B)^T in row
    order) and non-fused (the outer product A B, then the transpose of its
    int64-pointer result).  Both sum every entry over j in increasing order,
    so they agree bit for bit, and with the oracle's Gustavson B^T A^T."""
    print(f"case: device fused vs non-fused transposed output method={method} {m}x{k}x{n}", flush=True)
    from sparsews_amd.device import csr64_transpose
    A = powerlaw_rows(m, k, 1.8, 120, seed, 0)   # m x k (stored here as CSC: A^T's CSR arrays)
    B = powerlaw_rows(k, n, 1.8, 150, seed, 1)   # k x n
    set_outer_method(ctx, method)
    try:
        dAc = DeviceDCSR.from_csr(ctx, A.transpose())   # A's columns j = stored outer level
        dB = DeviceDCSR.from_csr(ctx, B)
        F = spgemm_outer(ctx, dB, dAc, n, k, m).download()        # (A B)^T = B^T A^T
        N = csr64_transpose(ctx, spgemm_outer(ctx, dAc, dB, m, k, n)).download()
    finally:
        set_outer_method(ctx, 0)
    G = oracle.spgemm_gustavson(B.transpose(), A.transpose())
    _, absum = oracle.spgemm_refsum(B.transpose(), A.transpose(), G.pos, G.crd)
    for X in (F, N):
        assert (X.nrows, X.ncols) == (n, m)
        np.testing.assert_array_equal(np.asarray(X.pos, np.int64), np.asarray(G.pos, np.int64))
        np.testing.assert_array_equal(X.crd, G.crd)
        if method in UNORDERED:  # atomically added: within the tolerance
            assert oracle.max_value_error(X.vals, G.vals, absum) <= TOL
        else:
            np.testing.assert_array_equal(X.vals, G.vals)
