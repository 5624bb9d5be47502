"""Row-wise hash SpGEMM on the GPU vs the CPU oracle, through the C ABI.

Bar: structure (row pointers, column indices, nnz) bit-exact; FP64 values
within |x - y| <= 1e-12 * max(|y|, sum|terms|) of both the oracle's
reference-order result and its long-double sorted sum.
"""
import numpy as np
import pytest

import oracle
from helpers import powerlaw_rows, random_csr, uniform_rows
from sparsews_amd import taco
from sparsews_amd.device import DeviceCSR, spgemm_rowwise, spgemm_rowwise_symbolic
from sparsews_amd.gen import CSR, rowwise_pair

pytestmark = pytest.mark.gpu
TOL = 1e-12


def run(ctx, A, B):
    dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
    return spgemm_rowwise(ctx, dA, dB).download()


def check(C, A, B, w_cap=None):
    G = oracle.spgemm_gustavson(A, B)
    assert C.nnz == G.nnz
    np.testing.assert_array_equal(C.pos, G.pos)
    np.testing.assert_array_equal(C.crd, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(C.vals, G.vals, absum) <= TOL
    assert oracle.max_value_error(C.vals, exact, absum) <= TOL
    if w_cap is not None:
        H = oracle.spgemm_hash(A, B, w_cap)
        assert oracle.max_value_error(C.vals, H.vals, absum) <= TOL


def test_rowwise_workload_10k(ctx):
    print("case: 10k x 10k, 8 per row, seed 0", flush=True)
    A, B = rowwise_pair()
    C = run(ctx, A, B)
    check(C, A, B, w_cap=1 << 16)
    st = ctx.stats()
    assert st["products"] == 640_000
    assert st["rows_per_path"][0] == 10_000


@pytest.mark.parametrize("m,seed", [(1, 11), (63, 12), (64, 13), (65, 14), (3000, 15), (20000, 16)])
def test_single_pass_short_rows(ctx, m, seed):
    """Every row has ub <= 64: k_row_wave writes each row into a fixed
    64-entry slot region and one compaction (k_compact_fixed, tiles of 64
    rows) forms C.  Includes empty rows and row counts around the tile size."""
    print(f"case: single-pass m={m} seed={seed}", flush=True)
    A = uniform_rows(m, 700, 0, 8, seed, 0)
    B = uniform_rows(700, 900, 0, 8, seed, 1)
    assert oracle.spgemm_row_ub(A, B).max() <= 64
    C = run(ctx, A, B)
    check(C, A, B)
    st = ctx.stats()
    assert st["rows_per_path"][0] + st["rows_per_path"][6] == m
    dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
    np.testing.assert_array_equal(spgemm_rowwise_symbolic(ctx, dA, dB), C.pos)


def test_fixed_slot_path_launched_ahead(ctx):
    """After a product that took the fixed-slot path, the next product on the
    context launches that path before its metadata reaches the host (rows
    over 64 products skipped); the result is kept only when every row fits,
    else the general path recomputes C.  A sequence of fitting and
    non-fitting products, numeric and symbolic, against the oracle."""
    print("case: fixed-slot path launched ahead", flush=True)
    short = [(uniform_rows(m, 700, 0, 8, 80 + m % 7, 0), uniform_rows(700, 900, 0, 8, 80 + m % 7, 1))
             for m in (3000, 64, 20000)]
    # one row of 20 x 8 = 160 products among rows of 32: not fitting
    longA = random_csr(2000, 700, np.where(np.arange(2000) == 1234, 20, 4), 91, 0)
    longB = random_csr(700, 900, np.full(700, 8), 91, 1)
    assert oracle.spgemm_row_ub(longA, longB).max() > 64
    seq = [short[0], (longA, longB), short[1], short[2], (longA, longB), (longA, longB), short[0]]
    for t, (A, B) in enumerate(seq):
        C = run(ctx, A, B)
        check(C, A, B)
        dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
        np.testing.assert_array_equal(spgemm_rowwise_symbolic(ctx, dA, dB), C.pos)
        fits = oracle.spgemm_row_ub(A, B).max() <= 64
        st = ctx.stats()
        assert (st["rows_per_path"][0] + st["rows_per_path"][6] == A.nrows) == fits, t


@pytest.mark.parametrize("m", [1, 63, 10_000, 40_000, 130_000])
def test_short_row_path(ctx, m):
    """Every row has at most 64 products: two launches, each wavefront of
    the first owning up to 8 consecutive rows (130,000 rows give 8 per
    wavefront on a 256-CU device) and packing their sorted entries, the
    second summing the earlier wavefronts' totals and copying into C.  Empty
    rows among them; three calls in a row (the second and third start on the
    short-row path with no upper-bound pass), each against the oracle, with
    the product count."""
    print(f"case: short-row path, {m} rows", flush=True)
    A = uniform_rows(m, 5000, 0, 8, 31, 0)
    B = uniform_rows(5000, 6000, 0, 8, 31, 1)
    ub = oracle.spgemm_row_ub(A, B)
    assert ub.max() <= 64
    try:
        ctx.set_profiling(True)
        for rep in range(3):
            ctx.reset_kernel_times()
            C = run(ctx, A, B)
            check(C, A, B)
            assert ctx.stats()["products"] == int(ub.sum()), rep
            times = ctx.kernel_times()
            assert "short_rows" in times and "short_compact" in times and "row_wave" not in times, (rep, times)
            if rep > 0:
                assert "rowstat" not in times, times
    finally:
        ctx.set_profiling(False)


def test_short_row_path_wide_columns(ctx):
    """B with 2^26 + 1000 columns: a packed sort key (column and A-entry
    index in 32 bits) no longer fits, so the short-row kernel sorts
    (column, value) pairs instead.  B's columns come from a pool of about
    400 values that includes 0, both sides of 2^26 and the last column, so
    rows repeat columns and the run sums are exercised.  Two calls against
    the single-threaded oracle (its dense workspace is one entry per
    column), with the product count."""
    print("case: short-row path, 3000 rows, 2^26 + 1000 columns", flush=True)
    n = (1 << 26) + 1000
    rng = np.random.default_rng(77)
    pool = np.unique(np.concatenate([[0, 1, (1 << 26) - 2, (1 << 26) - 1, 1 << 26, n - 1],
                                     rng.integers(0, n, 400)])).astype(np.int32)
    kdim = 200
    lengths = rng.integers(0, 9, kdim)
    cols = [np.sort(rng.choice(pool, int(L), replace=False)) for L in lengths]
    pos = np.zeros(kdim + 1, dtype=np.int32)
    np.cumsum(lengths, out=pos[1:])
    B = CSR(kdim, n, pos, np.concatenate(cols).astype(np.int32), rng.uniform(-1.0, 1.0, int(pos[-1])))
    A = uniform_rows(3000, kdim, 0, 8, 57, 0)
    ub = oracle.spgemm_row_ub(A, B)
    assert ub.max() <= 64
    G = oracle.spgemm_gustavson(A, B, nthreads=1)
    assert int(G.crd.max()) >= 1 << 26
    assert G.nnz < int(ub.sum())  # repeated columns within rows
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    try:
        ctx.set_profiling(True)
        for rep in range(2):
            ctx.reset_kernel_times()
            C = run(ctx, A, B)
            np.testing.assert_array_equal(C.pos, G.pos)
            np.testing.assert_array_equal(C.crd, G.crd)
            assert oracle.max_value_error(C.vals, G.vals, absum) <= TOL
            assert oracle.max_value_error(C.vals, exact, absum) <= TOL
            assert ctx.stats()["products"] == int(ub.sum()), rep
            assert "short_rows" in ctx.kernel_times(), rep
    finally:
        ctx.set_profiling(False)


def test_short_rows_beyond_short_path(ctx):
    """300,000 short rows need more first-kernel wavefronts than the
    compaction sums: the short-row path declines before launching and the
    general pipeline computes the product, twice in a row, each against the
    oracle."""
    print("case: short rows beyond the short-row path, 300k rows", flush=True)
    A = uniform_rows(300_000, 5000, 0, 8, 31, 0)
    B = uniform_rows(5000, 6000, 0, 8, 31, 1)
    ub = oracle.spgemm_row_ub(A, B)
    for rep in range(2):
        C = run(ctx, A, B)
        check(C, A, B)
        assert ctx.stats()["products"] == int(ub.sum()), rep


def test_short_path_declines_rows_of_many_a_entries(ctx):
    """Rows with more than 64 A entries but at most 64 products (most B rows
    empty): the short-row path declines them and the general pipeline
    computes the product; then a short product takes the short-row path
    again."""
    print("case: short-row path declines long A rows", flush=True)
    lensB = np.where(np.arange(900) % 10 == 0, 3, 0)
    B = random_csr(900, 700, lensB, 32, 1)
    A = random_csr(500, 900, np.where(np.arange(500) % 7 == 0, 100, 5), 32, 0)
    assert oracle.spgemm_row_ub(A, B).max() <= 64 and A.row_lengths().max() > 64
    check(run(ctx, A, B), A, B)
    A2, B2 = uniform_rows(4000, 700, 0, 8, 33, 0), uniform_rows(700, 900, 0, 8, 33, 1)
    for _ in range(2):
        check(run(ctx, A2, B2), A2, B2)
    check(run(ctx, A, B), A, B)


def test_matrix_market_file(ctx, tmp_path):
    """A symmetric Matrix Market file (lower triangle stored) read with its
    values, squared on the device."""
    print("case: row-wise on a Matrix Market file", flush=True)
    import scipy.io
    import scipy.sparse as sp
    from sparsews_amd import io as spio
    M = sp.tril(sp.random(800, 800, density=0.01, random_state=31)).tocoo()
    path = tmp_path / "sym.mtx"
    scipy.io.mmwrite(str(path), M, symmetry="symmetric")
    A = spio.read_mtx(path)
    C = spgemm_rowwise(ctx, DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, A)).download()
    check(C, A, A)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_uniform_small(ctx, seed):
    print(f"case: uniform seed {seed}", flush=True)
    A, B = uniform_rows(500, 400, 0, 24, seed, 0), uniform_rows(400, 700, 0, 24, seed, 1)
    check(run(ctx, A, B), A, B)


@pytest.mark.parametrize("seed", [4, 5])
def test_powerlaw_mixed_paths(ctx, seed):
    print(f"case: power-law seed {seed}", flush=True)
    A, B = powerlaw_rows(1500, 1500, 1.8, 1500, seed, 0), powerlaw_rows(1500, 1500, 1.8, 1500, seed, 1)
    C = run(ctx, A, B)
    check(C, A, B)
    st = ctx.stats()
    used = [i for i in range(6) if st["rows_per_path"][i] > 0]
    assert len(used) >= 3, st


def test_wide_rows_natural(ctx):
    """Rows with ub > 4096 and ub*16 < ncols take the HBM hash path."""
    print("case: wide rows", flush=True)
    n = 400_000
    A = random_csr(40, 3000, np.full(40, 400), 6, 0)
    B = random_csr(3000, n, np.full(3000, 15), 6, 1)
    C = run(ctx, A, B)
    check(C, A, B)
    assert ctx.stats()["rows_per_path"][5] == 40


@pytest.mark.parametrize("ncols", [3000, 140_000])
def test_dense_rows_natural(ctx, ncols):
    """ub * 16 >= ncols: the column bitmap with HBM accumulation up to 2^17
    columns, the dense HBM accumulator above."""
    print(f"case: dense rows ncols={ncols}", flush=True)
    A = random_csr(30, 500, np.full(30, 300), 7, 0)
    B = random_csr(500, ncols, np.full(500, 40), 7, 1)
    C = run(ctx, A, B)
    check(C, A, B)
    assert ctx.stats()["rows_per_path"][4] == 30


def test_long_rows_bitmap_hbm_accumulation(ctx):
    """Rows of ub > 4096 with at most 2^17 columns, some of them with few
    distinct columns (many products per entry)."""
    print("case: long rows, bitmap + HBM accumulation", flush=True)
    A = random_csr(64, 3000, np.full(64, 400), 9, 0)
    B = random_csr(3000, 100_000, np.full(3000, 15), 9, 1)
    check(run(ctx, A, B), A, B)
    Bn = random_csr(3000, 700, np.full(3000, 15), 9, 2)
    check(run(ctx, A, Bn), A, Bn)


@pytest.mark.parametrize("ncols", [3000, 131072, 131073, 200_000])
@pytest.mark.parametrize("per_row", [(20, 30), (40, 60)])
def test_block_bins_bitmap_and_hash(ctx, ncols, per_row):
    """Rows in the block bins (256 < ub <= 4096): the LDS bitmap kernel up to
    2^17 output columns, the LDS hash table and bitonic sort above."""
    a_len, b_len = per_row
    print(f"case: block bins ncols={ncols} a={a_len} b={b_len}", flush=True)
    A = random_csr(300, 2000, np.full(300, a_len), 8 + a_len, 0)
    B = random_csr(2000, ncols, np.full(2000, b_len), 8 + a_len, 1)
    C = run(ctx, A, B)
    check(C, A, B)
    st = ctx.stats()
    assert st["rows_per_path"][2] + st["rows_per_path"][3] == 300
    dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
    np.testing.assert_array_equal(spgemm_rowwise_symbolic(ctx, dA, dB), C.pos)


@pytest.mark.parametrize("a_len,b_len,ncols,path", [
    (30, 50, 5000, 3),       # second block bin, 1500 products held in registers
    (700, 1, 5000, 2),       # block bin, 11 chunks of A entries into registers
    (1800, 1, 5000, 3),      # second block bin, 29 chunks into registers
    (40, 60, 5000, 3),       # 2400 products: beyond the register capacity, two gathers
    (600, 1, 200_000, 2),    # LDS hash kernel, wavefronts on chunks of their own
    (3000, 3, 100_000, 4),   # long rows: staged products, LDS accumulation, 16 wavefronts on own chunks
    (2000, 15, 100_000, 4),  # long rows of ~26k entries: more than the LDS accumulator holds, HBM atomics
    (3000, 3, 400_000, 5),   # long rows: HBM hash table, wavefronts on own chunks
])
def test_rows_of_many_a_entries(ctx, a_len, b_len, ncols, path):
    """Rows whose products span many 64-entry chunks of A entries, in every
    kernel that enumerates them."""
    print(f"case: many A entries a={a_len} b={b_len} ncols={ncols}", flush=True)
    m, k = 40, 4000
    A = random_csr(m, k, np.full(m, a_len), 30 + a_len, 0)
    B = random_csr(k, ncols, np.full(k, b_len), 30 + a_len, 1)
    C = run(ctx, A, B)
    check(C, A, B)
    assert ctx.stats()["rows_per_path"][path] == m
    dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
    np.testing.assert_array_equal(spgemm_rowwise_symbolic(ctx, dA, dB), C.pos)


@pytest.mark.parametrize("every,path", [(10, 0), (2, 1)])
def test_wave_bins_with_many_a_entries(ctx, every, path):
    """Rows of 300 A entries whose B rows are mostly empty: at most 64 (or
    256) products over five chunks of A entries, gathered chunk by chunk by
    the wavefront-per-row kernels."""
    print(f"case: wave bin {path}, 300 A entries, B rows nonempty every {every}", flush=True)
    m, k = 50, 3000
    A = random_csr(m, k, np.full(m, 300), 60 + every, 0)
    lengths = np.zeros(k, dtype=np.int64)
    lengths[::every] = 1
    B = random_csr(k, 700, lengths, 60 + every, 1)
    assert oracle.spgemm_row_ub(A, B).max() <= (64 if path == 0 else 256)
    C = run(ctx, A, B)
    check(C, A, B)
    assert ctx.stats()["rows_per_path"][path] == m


@pytest.mark.parametrize("path", [0, 1, 2, 3, 4, 5])
def test_forced_paths(ctx, path):
    print(f"case: forced path {path}", flush=True)
    A, B = uniform_rows(300, 300, 0, 8, 20 + path, 0), uniform_rows(300, 500, 0, 8, 20 + path, 1)
    saved = ctx.rowwise_tuning()
    try:
        ctx.set_rowwise_tuning(force_path=path)
        C = run(ctx, A, B)
    finally:
        ctx.set_rowwise_tuning(**saved)
    check(C, A, B)
    st = ctx.stats()
    assert st["rows_per_path"][path] == int((oracle.spgemm_row_ub(A, B) > 0).sum())


def test_forced_path_too_small_is_an_error(ctx):
    from sparsews_amd import SpwsError
    A, B = uniform_rows(50, 300, 20, 20, 30, 0), uniform_rows(300, 300, 20, 20, 30, 1)
    saved = ctx.rowwise_tuning()
    try:
        ctx.set_rowwise_tuning(force_path=0)
        with pytest.raises(SpwsError):
            run(ctx, A, B)
    finally:
        ctx.set_rowwise_tuning(**saved)


def test_edge_cases(ctx):
    print("case: edge cases", flush=True)
    # empty A
    A = CSR(7, 5, np.zeros(8, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    B = uniform_rows(5, 9, 1, 4, 31, 1)
    C = run(ctx, A, B)
    assert C.nnz == 0 and np.all(C.pos == 0)
    # zero rows
    A0 = CSR(0, 5, np.zeros(1, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    C0 = run(ctx, A0, B)
    assert C0.nnz == 0 and len(C0.pos) == 1
    # identity
    n = 200
    I = CSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    Bm = uniform_rows(n, 150, 0, 30, 32, 1)
    CI = run(ctx, I, Bm)
    np.testing.assert_array_equal(CI.pos, Bm.pos)
    np.testing.assert_array_equal(CI.crd, Bm.crd)
    np.testing.assert_array_equal(CI.vals, Bm.vals)
    # one fully dense row in A
    A1 = random_csr(3, 64, np.array([64, 0, 5]), 33, 0)
    B1 = uniform_rows(64, 80, 0, 6, 33, 1)
    check(run(ctx, A1, B1), A1, B1)


def test_buffer_reuse_large_then_small(ctx):
    print("case: reuse", flush=True)
    A, B = uniform_rows(800, 800, 0, 30, 40, 0), uniform_rows(800, 800, 0, 30, 40, 1)
    check(run(ctx, A, B), A, B)
    a, b = uniform_rows(90, 60, 0, 9, 41, 0), uniform_rows(60, 70, 0, 9, 41, 1)
    check(run(ctx, a, b), a, b)


def test_symbolic_phase(ctx):
    print("case: symbolic", flush=True)
    A, B = powerlaw_rows(700, 700, 1.8, 700, 42, 0), powerlaw_rows(700, 700, 1.8, 700, 42, 1)
    pos = spgemm_rowwise_symbolic(ctx, DeviceCSR.upload(ctx, A, False), DeviceCSR.upload(ctx, B, False))
    np.testing.assert_array_equal(pos, oracle.spgemm_symbolic(A, B))


def test_taco_entry_points():
    print("case: taco compute / assemble+compute", flush=True)
    A, B = uniform_rows(300, 250, 0, 12, 43, 0), uniform_rows(250, 260, 0, 12, 43, 1)
    At, Bt = taco.DC_to_taco_tensor(A), taco.DC_to_taco_tensor(B)
    Ct = taco.init_taco_tensor_DC(A.nrows, B.ncols)
    taco.compute(Ct, At, Bt, 1 << 10)
    C = taco.taco_to_csr(Ct)
    taco.free_outputs(Ct)
    check(C, A, B)
    C2t = taco.init_taco_tensor_DC(A.nrows, B.ncols)
    taco.assemble(C2t, At, Bt)
    np.testing.assert_array_equal(C2t.index(1, 0, A.nrows + 1), oracle.spgemm_symbolic(A, B))
    taco.compute_split(C2t, At, Bt)
    C2 = taco.taco_to_csr(C2t)
    taco.free_outputs(C2t)
    check(C2, A, B)


def test_kernel_timing_filter(ctx):
    """Every launch is timed by default; spws_ctx_set_profiling_only limits
    the events to one kernel name.  Products alternate between the short-row
    path and the binned path, which also checks that each leaves the bin
    metadata ready for the next."""
    print("case: kernel timing filter", flush=True)
    short = (uniform_rows(400, 300, 0, 8, 44, 0), uniform_rows(300, 350, 0, 8, 44, 1))
    binned = (uniform_rows(300, 400, 10, 30, 45, 0), uniform_rows(400, 600, 10, 30, 45, 1))
    try:
        ctx.set_profiling(True)
        ctx.reset_kernel_times()
        check(run(ctx, *short), *short)
        every = ctx.kernel_times()
        assert "short_rows" in every and "compact_fixed" not in every, every
        ctx.set_profiling_only("short_rows")
        ctx.reset_kernel_times()
        for A, B in (short, binned, short):
            check(run(ctx, A, B), A, B)
        only = ctx.kernel_times()
        # each short product takes the short-row kernels; the binned one,
        # following a short one, tries it first (it declines, then the bins run)
        assert set(only) == {"short_rows"}, only
        assert only["short_rows"][1] == 3 and only["short_rows"][0] > 0.0
        # a binned product after a binned one does not try it
        check(run(ctx, *binned), *binned)
        ctx.reset_kernel_times()
        check(run(ctx, *binned), *binned)
        assert ctx.kernel_times() == {}, ctx.kernel_times()
    finally:
        ctx.set_profiling_only(None)
        ctx.set_profiling(False)
