"""Batched SpGEMM (one hipGraph launch for many requests) and the serving
loop, on the GPU, through the C ABI.

Every request's result must match the oracle (structure bit-exact, values
within 1e-12), match the plain one-call path, and be the same alone and in a
mix.  The graph is replayed several times with eager work between replays,
and every replay is checked.
"""
import numpy as np
import pytest

import oracle
from helpers import powerlaw_rows, random_csr, uniform_rows
from sparsews_amd import SpwsError, gen
from sparsews_amd.device import Context, DeviceCSR, spgemm_rowwise
from sparsews_amd.serve import DeviceRequest, Engine, default_caps, latency_summary, serve

pytestmark = pytest.mark.gpu
TOL = 1e-12

SMALL_CAPS = dict(max_requests=64, max_rows=1 << 16, max_nnz=1 << 22, max_products=1 << 25,
                  max_row_products=1 << 20)


def make_requests(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(rng.integers(40, 1500))
        if t % 2 == 0:
            A, B = uniform_rows(n, n, 0, 30, seed * 100 + t, 0), uniform_rows(n, n, 0, 30, seed * 100 + t, 1)
        else:
            A = powerlaw_rows(n, n, 1.8, min(n, 400), seed * 100 + t, 0)
            B = powerlaw_rows(n, n, 1.8, min(n, 400), seed * 100 + t, 1)
        out.append((A, B))
    return out


def check_result(C, A, B):
    G = oracle.spgemm_gustavson(A, B)
    np.testing.assert_array_equal(C.pos, G.pos)
    np.testing.assert_array_equal(C.crd, G.crd)
    exact, absum = oracle.spgemm_refsum(A, B, G.pos, G.crd)
    assert oracle.max_value_error(C.vals, exact, absum) <= TOL
    return G, absum


@pytest.fixture(scope="module")
def engine(ctx):
    e = Engine(ctx, default_caps(**SMALL_CAPS))
    yield e
    e.close()


def run_batch(engine, ctx, pairs):
    engine.reset()
    reqs = [DeviceRequest.upload(ctx, A, B) for A, B in pairs]
    slots = [engine.add(r) for r in reqs]
    engine.launch()
    engine.wait()
    return [engine.result(s, A.nrows, B.ncols) for s, (A, B) in zip(slots, pairs)], reqs


def test_batch_matches_oracle_and_plain(ctx, engine):
    print("case: batch of 24 mixed requests", flush=True)
    pairs = make_requests(1, 24)
    results, reqs = run_batch(engine, ctx, pairs)
    total = 0
    for (A, B), C, r in zip(pairs, results, reqs):
        G, absum = check_result(C, A, B)
        P = spgemm_rowwise(ctx, r.A, r.B).download()
        np.testing.assert_array_equal(P.crd, C.crd)
        assert oracle.max_value_error(C.vals, P.vals, absum) <= TOL
        total += C.nnz
    assert engine.nnz() == total


def test_graph_replays_with_eager_work_between(ctx, engine):
    for rep, seed in enumerate([2, 3, 4]):
        print(f"case: graph replay {rep}", flush=True)
        pairs = make_requests(seed, 10 + 4 * rep)
        results, _ = run_batch(engine, ctx, pairs)
        # eager work on the context's stream between replays
        a, b = uniform_rows(300, 300, 1, 9, 90 + rep, 0), uniform_rows(300, 300, 1, 9, 90 + rep, 1)
        check_result(spgemm_rowwise(ctx, DeviceCSR.upload(ctx, a), DeviceCSR.upload(ctx, b)).download(), a, b)
        for (A, B), C in zip(pairs, results):
            check_result(C, A, B)


def test_same_alone_and_in_the_mix(ctx, engine):
    print("case: alone vs mix", flush=True)
    pairs = make_requests(5, 16)
    A0, B0 = pairs[7]
    alone, _ = run_batch(engine, ctx, [(A0, B0)])
    mix, _ = run_batch(engine, ctx, pairs)
    C1, C2 = alone[0], mix[7]
    np.testing.assert_array_equal(C1.pos, C2.pos)
    np.testing.assert_array_equal(C1.crd, C2.crd)
    _, absum = check_result(C1, A0, B0)
    assert oracle.max_value_error(C1.vals, C2.vals, absum) <= TOL


def test_long_rows_take_the_hbm_hash_path_in_a_batch(ctx, engine):
    print("case: long rows in a batch", flush=True)
    A = random_csr(20, 600, np.full(20, 300), 6, 0)
    B = random_csr(600, 20000, np.full(600, 20), 6, 1)
    small = make_requests(6, 4)
    results, _ = run_batch(engine, ctx, [(A, B)] + small)
    check_result(results[0], A, B)
    for (a, b), C in zip(small, results[1:]):
        check_result(C, a, b)


def test_request_widths_in_one_batch(ctx, engine):
    """Rows of requests from 64 to 65,536 columns wide share the bitmap
    kernels' workgroups; each row's passes cover its own request's columns
    (block bins in registers and with two gathers, and rows beyond them)."""
    print("case: request widths 64..65536 in one batch", flush=True)
    pairs = []
    for t, (n, a_len, b_len) in enumerate([(64, 20, 30), (65536, 30, 40), (700, 25, 35), (30000, 40, 60),
                                           (5000, 400, 12), (96, 10, 50)]):
        k = max(2 * a_len, 800)
        A = random_csr(60, k, np.full(60, a_len), 70 + t, 0)
        B = random_csr(k, n, np.full(k, min(b_len, n)), 70 + t, 1)
        pairs.append((A, B))
    results, _ = run_batch(engine, ctx, pairs)
    for (A, B), C in zip(pairs, results):
        check_result(C, A, B)


def test_any_width_engine_and_width_limit(ctx):
    """max_cols = 0: the block bins hash and sort; max_cols > 0: the LDS
    column bitmap, and a wider request is refused (it runs plain)."""
    print("case: engine width settings", flush=True)
    pairs = make_requests(9, 10)
    wide = dict(SMALL_CAPS, max_cols=0)
    e = Engine(ctx, default_caps(**wide))
    try:
        results, _ = run_batch(e, ctx, pairs)
        for (A, B), C in zip(pairs, results):
            check_result(C, A, B)
    finally:
        e.close()
    narrow = dict(SMALL_CAPS, max_cols=1024)
    e2 = Engine(ctx, default_caps(**narrow))
    try:
        A, B = uniform_rows(100, 100, 1, 9, 95, 0), uniform_rows(100, 2000, 1, 9, 95, 1)
        with pytest.raises(SpwsError) as ex:
            e2.add(DeviceRequest.upload(ctx, A, B))
        assert ex.value.code == 4
    finally:
        e2.close()


def test_admission_limits(ctx):
    print("case: admission limits", flush=True)
    caps = default_caps(max_requests=4, max_rows=1 << 12, max_nnz=1 << 16, max_products=1 << 15,
                        max_row_products=256)
    e = Engine(ctx, caps)
    try:
        tiny = [(uniform_rows(50, 50, 0, 4, 700 + t, 0), uniform_rows(50, 50, 0, 4, 700 + t, 1)) for t in range(6)]
        reqs = [DeviceRequest.upload(ctx, A, B) for A, B in tiny]
        assert all(r.max_row_products <= 16 for r in reqs)
        for r in reqs[:4]:
            e.add(r)
        with pytest.raises(SpwsError) as ex:
            e.add(reqs[4])
        assert ex.value.code == 2  # the batch is full
        e.reset()
        A = random_csr(5, 300, np.full(5, 100), 8, 0)
        B = random_csr(300, 300, np.full(300, 10), 8, 1)
        with pytest.raises(SpwsError) as ex2:
            e.add(DeviceRequest.upload(ctx, A, B))
        assert ex2.value.code == 4  # a row beyond max_row_products
        # an emptied engine still runs a batch correctly
        e.reset()
        slots = [e.add(r) for r in reqs[:3]]
        e.launch()
        e.wait()
        for (a, b), sl in zip(tiny[:3], slots):
            check_result(e.result(sl, 50, 50), a, b)
    finally:
        e.close()


def test_serve_plain_and_batched(ctx):
    """Both loop modes hand back every request's product (copied into the
    request's own output buffer), equal to the oracle's and to the same
    request run alone; the batched loop runs two engines and a large lane."""
    print("case: serve loop", flush=True)
    pairs = make_requests(8, 48)
    reqs = [DeviceRequest.upload(ctx, A, B, output=True) for A, B in pairs]
    arrivals = np.cumsum(np.random.default_rng(8).exponential(2e-4, len(reqs)))
    arrivals -= arrivals[0]
    engines = [Engine(ctx, default_caps(**SMALL_CAPS)) for _ in range(3)]
    large_at = int(np.percentile([r.products for r in reqs], 80))
    try:
        f_plain, b_plain, st_plain, n_plain = serve(ctx, [], reqs, arrivals, "plain", with_nnz=True)
        plain_results = [r.result(int(k)) for r, k in zip(reqs, n_plain)]
        f_bat, b_bat, st_bat, n_bat = serve(ctx, engines, reqs, arrivals, "batched", large_products=large_at,
                                            large_engines=1, with_nnz=True)
        batch_results = [r.result(int(k)) for r, k in zip(reqs, n_bat)]
    finally:
        for e in engines:
            e.close()
    assert np.all(f_plain >= arrivals) and np.all(f_bat >= arrivals)
    assert np.all(b_plain == -1)
    assert st_bat["batches"] >= 1 and st_bat["batches"] <= len(reqs)
    assert np.all(b_bat >= 0)
    for (A, B), r, Cp, Cb in zip(pairs, reqs, plain_results, batch_results):
        _, absum = check_result(Cp, A, B)
        check_result(Cb, A, B)
        alone = spgemm_rowwise(ctx, r.A, r.B).download()
        np.testing.assert_array_equal(alone.crd, Cb.crd)
        assert oracle.max_value_error(Cb.vals, alone.vals, absum) <= TOL
    print("plain", latency_summary(arrivals, f_plain), "batched", latency_summary(arrivals, f_bat), st_bat)


def test_output_capacity_too_small(ctx):
    """A request whose output buffer is smaller than its product reports -1
    and the rest of its batch is delivered."""
    print("case: output capacity", flush=True)
    pairs = make_requests(11, 6)
    reqs = [DeviceRequest.upload(ctx, A, B, output=True) for A, B in pairs]
    nnz_true = [oracle.spgemm_gustavson(A, B).nnz for A, B in pairs]
    reqs[2].alloc_output(ctx, capacity=max(nnz_true[2] - 1, 0))
    engines = [Engine(ctx, default_caps(**SMALL_CAPS))]
    try:
        _, _, _, n_bat = serve(ctx, engines, reqs, np.zeros(len(reqs)), "batched", with_nnz=True)
        _, _, _, n_plain = serve(ctx, [], reqs, np.zeros(len(reqs)), "plain", with_nnz=True)
    finally:
        engines[0].close()
    for n_out in (n_bat, n_plain):
        assert n_out[2] == -1
        for t in (0, 1, 3, 4, 5):
            assert n_out[t] == nnz_true[t]
            check_result(reqs[t].result(int(n_out[t])), *pairs[t])


def test_underreported_bounds_are_refused_on_the_device(ctx):
    """Admission trusts the caller's products; the graph recomputes them and
    refuses a batch whose true bounds exceed the pools (nothing is written),
    and the engine computes the next batch correctly."""
    print("case: under-reported upper bound", flush=True)
    caps = default_caps(max_requests=4, max_rows=1 << 12, max_nnz=1 << 16, max_products=1 << 14,
                        max_row_products=1 << 10)
    e = Engine(ctx, caps)
    try:
        A, B = uniform_rows(400, 400, 8, 12, 801, 0), uniform_rows(400, 400, 8, 12, 801, 1)
        r = DeviceRequest.upload(ctx, A, B)
        assert r.products > caps.max_products
        r.products = 1000  # the caller's claim
        e.reset()
        e.add(r)
        e.launch()
        e.wait()
        with pytest.raises(SpwsError) as ex:
            e.nnz()
        assert ex.value.code == 1
        with pytest.raises(SpwsError):
            e.result_nnz(0)
        a, b = uniform_rows(60, 60, 0, 5, 802, 0), uniform_rows(60, 60, 0, 5, 802, 1)
        results, _ = run_batch(e, ctx, [(a, b)])
        check_result(results[0], a, b)
    finally:
        e.close()


def config4_selection(stream, prefix, largest):
    """Indices of the first `prefix` requests of the stream and of its
    `largest` largest requests of each kind (by rows; the uniform ones hold
    the stream's q99-and-above product counts)."""
    pick = set(range(prefix))
    for kind in ("uniform", "powerlaw"):
        ranked = sorted((r for r in stream if r.kind == kind), key=lambda r: -r.n)
        pick.update(r.rid for r in ranked[:largest])
    return sorted(pick)


def test_config4_stream_served_continuously(ctx):
    """The served-SpGEMM request mix itself (n log-uniform in 256..65,536,
    half uniform 2..64 entries per row, half Zipf(1.8) capped at 4,096; seed
    0): its first 120 requests and the largest ones (n up to 65,517, about
    7e7 products), arriving as a Poisson stream fast enough that batches
    overlap, through the continuously batched loop with four engines and a
    large lane.  Every served product is checked against the oracle
    (structure bit-exact, values within 1e-12) and against the same request
    computed alone."""
    stream = gen.request_stream(count=4096, seed=0)
    idx = config4_selection(stream, 120, 4)
    print(f"case: config-4 stream, {len(idx)} requests", flush=True)
    pairs = [gen.request_operands(stream[i]) for i in idx]
    reqs = [DeviceRequest.upload(ctx, A, B, output=True) for A, B in pairs]
    products = np.array([r.products for r in reqs])
    assert max(r.n for r in reqs) >= 65_000 and products.max() >= 6e7
    gaps = np.random.default_rng(44).exponential(1.0 / 20_000.0, len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    engines = [Engine(ctx) for _ in range(4)]
    try:
        fin, bid, st, nnz = serve(ctx, engines, reqs, arrivals, "batched", max_batch_products=1 << 26,
                                  large_products=1 << 24, large_engines=1, with_nnz=True)
    finally:
        for e in engines:
            e.close()
    print("served:", st, latency_summary(arrivals, fin), flush=True)
    assert st["fallback_requests"] == 0
    # batches carried several requests each
    assert st["batches"] < len(reqs)
    for t, ((A, B), r) in enumerate(zip(pairs, reqs)):
        C = r.result(int(nnz[t]))
        _, absum = check_result(C, A, B)
        alone = spgemm_rowwise(ctx, r.A, r.B).download()
        np.testing.assert_array_equal(alone.pos, C.pos)
        np.testing.assert_array_equal(alone.crd, C.crd)
        assert oracle.max_value_error(C.vals, alone.vals, absum) <= TOL
        for arr in (r.C_pos, r.C_crd, r.C_vals):
            arr.free()


def test_tuned_settings_serve_correctly():
    """The autotuner's table (sparsews_amd/tuning.json, the defaults when it
    is absent) applied as the serve bench applies it -- row bins on the
    context, engines, products per batch and the large lane -- on a prefix
    of the served request mix: every served product against the oracle."""
    from sparsews_amd import autotune
    print("case: tuned settings serve correctly", flush=True)
    with Context(0) as tctx:
        shape = autotune.apply(tctx, autotune.load())
        stream = gen.request_stream(count=4096, seed=0)
        idx = config4_selection(stream, 60, 2)
        pairs = [gen.request_operands(stream[i]) for i in idx]
        reqs = [DeviceRequest.upload(tctx, A, B, output=True) for A, B in pairs]
        gaps = np.random.default_rng(45).exponential(1.0 / 20_000.0, len(reqs))
        arrivals = np.cumsum(gaps) - gaps[0]
        engines = [Engine(tctx) for _ in range(shape["engines"])]
        try:
            fin, bid, st, nnz = serve(tctx, engines, reqs, arrivals, "batched",
                                      max_batch_products=shape["max_batch_products"],
                                      large_products=shape["large_products"],
                                      large_engines=min(shape["large_engines"], shape["engines"] - 1),
                                      with_nnz=True)
        finally:
            for e in engines:
                e.close()
        assert st["fallback_requests"] == 0
        for t, ((A, B), r) in enumerate(zip(pairs, reqs)):
            check_result(r.result(int(nnz[t])), A, B)
            for arr in (r.C_pos, r.C_crd, r.C_vals):
                arr.free()
