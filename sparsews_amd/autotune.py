"""Autotuner of the batched / served row-wise SpGEMM.

    python -m sparsews_amd.autotune [--requests 1024] [--out PATH]

Runs on the GPU over a seeded sample of the served request mix (square
requests, n log-uniform in 256..65,536, half uniform 2..64 and half Zipf(1.8)
entries per row) and writes a table that the serving code loads
(``load``/``apply``; the default path is sparsews_amd/tuning.json):

1. Row bins.  The bin thresholds decide which kernel (and so which LDS table
   size and workgroup shape) each row takes: wave_ub rows run one wavefront
   with a 128-slot LDS table, wave2_ub rows one wavefront with 512 slots,
   block_ub / block2_ub rows one 256-thread workgroup with the LDS column
   bitmap and up to 4 / 8 products per thread in registers, longer rows the
   1,024-thread staged bitmap kernel.  Coordinate descent over the
   thresholds, scored by the batched saturation throughput (every request
   queued at t = 0) of 4 engines.
2. Serving shape.  Engines, products per batch, and the large lane
   (threshold and reserved engines), scored by the highest Poisson arrival
   rate whose p99 latency stays within 2 ms (bisection over the rate, the
   same unit-rate arrival pattern scaled).

The arithmetic does not depend on any of these settings: every choice
computes each entry in the same order, so the tuning changes only speed.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import gen
from .device import Context
from .serve import DeviceRequest, Engine, serve

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")
P99_TARGET_MS = 2.0

# the engine's defaults (spws_ctx's row-wise tuning and the bench's serve options)
BASE_BINS = {"wave_ub": 64, "wave2_ub": 256, "block_ub": 1024, "block2_ub": 4096}
BASE_SERVE = {"engines": 4, "max_batch_products": 1 << 26, "large_products": 1 << 24, "large_engines": 1}
BIN_CHOICES = {"wave_ub": [32, 48, 64], "wave2_ub": [128, 192, 256], "block_ub": [512, 768, 1024],
               "block2_ub": [2048, 3072, 4096]}
SERVE_CHOICES = {"engines": [3, 4, 6, 8], "max_batch_products": [1 << 24, 1 << 25, 1 << 26, 1 << 27],
                 "large_products": [0, 1 << 22, 1 << 23, 1 << 24, 1 << 25], "large_engines": [1, 2]}


def load(path: str = DEFAULT_PATH) -> dict | None:
    """The tuning table, or None when there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def apply(ctx: Context, table: dict | None) -> dict:
    """Sets the table's row bins on ctx (engines created afterwards copy
    them) and returns the serve options (engines, max_batch_products,
    large_products, large_engines); the defaults when table is None."""
    if table is None:
        return dict(BASE_SERVE)
    ctx.set_rowwise_tuning(**table["bins"])
    return dict(table["serve"])


def request_sample(count: int, seed: int = 0, stride: int = 4):
    """Every stride-th request of the first count * stride of the stream
    (the stream's size mix, fewer requests)."""
    stream = gen.request_stream(count=count * stride, seed=seed)
    return stream[::stride][:count]


def _engines(ctx, k):
    return [Engine(ctx) for _ in range(k)]


def _close(engines):
    for e in engines:
        e.close()


def saturation(ctx, reqs, bins: dict, serve_opts: dict) -> float:
    """Batched requests per second with every request queued at t = 0."""
    ctx.set_rowwise_tuning(**bins)
    engines = _engines(ctx, serve_opts["engines"])
    try:
        kw = dict(max_batch_products=serve_opts["max_batch_products"],
                  large_products=serve_opts["large_products"],
                  large_engines=min(serve_opts["large_engines"], serve_opts["engines"] - 1))
        zero = np.zeros(len(reqs))
        serve(ctx, engines, reqs[:32], zero[:32], "batched", **kw)  # warm-up
        best = 0.0
        for _ in range(2):
            _, _, st = serve(ctx, engines, reqs, zero, "batched", **kw)
            best = max(best, len(reqs) / st["wall_seconds"])
        return best
    finally:
        _close(engines)


def latency_bound(ctx, reqs, serve_opts: dict, sat: float, steps: int = 6):
    """Highest Poisson rate with p99 within P99_TARGET_MS (and that p99)."""
    engines = _engines(ctx, serve_opts["engines"])
    kw = dict(max_batch_products=serve_opts["max_batch_products"], large_products=serve_opts["large_products"],
              large_engines=min(serve_opts["large_engines"], serve_opts["engines"] - 1))
    gaps = np.random.default_rng(777).exponential(1.0, len(reqs))
    unit = np.cumsum(gaps) - gaps[0]

    def p99_at(rate):
        arr = unit / rate
        fin, _, _ = serve(ctx, engines, reqs, arr, "batched", **kw)
        return float(np.percentile((fin - arr) * 1e3, 99))

    try:
        lo, hi = 0.2 * sat, sat
        p_lo = p99_at(lo)
        if p_lo > P99_TARGET_MS:
            return 0.0, p_lo
        p_hi = p99_at(hi)
        if p_hi <= P99_TARGET_MS:
            return hi, p_hi
        for _ in range(steps):
            mid = (lo * hi) ** 0.5
            p = p99_at(mid)
            if p <= P99_TARGET_MS:
                lo, p_lo = mid, p
            else:
                hi = mid
        return lo, p_lo
    finally:
        _close(engines)


def tune(count: int = 1024, log=print) -> dict:
    ctx = Context(0)
    t0 = time.perf_counter()
    reqs = []
    for r in request_sample(count):
        A, B = gen.request_operands(r)
        reqs.append(DeviceRequest.upload(ctx, A, B))
    log(f"autotune: {len(reqs)} requests uploaded ({time.perf_counter() - t0:.0f} s)")
    bins = dict(BASE_BINS)
    trials = []
    base_rate = saturation(ctx, reqs, bins, BASE_SERVE)
    trials.append({"bins": dict(bins), "saturation_req_per_s": round(base_rate, 1)})
    log(f"autotune: default bins {bins}: {base_rate:.0f} req/s")
    # coordinate descent over the bin thresholds, largest bins first; each
    # threshold stays between its neighbours
    order = ["block2_ub", "block_ub", "wave2_ub", "wave_ub"]
    best_rate = base_rate
    for key in order:
        for v in BIN_CHOICES[key]:
            if v == bins[key]:
                continue
            cand = dict(bins, **{key: v})
            if not (cand["wave_ub"] <= cand["wave2_ub"] <= cand["block_ub"] <= cand["block2_ub"]):
                continue
            rate = saturation(ctx, reqs, cand, BASE_SERVE)
            trials.append({"bins": dict(cand), "saturation_req_per_s": round(rate, 1)})
            log(f"autotune: {key}={v}: {rate:.0f} req/s")
            if rate > best_rate * 1.01:
                best_rate, bins = rate, cand
    ctx.set_rowwise_tuning(**bins)
    # serving shape, scored by the latency-bound rate
    shape = dict(BASE_SERVE)
    sat = saturation(ctx, reqs, bins, shape)
    best_bound, best_p99 = latency_bound(ctx, reqs, shape, sat)
    trials.append({"serve": dict(shape), "latency_bound_req_per_s": round(best_bound, 1)})
    log(f"autotune: default serve shape {shape}: {best_bound:.0f} req/s at p99 {best_p99:.2f} ms")
    for key in ("large_products", "engines", "max_batch_products", "large_engines"):
        for v in SERVE_CHOICES[key]:
            if v == shape[key]:
                continue
            cand = dict(shape, **{key: v})
            if cand["large_engines"] >= cand["engines"]:
                continue
            sat_c = saturation(ctx, reqs, bins, cand)
            rate, p99 = latency_bound(ctx, reqs, cand, sat_c)
            trials.append({"serve": dict(cand), "latency_bound_req_per_s": round(rate, 1),
                           "saturation_req_per_s": round(sat_c, 1)})
            log(f"autotune: {key}={v}: {rate:.0f} req/s at p99 {p99:.2f} ms (saturation {sat_c:.0f})")
            if rate > best_bound * 1.02:
                best_bound, best_p99, shape = rate, p99, cand
    ctx.close()
    return {"bins": bins, "serve": shape,
            "measured": {"requests": len(reqs), "saturation_req_per_s": round(best_rate, 1),
                         "latency_bound_req_per_s": round(best_bound, 1), "p99_ms": round(best_p99, 3),
                         "p99_target_ms": P99_TARGET_MS},
            "trials": trials}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--requests", type=int, default=1024)
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args(argv)
    table = tune(a.requests, log=lambda *x: print(*x, flush=True))
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({k: table[k] for k in ("bins", "serve", "measured")}), flush=True)


if __name__ == "__main__":
    main()
