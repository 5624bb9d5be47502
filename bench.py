#!/usr/bin/env python3
"""Benchmark of the sparse-workspace kernels on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hot path over one synthetic batch with the
operands already resident in HBM.  Rank 0 prints ONE JSON line with the
metric, the roofline of the dominant kernel (HIP events around every launch
on the context's stream, recorded inside the timed region) and the CPU
baseline (the oracle on the host cores, on a bounded sample).

Workloads (sizes in the docstring of each runner):
  outer     outer-product CSC x CSR SpGEMM, 200,000 x 200,000, Zipf 1.8 rows
  rowwise   row-wise CSR x CSR hash SpGEMM, 10,000 x 10,000, 8 per row
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(ctx, step, steps, warmup):
    for _ in range(warmup):
        step()
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    times = ctx.kernel_times()
    ctx.set_profiling(False)
    return (t1 - t0) / steps, times


def roofline(times, name, bytes_per_launch):
    ms, launches = times[name]
    avg_s = ms / launches / 1e3
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": name, "avg_launch_us": round(avg_s * 1e6, 3),
            "algorithmic_bytes_per_launch": int(bytes_per_launch)}


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# --------------------------------------------------------------------------- row-wise

def run_rowwise(args):
    """Row-wise hash SpGEMM, A and B independent 10,000 x 10,000 CSR with 8
    distinct uniform columns per row, FP64 U[-1,1], seed 0."""
    from sparsews_amd import gen
    from sparsews_amd.device import Context, DeviceCSR, spgemm_rowwise
    A, B = gen.rowwise_pair()
    ctx = Context(0)
    dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
    out = {}

    def step():
        out["C"] = spgemm_rowwise(ctx, dA, dB)

    sec, times = timed_steps(ctx, step, args.steps, args.warmup)
    C = out["C"]
    P = ctx.stats()["products"]
    nnz = C.nnz
    m = A.nrows
    # dominant kernel: the numeric kernel of the short-row bin.  Algorithmic
    # bytes per launch: A rows (pos 8 B/row + 12 B/entry), B_pos gathers (8 B per
    # A entry), B entries gathered (12 B per product), slot writes (12 B per
    # output entry), per-row metadata (row id 4 + offset 8 + nnz 4 B).
    dom = max((k for k in times if k.startswith("row_")), key=lambda k: times[k][0])
    dom_bytes = m * 8 + A.nnz * 12 + A.nnz * 8 + P * 12 + nnz * 12 + m * 16
    compulsory = (4 * (m + 1) + 12 * A.nnz) + (4 * (B.nrows + 1) + 12 * B.nnz) + (4 * (m + 1) + 12 * nnz)
    res = {
        "metric": "SpGEMM GFLOP/s",
        "value": round(2 * P / sec / 1e9, 3),
        "unit": "GFLOP/s",
        "ms_per_step": round(sec * 1e3, 4),
        "dtype": "f64",
        "config": {"workload": "rowwise_hash_spgemm_10k_r8", "m": m, "n": B.ncols, "nnz_per_row": 8,
                   "products": int(P), "nnz_C": int(nnz), "seed": 0},
        "roofline": roofline(times, dom, dom_bytes),
        "compulsory_bytes": int(compulsory),
        "compulsory_frac_of_hbm_roofline": round(compulsory / sec / 1e9 / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in sorted(times.items())},
    }
    if not args.no_cpu_baseline:
        import oracle
        cores = cpu_cores()
        oracle.spgemm_gustavson(A, B, cores)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            oracle.spgemm_gustavson(A, B, cores)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        res["cpu_baseline"] = {"value": round(2 * P / dt / 1e9, 3), "unit": "GFLOP/s", "cores": cores,
                               "kind": "port",
                               "sample": f"full workload, OpenMP Gustavson restatement "
                                         f"(CSR_CSR_taco_parallel.h), {reps} runs"}
        t0 = time.perf_counter()
        oracle.spgemm_hash(A, B, 1 << 16)
        dt1 = time.perf_counter() - t0
        res["cpu_baseline_hash_sequential"] = {"value": round(2 * P / dt1 / 1e9, 3), "unit": "GFLOP/s",
                                               "cores": 1, "kind": "port",
                                               "sample": "full workload, w_cap 2^16, 1 run"}
    ctx.close()
    return res


WORKLOADS = {"rowwise": run_rowwise}
DEFAULT_WORKLOAD = "rowwise"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0
    res = WORKLOADS[args.workload](args)
    line = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic", "config": res.pop("config")}
    line.update(res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
