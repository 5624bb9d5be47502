#!/usr/bin/env python3
"""Benchmark of the sparse-workspace kernels on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hot path over one synthetic batch with the
operands already resident in HBM.  Rank 0 prints ONE JSON line with the
metric, the roofline of the dominant kernel (HIP events around every launch
on the context's stream, recorded inside the timed region) and the CPU
baseline (the oracle on the host cores, on a bounded sample).

Workloads (sizes in the docstring of each runner):
  all       (default) the outer product as the headline, with a record of
            each of rowwise, mttkrp, ttm and serve under "workloads"
  outer     outer-product DCSC x DCSR SpGEMM, 200,000 x 200,000,
            Zipf 1.8 row lengths, A = B^T
  rowwise   row-wise CSR x CSR hash SpGEMM, 10,000 x 10,000, 8 per row
  mttkrp    mode-1 MTTKRP, 50k x 40k x 30k, 20M nonzeros, rank 32
  ttm       mode-3 TTM, 100k x 80k x 2k, 10M nonzeros, U 2,000 x 16
  transpose CSR transpose on the device, 200k x 200k, 32 columns per row
  fusion    fused vs non-fused transposed-output product over 15 cases,
            output nnz 10^3 .. 10^8 (speedup per output-nnz decade)
  serve     4,096 SpGEMM requests, plain vs continuously batched
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


def timed_steps(ctx, step, steps, warmup, pick):
    """`warmup` untimed steps; one more untimed step with every launch timed
    (the per-kernel breakdown, from which `pick` names the dominant kernel);
    then exactly `steps` timed steps in which only the dominant kernel's
    launches carry HIP events (events on every launch would add their own
    cost to a short step).  Returns (seconds per timed step, dominant kernel,
    its (total ms, launches) over the timed steps, breakdown of one step)."""
    for _ in range(warmup):
        step()
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    step()
    ctx.sync()
    breakdown = ctx.kernel_times()
    dom = pick(breakdown)
    ctx.set_profiling_only(dom)
    ctx.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    dom_times = ctx.kernel_times()[dom]
    ctx.set_profiling(False)
    ctx.set_profiling_only(None)
    return (t1 - t0) / steps, dom, dom_times, breakdown


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` (the name the context's timer uses)
    for `workload`, from a committed profiles/*_traffic.md summary (two
    rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, calibrated by a plain
    copy kernel); None when no such summary exists."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(HERE, "profiles", "*_traffic.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"bench-traffic: workload=(\S*) kernel=(\S+) bytes_per_launch=(\d+)", line.strip())
                if m and m.group(1) == workload and re.sub(r"<.*", "", m.group(2)) in ("k_" + kernel, kernel):
                    return int(m.group(3))
    return None


def calibration_copy(ctx, nbytes):
    """One plain device copy of nbytes (reads and writes exactly nbytes):
    the reference point for the FETCH_SIZE / WRITE_SIZE counters."""
    if nbytes <= 0:
        return
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.copy(dst, src, nbytes)
    ctx.sync()
    ctx.free(src)
    ctx.free(dst)


def roofline(dom, dom_times, steps, compulsory_per_step, kernel_io_per_step, workload):
    """Roofline object of the dominant kernel `dom`.  `achieved` is the
    workload's compulsory bytes (inputs read once, outputs written once) per
    launch of `dom` divided by its average launch time from HIP events over
    the timed steps.  `kernel_io` is the same kernel measured against the
    bytes it reads and writes itself (workspace round trips included), and
    `traffic` the HBM bytes per launch that rocprofv3 counted (committed
    traffic summary), or None."""
    ms, launches = dom_times
    avg_s = ms / launches / 1e3
    alg = compulsory_per_step * steps / launches
    achieved = alg / avg_s / 1e9
    io = kernel_io_per_step * steps / launches
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(workload, dom), "kernel": dom,
            "avg_launch_us": round(avg_s * 1e6, 3), "algorithmic_bytes_per_launch": int(alg),
            "kernel_io": {"bytes_per_launch": int(io), "achieved": round(io / avg_s / 1e9, 2),
                          "frac": round(io / avg_s / 1e9 / HBM_PEAK_GBS, 4)}}


def rocsparse_leg(ctx, dA, dB, C, products, steps, warmup):
    """rocSPARSE's SpGEMM (installed with ROCm) on the same device operands,
    for context: wall time per product over `steps` products (its nnz and
    compute stages, buffers reused), and whether its output has our
    structure and values (within 1e-12 relative)."""
    from sparsews_amd import rocsparse_cmp
    if not rocsparse_cmp.available():
        return {"error": "librocsparse.so not found"}
    try:
        r = rocsparse_cmp.RocsparseSpgemm(ctx, dA, dB)
        for _ in range(warmup):
            r.run()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.run()
        ctx.sync()
        sec = (time.perf_counter() - t0) / steps
        pos, crd, vals = r.result()
        r.close()
    except RuntimeError as e:
        return {"error": str(e)}
    ours = C.download()
    same = bool(len(crd) == ours.nnz and np.array_equal(pos, np.asarray(ours.pos)) and np.array_equal(crd, ours.crd))
    rel = float(np.max(np.abs(vals - ours.vals) / np.maximum(np.abs(ours.vals), 1e-300))) if same and len(crd) else None
    return {"ms_per_step": round(sec * 1e3, 4), "value": round(2 * products / sec / 1e9, 3), "unit": "GFLOP/s",
            "same_structure": same, "max_rel_value_diff": rel}


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

    # dominant kernel: the longest of the row kernels (the short-row path's
    # two kernels, or the binned path's per-row kernels)
    sec, dom, dom_times, bd = timed_steps(
        ctx, step, args.steps, args.warmup,
        lambda bd: max((k for k in bd if k.startswith(("row_", "short_"))), key=lambda k: bd[k][0]))
    calibration_copy(ctx, args.calibrate_copy)
    C = out["C"]
    P = ctx.stats()["products"]
    nnz = C.nnz
    m = A.nrows
    # bytes each row kernel reads and writes itself, per launch:
    #   short_rows     A row pointers (4 B/row) and entries (12 B), B_pos
    #                  gathers (8 B per A entry), B entries (12 B per
    #                  product), packed entries (12 B per output entry) and
    #                  row counts (4 B/row) written
    #   short_compact  row counts (4 B/row) and packed entries (12 B) read, C
    #                  (4 B/row + 12 B per entry) written
    #   row_*          as short_rows plus per-row metadata (row id 4 + offset
    #                  8 + nnz 4 B)
    io = {"short_rows": m * 8 + A.nnz * 20 + P * 12 + nnz * 12,
          "short_compact": m * 8 + nnz * 24}
    dom_bytes = io.get(dom, m * 8 + A.nnz * 12 + A.nnz * 8 + P * 12 + nnz * 12 + m * 16)
    # compulsory bytes: A and B read once, C written once (int32 pointers)
    compulsory = (4 * (m + 1) + 12 * A.nnz) + (4 * (B.nrows + 1) + 12 * B.nnz) + (4 * (m + 1) + 12 * nnz)
    res = {
        "metric": "SpGEMM GFLOP/s",
        "value": round(2 * P / sec / 1e9, 3),
        "unit": "GFLOP/s",
        "ms_per_step": round(sec * 1e3, 4),
        "dtype": "f64",
        "config": {"workload": "rowwise_hash_spgemm_10k_r8", "m": m, "n": B.ncols, "nnz_per_row": 8,
                   "products": int(P), "nnz_C": int(nnz), "seed": 0},
        "roofline": roofline(dom, dom_times, args.steps, compulsory, dom_bytes, "rowwise"),
        "compulsory_bytes": int(compulsory),
        "compulsory_frac_of_hbm_roofline": round(compulsory / sec / 1e9 / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in sorted(bd.items())},
    }
    if not args.no_rocsparse:
        res["rocsparse"] = rocsparse_leg(ctx, dA, dB, C, P, args.steps, args.warmup)
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
        # the CAP-parallel hash (CSR_CSR_hash_parallel.h): cap = cores - 1
        # workers and the merger on the last core; at least 5 workers keep
        # each one's entries under 2 w_cap, so no flush runs
        cap = max(5, cores - 1)
        G = oracle.spgemm_gustavson(A, B, cores)
        oracle.spgemm_hash_parallel(A, B, 1 << 16, cap)
        reps, t0 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t0 < min(args.cpu_seconds, 3.0):
            Hp, flushes = oracle.spgemm_hash_parallel(A, B, 1 << 16, cap)
            reps += 1
        dtp = (time.perf_counter() - t0) / reps
        same = bool(np.array_equal(Hp.pos, G.pos) and np.array_equal(Hp.crd, G.crd) and np.array_equal(Hp.vals, G.vals))
        res["cpu_baseline_hash_parallel"] = {"value": round(2 * P / dtp / 1e9, 3), "unit": "GFLOP/s",
                                             "cores": cap + 1, "kind": "port", "flushes": flushes,
                                             "same_as_gustavson": same,
                                             "sample": f"full workload, {cap} workers + merger, w_cap 2^16, "
                                                       f"{reps} runs"}
    ctx.close()
    return res


# --------------------------------------------------------------------------- outer product

def run_outer(args):
    """Outer-product SpGEMM C = A*B, B a 200,000 x 200,000 CSR with Zipf(1.8)
    row lengths on [1, 20,000] and uniform distinct columns, FP64 U[-1,1],
    seed 1; A = B^T, so DCSC(A) shares B's arrays.  Output sorted CSR with
    int64 row pointers."""
    from sparsews_amd import formats, gen
    from sparsews_amd.device import Context, DeviceDCSR, set_outer_method, spgemm_outer
    B = gen.outer_operand()
    N = B.nrows
    Bd = formats.dcsr_from_csr(B)
    ctx = Context(0)
    set_outer_method(ctx, args.outer_method, args.outer_window)
    dA, dB = DeviceDCSR.upload(ctx, Bd), DeviceDCSR.upload(ctx, Bd)
    out = {}

    def step():
        out["C"] = spgemm_outer(ctx, dA, dB, N, N, N)

    def pick(bd):
        names = {1: ("radix_scatter", "radix_hist", "outer_expand", "outer_reduce"),
                 2: ("bk_expand", "bk_count", "bk_reduce"),
                 3: ("fz_count", "fz_reduce"), 4: ("fz_count", "fz_reduce")}[ctx.stats()["outer_method"]]
        return max(names, key=lambda k: bd.get(k, (0, 1))[0])

    sec, dom, dom_times, bd = timed_steps(ctx, step, args.steps, args.warmup, pick)
    calibration_copy(ctx, args.calibrate_copy)
    st = ctx.stats()
    P, nnz = st["products"], out["C"].nnz
    key_bytes = 4
    # algorithmic bytes of each kernel over the timed region (per step):
    # LSD method
    #   radix_scatter  read + write of key (4 B) and value (8 B) per product per pass
    #   radix_hist     read of the key per product per pass
    #   outer_expand   12 B written + 12 B of B gathered per product
    #   outer_reduce   12 B read per product + 12 B written per distinct entry
    # bucket method
    #   bk_expand      10 B written (2 B window column + 8 B value) + 12 B of B gathered per product
    #   bk_count       2 B read per product
    #   bk_reduce      10 B read per product + 12 B written per distinct entry
    passes = bd["radix_scatter"][1] // st["chunks"] if "radix_scatter" in bd else 0
    # fused method
    #   fz_count       4 B of B columns gathered per product
    #   fz_reduce      12 B of B gathered per product + 12 B written per distinct entry
    if st["outer_method"] in (3, 4):
        per_step = {"fz_count": 4 * P, "fz_reduce": 12 * P + 12 * nnz}
    elif st["outer_method"] == 2:
        per_step = {"bk_expand": 22 * P, "bk_count": 2 * P, "bk_reduce": 10 * P + 12 * nnz}
    else:
        per_step = {
            "radix_scatter": 2 * (key_bytes + 8) * P * passes,
            "radix_hist": key_bytes * P * passes,
            "outer_expand": 24 * P,
            "outer_reduce": 12 * P + 12 * nnz,
        }
    # compulsory bytes: both operands read once, int64 row pointers and
    # (column, value) of every output entry written once
    n_in = 2 * (4 * (N + 1) + 12 * B.nnz)
    compulsory = n_in + 8 * (N + 1) + 12 * nnz
    roof = roofline(dom, dom_times, args.steps, compulsory, per_step[dom], "outer")
    res = {
        "metric": "SpGEMM GFLOP/s",
        "value": round(2 * P / sec / 1e9, 3),
        "unit": "GFLOP/s",
        "ms_per_step": round(sec * 1e3, 3),
        "dtype": "f64",
        "config": {"workload": "outer_product_spgemm_200k_zipf1.8", "n": N, "zipf_s": 1.8, "row_cap": 20000,
                   "nnz_B": int(B.nnz), "products": int(P), "nnz_C": int(nnz), "chunks": st["chunks"],
                   "sort_method": {1: "lsd_radix", 2: "row_window_buckets", 3: "fused_row_windows",
                                   4: "fused_row_windows_atomic"}[st["outer_method"]],
                   "window_bits": st["window_bits"], "sort_key_bits": st["sort_bits"],
                   "radix_passes": int(passes), "seed": 1},
        "roofline": roof,
        "compulsory_bytes": int(compulsory),
        "compulsory_frac_of_hbm_roofline": round(compulsory / sec / 1e9 / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0], 3) for k, v in sorted(bd.items())},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        res["cpu_baseline"], res["cpu_baseline_coord_sequential"] = outer_cpu_baseline(B, args.cpu_seconds)
    return res


def outer_cpu_baseline(B, seconds):
    """The oracle on bounded row samples of the same product (rows of A = B^T):
    OpenMP Gustavson on all cores, and the reference's sequential coordinate
    workspace (w_cap = 2^floor(log2 nnz(B))) on one core."""
    import oracle
    from sparsews_amd import formats
    from sparsews_amd.gen import CSR
    cores = cpu_cores()
    AT = B.transpose()
    rng = np.random.default_rng(7)
    L = B.row_lengths()
    ub = np.zeros(B.ncols, dtype=np.int64)
    np.add.at(ub, B.crd, np.repeat(L, L))

    def rows_sample(budget):
        order = rng.permutation(B.nrows)
        csum = np.cumsum(ub[order])
        k = int(np.searchsorted(csum, budget)) + 1
        rows = np.sort(order[:k])
        lens = AT.row_lengths()[rows]
        idx = np.concatenate([np.arange(AT.pos[r], AT.pos[r + 1]) for r in rows])
        return CSR(len(rows), B.nrows, np.concatenate([[0], np.cumsum(lens)]).astype(np.int32),
                   AT.crd[idx], AT.vals[idx]), int(ub[rows].sum())

    A1, P1 = rows_sample(2.5e8)
    t0 = time.perf_counter()
    oracle.spgemm_gustavson(A1, B, cores)
    t1 = time.perf_counter() - t0
    scale = max(1.0, seconds / max(t1, 1e-3))
    # at most 1e9 products: the oracle's CSR output has int32 row pointers
    A2, P2 = rows_sample(min(2.5e8 * scale, 1e9))
    t0 = time.perf_counter()
    oracle.spgemm_gustavson(A2, B, cores)
    t2 = time.perf_counter() - t0
    omp = {"value": round(2 * P2 / t2 / 1e9, 4), "unit": "GFLOP/s", "cores": cores, "kind": "port",
           "sample": f"{A2.nrows} random output rows ({P2} products) of the 200k workload, "
                     f"OpenMP Gustavson restatement (CSR_CSR_taco_parallel.h), {t2:.1f} s"}
    A3, P3 = rows_sample(2e6)
    Ac, Bd = formats.dcsc_from_csr(A3), formats.dcsr_from_csr(B)
    w_cap = 1 << int(np.log2(B.nnz))
    t0 = time.perf_counter()
    oracle.outer_coord(Ac, Bd, A3.nrows, B.ncols, w_cap)
    t3 = time.perf_counter() - t0
    seq = {"value": round(2 * P3 / t3 / 1e9, 5), "unit": "GFLOP/s", "cores": 1, "kind": "port",
           "sample": f"{A3.nrows} random output rows ({P3} products), sequential coordinate workspace "
                     f"(DCSC_DCSR_coord_index_chase.h, w_cap {w_cap}), {t3:.1f} s"}
    return omp, seq


# --------------------------------------------------------------------------- tensors

def run_mttkrp(args):
    """Mode-1 MTTKRP, X 50,000 x 40,000 x 30,000 with 20M distinct nonzeros
    (per-mode Zipf 1.5 indices), FP64 U[0,1); dense factors of rank 32.  The
    CSF (levels j, k, i: output mode as leaf, as the reference lays it out)
    is resident in HBM; a step is the whole MTTKRP from that layout."""
    from sparsews_amd import formats, gen
    from sparsews_amd.device import Context, DeviceArray, DeviceCSF, mttkrp
    X, Bf, Cf = gen.mttkrp_tensor()
    I, J, K = X.dims
    R = Bf.shape[1]
    csf = formats.csf3_from_coo(X, (1, 2, 0))
    ctx = Context(0)
    dX = DeviceCSF.upload(ctx, csf)
    dB = DeviceArray.from_host(ctx, np.ascontiguousarray(Bf.ravel()))
    dC = DeviceArray.from_host(ctx, np.ascontiguousarray(Cf.ravel()))
    out = {}

    def step():
        out["r"] = mttkrp(ctx, dX, I, J, K, R, dB, dC)

    names = ("mttkrp_accum", "radix_scatter", "mttkrp_keys")
    sec, dom, dom_times, bd = timed_steps(ctx, step, args.steps, args.warmup,
                                          lambda bd: max(names, key=lambda k: bd.get(k, (0, 1))[0]))
    calibration_copy(ctx, args.calibrate_copy)
    nnz = X.nnz
    nrows = out["r"][0].count
    passes = bd["radix_scatter"][1] if "radix_scatter" in bd else 0
    per_step = {
        # sorted (key, value) records + factor matrices once + the output rows
        "mttkrp_accum": 16 * nnz + 8 * R * (J + K) + 8 * R * nrows,
        "radix_scatter": 2 * 16 * nnz * passes,
        "mttkrp_keys": 4 * nnz + 4 * nnz + 8 * nnz,
    }
    # compulsory bytes: X as COO (three int32 coordinates + FP64 value per
    # nonzero), both factor matrices once, the dense output once
    compulsory = 20 * nnz + 8 * R * (J + K) + 8 * R * I
    res = {
        "metric": "MTTKRP GFLOP/s", "value": round(3 * R * nnz / sec / 1e9, 3), "unit": "GFLOP/s",
        "ms_per_step": round(sec * 1e3, 4), "dtype": "f64",
        "config": {"workload": "mttkrp_mode1_50k_40k_30k_20M_R32", "dims": [I, J, K], "nnz": nnz, "rank": R,
                   "zipf_s": 1.5, "rows_out": int(nrows), "seed": 0},
        "roofline": roofline(dom, dom_times, args.steps, compulsory, per_step[dom], "mttkrp"),
        "compulsory_bytes": int(compulsory),
        "compulsory_frac_of_hbm_roofline": round(compulsory / sec / 1e9 / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in sorted(bd.items())},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        import oracle
        from sparsews_amd.gen import COO3
        rng = np.random.default_rng(9)
        sel = np.sort(rng.choice(nnz, 1_000_000, replace=False))
        Xs = COO3(X.dims, X.idx[sel], X.vals[sel])
        cs = formats.csf3_from_coo(Xs, (1, 2, 0))
        w_cap = 1 << int(np.log2(I * R * 0.1))
        t0 = time.perf_counter()
        oracle.mttkrp_hash(cs, formats.dense_as_dcsr(Bf), formats.dense_as_dcsr(Cf), I, R, w_cap)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(3 * R * Xs.nnz / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
                               "kind": "port",
                               "sample": f"1,000,000 random nonzeros of the tensor, sequential hash MTTKRP "
                                         f"(COO_CSF_DCSR_DCSR_hash_free.h, w_cap {w_cap}), {dt:.1f} s"}
    return res


def run_ttm(args):
    """TTM along mode 3, X 100,000 x 80,000 x 2,000 with 10M distinct nonzeros
    (modes 1-2 Zipf 1.5, mode 3 uniform), FP64 U[-1,1], times U 2,000 x 16.
    The CSF (levels i, k, j: contracted mode in the middle) is resident; the
    output is the sorted (i, j) fiber list with dense length-16 vectors."""
    from sparsews_amd import formats, gen
    from sparsews_amd.device import Context, DeviceArray, DeviceCSF, set_ttm_method, ttm
    X, U = gen.ttm_tensor()
    I, J, K = X.dims
    R = U.shape[1]
    csf = formats.csf3_from_coo(X, (0, 2, 1))
    ctx = Context(0)
    set_ttm_method(ctx, args.ttm_method)
    dX = DeviceCSF.upload(ctx, csf)
    dU = DeviceArray.from_host(ctx, np.ascontiguousarray(U.ravel()))
    out = {}

    def step():
        out["r"] = ttm(ctx, dX, I, J, K, R, dU)

    sec, dom, dom_times, bd = timed_steps(ctx, step, args.steps, args.warmup,
                                          lambda bd: max(bd, key=lambda k: bd[k][0]))
    calibration_copy(ctx, args.calibrate_copy)
    nnz = X.nnz
    F = out["r"][0].count
    roots_path = "tw_heavy" in bd or "tw_light" in bd
    # sorting path: the nonzeros are grouped by one stable radix sort on the
    # (i, j) bits of their packed (i, j, k) keys: 8-bit digits
    sort_bits = max(1, (I - 1).bit_length()) + max(1, (J - 1).bit_length())
    passes = -(-sort_bits // 8)
    # roots path: nonzeros of heavy roots (more than 256) and light roots
    rl = np.diff(csf.l3_pos[csf.l2_pos.astype(np.int64)])
    heavy_nnz = int(rl[rl > 256].sum())
    light_nnz = nnz - heavy_nnz
    per_step = {
        # sorting path
        #   ttm_accum      per nonzero key + value read; per fiber R values written; U once
        #   radix_scatter  per nonzero and pass key + value read and written
        #   ttm_keys       per nonzero leaf fiber, fiber root, three coordinates and
        #                  value read; key + value written
        "ttm_accum": nnz * (8 + 8) + 8 * R * F + 8 * R * K,
        "radix_scatter": nnz * 32 * passes,
        "ttm_keys": nnz * (4 + 4 + 4 + 4 + 4 + 8 + 16),
        # roots path
        #   tw_scatter     per heavy nonzero column, value and k read, key + value written
        #   tw_heavy       per heavy nonzero key + value read; its fibers' R values and
        #                  (i, j) written
        #   tw_light       per light nonzero column, value and k read; its fibers written
        #   tw_hist        per heavy nonzero column read
        "tw_scatter": heavy_nnz * (4 + 8 + 4 + 4 + 8),
        "tw_heavy": heavy_nnz * (4 + 8),
        "tw_light": light_nnz * (4 + 8 + 4),
        "tw_hist": heavy_nnz * 4,
    }
    # compulsory bytes: X as COO once, U once, every output fiber's (i, j)
    # and R values written once
    compulsory = 20 * nnz + 8 * R * K + F * (8 + 8 * R)
    res = {
        "metric": "TTM GFLOP/s", "value": round(2 * R * nnz / sec / 1e9, 3), "unit": "GFLOP/s",
        "ms_per_step": round(sec * 1e3, 4), "dtype": "f64",
        "config": {"workload": "ttm_mode3_100k_80k_2k_10M_R16", "dims": [I, J, K], "nnz": nnz, "rank": R,
                   "fibers_out": int(F), "seed": 0,
                   "method": "roots_and_windows" if roots_path else "radix_sort",
                   **({"heavy_root_nnz": heavy_nnz, "light_root_nnz": light_nnz} if roots_path else
                      {"sort_bits": sort_bits, "radix_passes": passes})},
        "roofline": roofline(dom, dom_times, args.steps, compulsory, per_step.get(dom, compulsory), "ttm"),
        "compulsory_bytes": int(compulsory),
        "compulsory_frac_of_hbm_roofline": round(compulsory / sec / 1e9 / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in sorted(bd.items())},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        import oracle
        from sparsews_amd.gen import COO3
        rng = np.random.default_rng(10)
        sel = np.sort(rng.choice(nnz, 1_000_000, replace=False))
        Xs = COO3(X.dims, X.idx[sel], X.vals[sel])
        cs = formats.csf3_from_coo(Xs, (0, 2, 1))
        t0 = time.perf_counter()
        oracle.ttm_hash_il(cs, formats.dense_as_dcsr(U), I, J, R)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(2 * R * Xs.nnz / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
                               "kind": "port",
                               "sample": f"1,000,000 random nonzeros of the tensor, sequential hash_il TTM "
                                         f"(COO_CSF_DCSR_hash_il.h), {dt:.1f} s"}
    return res


# --------------------------------------------------------------------------- batched / served

P99_TARGET_MS = 2.0


def latency_bound_rate(run, n, sat, steps=7):
    """Highest Poisson arrival rate at which run(arrivals) keeps the p99
    latency of its n requests within P99_TARGET_MS: bisection (geometric)
    between 2% and 100% of the saturation rate `sat`, the same unit-rate
    arrival pattern scaled.  Returns (rate, p99 at that rate); rate 0 when
    even the lowest fails."""
    gaps = np.random.default_rng(777).exponential(1.0, n)
    unit = np.cumsum(gaps) - gaps[0]

    def p99_at(rate):
        arr = unit / rate
        fin = run(arr)
        return float(np.percentile((fin - arr) * 1e3, 99))

    lo, hi = 0.02 * sat, sat
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


def run_serve(args):
    """Batched / served SpGEMM: a stream of requests A_b * B_b, square n_b
    log-uniform in [256, 65,536], half with uniform 2..64 nonzeros per row
    and half Zipf(1.8) capped at 4,096, FP64 U[-1,1], seed 0.  Operands are
    resident in HBM before the run; results stay on the device (in the
    engines' pools, or the context's output for the plain path).  Measured:
    saturation throughput (all requests queued at t = 0) of the plain path
    (one row-wise call per request) and of the continuously batched path
    (engines with one hipGraph launch per batch, a products cap per batch,
    and a large lane: requests of at least 2^serve_large_log2 products run
    one per engine on serve_large_engines reserved engines); then Poisson
    arrivals at 80% of the plain path's capacity for both, and at 3x that
    rate for the batched path, with p50/p99 latency (finish - arrival).
    Last, the latency-bound rate of each path: the highest Poisson arrival
    rate (bisection over the rate, the first 2,048 requests) at which the
    p99 latency stays within 2 ms."""
    from sparsews_amd import gen
    from sparsews_amd.device import Context
    from sparsews_amd.serve import DeviceRequest, Engine, latency_summary, serve
    stream = gen.request_stream(count=args.requests, seed=0)
    ctx = Context(0)
    t0 = time.perf_counter()
    reqs, products = [], 0
    in_bytes = 0
    for r in stream:
        A, B = gen.request_operands(r)
        d = DeviceRequest.upload(ctx, A, B)
        reqs.append(d)
        products += d.products
        in_bytes += 4 * (A.nrows + 1) + 12 * A.nnz + 4 * (B.nrows + 1) + 12 * B.nnz
        del A, B
        if len(reqs) % 1024 == 0:
            log(f"serve: {len(reqs)} requests generated and uploaded ({time.perf_counter() - t0:.0f} s)")
    N = len(reqs)
    # the autotuner's table (sparsews_amd/tuning.json) when present: row bins
    # on the context (the engines copy them) and the serving shape; options
    # given on the command line win
    from sparsews_amd import autotune
    table = None if args.no_tuning else autotune.load()
    shape = autotune.apply(ctx, table)
    n_engines = args.serve_engines if args.serve_engines is not None else shape["engines"]
    cap = (1 << args.serve_cap_log2) if args.serve_cap_log2 is not None else shape["max_batch_products"]
    large_engines = (args.serve_large_engines if args.serve_large_engines is not None
                     else shape["large_engines"])
    large = (((1 << args.serve_large_log2) if args.serve_large_log2 is not None else shape["large_products"])
             if large_engines > 0 else 0)
    lanes = dict(max_batch_products=cap, large_products=large, large_engines=large_engines)
    engines = [Engine(ctx) for _ in range(n_engines)]
    zero = np.zeros(N)

    def batched(reqs_, arr_, **kw):
        return serve(ctx, engines, reqs_, arr_, "batched", **lanes, **kw)

    # warm-up: one pass of each mode on a prefix
    serve(ctx, [], reqs[:32], zero[:32], "plain")
    batched(reqs[:32], zero[:32])
    f_p, _, st_p, nnz_p = serve(ctx, [], reqs, zero, "plain", with_nnz=True)
    f_b, bid, st_b, nnz_b = batched(reqs, zero, with_nnz=True)
    assert np.array_equal(nnz_p, nnz_b), "plain and batched output sizes differ"
    sat_plain = N / st_p["wall_seconds"]
    sat_batch = N / st_b["wall_seconds"]
    log(f"serve: saturation plain {sat_plain:.1f} req/s, batched {sat_batch:.1f} req/s")
    rng = np.random.default_rng(12345)
    lam = 0.8 * sat_plain
    gaps = rng.exponential(1.0 / lam, N)
    arr = np.cumsum(gaps) - gaps[0]
    f_pl, _, st_pl = serve(ctx, [], reqs, arr, "plain")
    f_bl, _, st_bl = batched(reqs, arr)
    arr3 = arr / 3.0
    f_b3, _, st_b3 = batched(reqs, arr3)
    nb = min(N, 2048)
    bound = {"p99_target_ms": P99_TARGET_MS, "requests": nb}
    runs = {"plain": lambda a: serve(ctx, [], reqs[:nb], a, "plain")[0],
            "batched": lambda a: batched(reqs[:nb], a)[0]}
    for mode, sat in (("plain", sat_plain), ("batched", sat_batch)):
        rate, p99 = latency_bound_rate(runs[mode], nb, sat)
        bound[mode] = {"req_per_s": round(rate, 2), "p99_ms": round(p99, 4)}
        log(f"serve: {mode} sustains {rate:.1f} req/s at p99 {p99:.3f} ms")
    bound["speedup"] = round(bound["batched"]["req_per_s"] / max(bound["plain"]["req_per_s"], 1e-9), 3)
    for e in engines:
        e.close()
    ctx_bins = {k: v for k, v in ctx.rowwise_tuning().items() if k in ("wave_ub", "wave2_ub", "block_ub", "block2_ub")}
    sizes = np.array([r.products for r in reqs])
    del reqs[:], runs
    ctx.close()
    big = sizes >= np.percentile(sizes, 99)
    lat_b3 = (f_b3 - arr3) * 1e3
    # compulsory bytes of the stream: every request's operands read once and
    # its output (int32 row pointers, column and value per entry) written
    # once; over the batched run's wall time at saturation (the batch
    # pipeline's kernels run inside graphs, on several streams)
    out_bytes = int(sum(4 * (r.n + 1) for r in stream)) + 12 * int(nnz_b.sum())
    compulsory = in_bytes + out_bytes
    achieved = compulsory / st_b["wall_seconds"] / 1e9
    res = {
        "metric": "batched SpGEMM req/s", "value": round(sat_batch, 2), "unit": "req/s",
        "ms_per_step": round(st_b["wall_seconds"] * 1e3, 3), "dtype": "f64",
        "config": {"workload": "served_spgemm_4096_requests", "requests": N, "n_range": [256, 65536],
                   "products_total": int(products), "nnz_total": int(nnz_b.sum()),
                   "products_median": int(np.median(sizes)), "products_q99": int(np.percentile(sizes, 99)),
                   "engines": n_engines, "max_batch_products": cap, "large_lane_products": large,
                   "large_lane_engines": large_engines, "seed": 0,
                   "tuning": "sparsews_amd/tuning.json" if table is not None else "defaults",
                   "row_bins": ctx_bins},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "batch pipeline (all graphs, wall time of the saturated batched run)",
                     "algorithmic_bytes_per_launch": None, "compulsory_bytes_stream": int(compulsory)},
        "saturation": {"plain_req_per_s": round(sat_plain, 2), "batched_req_per_s": round(sat_batch, 2),
                       "speedup": round(sat_batch / sat_plain, 3), "batches": st_b["batches"],
                       "fallback_requests": st_b["fallback_requests"],
                       "plain_products_per_s": round(products / st_p["wall_seconds"] / 1e9, 3),
                       "batched_products_per_s": round(products / st_b["wall_seconds"] / 1e9, 3),
                       "unit_products": "G products/s"},
        "poisson_at_80pct_plain": {"rate_req_per_s": round(lam, 2),
                                   "plain": {**latency_summary(arr, f_pl),
                                             "req_per_s": round(N / st_pl["wall_seconds"], 2)},
                                   "batched": {**latency_summary(arr, f_bl),
                                               "req_per_s": round(N / st_bl["wall_seconds"], 2),
                                               "batches": st_bl["batches"]}},
        "latency_bound": bound,
        "poisson_batched_at_3x": {"rate_req_per_s": round(3 * lam, 2), **latency_summary(arr3, f_b3),
                                  "req_per_s": round(N / st_b3["wall_seconds"], 2), "batches": st_b3["batches"],
                                  "p99_ms_excluding_top1pct_requests": round(float(np.percentile(lat_b3[~big], 99)), 4)},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = serve_cpu_baseline(stream, args.cpu_seconds)
    return res


def serve_cpu_baseline(stream, seconds):
    """The oracle's OpenMP Gustavson (CSR_CSR_taco_parallel.h) on all host
    cores, one request after another from the start of the same request
    stream (operands regenerated from their seeds), until about `seconds`
    of work: requests per second on the CPU."""
    import oracle
    from sparsews_amd import gen
    cores = cpu_cores()
    done, products, busy = 0, 0, 0.0
    for r in stream:
        A, B = gen.request_operands(r)
        t0 = time.perf_counter()
        oracle.spgemm_gustavson(A, B, cores)
        busy += time.perf_counter() - t0
        products += gen.row_products(A, B)[0]
        done += 1
        if busy >= seconds:
            break
    return {"value": round(done / busy, 2), "unit": "req/s", "cores": cores, "kind": "port",
            "sample": f"first {done} requests of the stream ({products} products), OpenMP Gustavson restatement "
                      f"(CSR_CSR_taco_parallel.h), {busy:.1f} s"}


def run_transpose(args):
    """CSR transpose on the device (the transpose pass of the non-fused
    transposed-output product), 200,000 x 200,000 with 32 uniform columns per
    row (6.4M entries), FP64.  Input and preallocated outputs are resident;
    a step is one spws_csr_transpose_dev call.  Value = algorithmic bytes per
    second: every entry read (12 B) and written (12 B) once, both row-pointer
    arrays once."""
    import ctypes as C
    from sparsews_amd import gen
    from sparsews_amd._lib import check, lib
    from sparsews_amd.device import Context, DeviceArray, DeviceCSR
    m = n = 200_000
    A = gen.uniform_csr(m, n, 32, 3, 0)
    nnz = int(A.pos[-1])
    ctx = Context(0)
    dA = DeviceCSR.upload(ctx, A)
    out_pos = DeviceArray(ctx, n + 1, np.int32)
    out_crd = DeviceArray(ctx, nnz, np.int32)
    out_val = DeviceArray(ctx, nnz, np.float64)

    def step():
        check(lib().spws_csr_transpose_dev(ctx.h, m, n, nnz, dA.pos.ptr, dA.crd.ptr, dA.vals.ptr, out_pos.ptr,
                                           out_crd.ptr, out_val.ptr))

    names = ("radix_scatter", "tr_keys", "tr_out")
    sec, dom, dom_times, bd = timed_steps(ctx, step, args.steps, args.warmup,
                                          lambda bd: max(names, key=lambda k: bd.get(k, (0, 1))[0]))
    calibration_copy(ctx, args.calibrate_copy)
    # spot check against the host transpose (the first 1,000 output rows)
    T = DeviceCSR(n, m, nnz, out_pos, out_crd, out_val).download()
    cols = A.crd.astype(np.int64)
    order = np.argsort(cols, kind="stable")
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(A.pos))
    lim = int(np.searchsorted(cols[order], 1000))
    assert np.array_equal(T.crd[:lim], rows[order][:lim]) and np.array_equal(T.vals[:lim], A.vals[order][:lim])
    passes = -(-max(1, (n - 1).bit_length()) // 8)
    per_step = {
        # per entry and pass: key + value read and written
        "radix_scatter": nnz * 32 * passes,
        # per entry: column and value read, key and value written; row pointers
        "tr_keys": nnz * (4 + 8 + 8 + 8) + 4 * (m + 1),
        # per entry: key and value read, row index and value written
        "tr_out": nnz * (8 + 8 + 4 + 8),
    }
    algorithmic = 24 * nnz + 4 * (m + 1) + 4 * (n + 1)
    res = {
        "metric": "CSR transpose GB/s", "value": round(algorithmic / sec / 1e9, 2), "unit": "GB/s",
        "ms_per_step": round(sec * 1e3, 4), "dtype": "f64",
        "config": {"workload": "csr_transpose_200k_32_per_row", "rows": m, "cols": n, "nnz": nnz,
                   "radix_passes": passes, "seed": 3},
        "roofline": roofline(dom, dom_times, args.steps, algorithmic, per_step[dom], "transpose"),
        "compulsory_bytes": int(algorithmic),
        "compulsory_frac_of_hbm_roofline": round(algorithmic / sec / 1e9 / HBM_PEAK_GBS, 5),
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in sorted(bd.items())},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        import oracle
        reps, t0 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t0 < 3.0:
            oracle.csr_transpose(A)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        res["cpu_baseline"] = {"value": round(algorithmic / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                               "sample": f"full workload, transpose() restatement (nonfuse/CSC_CSR_T_hash.h:352), "
                                         f"{reps} runs"}
    return res


# (rows m, inner k, columns n, entries per row r of both operands, seed);
# nnz(C) is about m * r^2 while r^2 is far below n, so the cases span the
# output-nnz decades 10^3 .. 10^8
FUSION_CASES = [
    (1_000, 1_000, 1_000, 2, 11), (2_000, 2_000, 2_000, 2, 12), (600, 600, 600, 3, 13),
    (4_000, 4_000, 4_000, 3, 14), (10_000, 10_000, 10_000, 2, 15), (2_000, 2_000, 2_000, 6, 16),
    (20_000, 20_000, 20_000, 4, 17), (50_000, 50_000, 50_000, 3, 18), (10_000, 10_000, 10_000, 8, 19),
    (100_000, 100_000, 100_000, 4, 20), (100_000, 100_000, 100_000, 8, 21), (50_000, 50_000, 50_000, 12, 22),
    (200_000, 200_000, 200_000, 8, 23), (500_000, 500_000, 500_000, 8, 24), (1_000_000, 1_000_000, 1_000_000, 8, 25),
]
FUSION_BINS = [0, 1e4, 1e5, 1e6, 1e7, 1e8]


def _median_time(ctx, fn, budget_s=0.25, lo=3, hi=40):
    """Median wall time of fn (each call followed by a stream sync): at
    least lo calls, more while the budget lasts, at most hi."""
    times = []
    t_start = time.perf_counter()
    while len(times) < lo or (len(times) < hi and time.perf_counter() - t_start < budget_s):
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def _bin_stats(xs, ys):
    out = []
    for lo, hi in zip(FUSION_BINS[:-1], FUSION_BINS[1:]):
        sel = [y for x, y in zip(xs, ys) if lo <= x < hi]
        if sel:
            out.append({"output_nnz": [lo, hi], "cases": len(sel),
                        "gmean": round(float(np.exp(np.mean(np.log(sel)))), 3),
                        "min": round(min(sel), 3), "max": round(max(sel), 3)})
    return out


def run_fusion(args):
    """Fused vs non-fused transposed-output product, C = (A B)^T with A CSC
    (m x k) and B CSR (k x n), over FUSION_CASES (uniform random operands, r
    distinct columns per row, FP64): the reference's one published figure,
    the fusion speedup binned by output nnz (scripts/draw_bars.py:7-58, bins
    [0, 1e4, 1e5, 1e6, 1e7, 1e8), geometric mean, min and max per bin).
    Every array stays on the device.  Fused: one outer product B^T A^T, whose
    row-major result is already (A B)^T.  Non-fused (nonfuse/CSC_CSR_T_hash.h:
    compute_nofuse at :232, transpose at :352): the outer product A B, then
    the transpose of its result into preallocated arrays.  A leg's time is
    the median of repeated calls, each ended by a stream sync; fused and
    non-fused outputs are compared on every case up to 2*10^7 entries (bit
    for bit when the automatic method sums in order, within 1e-12 when it is
    the atomic-add variant).  The CPU columns run the same two pipelines through the oracle
    (the reference's coordinate workspace and transpose, one core) on the
    cases up to 2*10^6 products."""
    from sparsews_amd import formats, gen
    from sparsews_amd._lib import check, lib
    from sparsews_amd.device import Context, DeviceArray, DeviceDCSR, spgemm_outer
    ctx = Context(0)
    cases = []
    big = None
    for (m, k, n, r, seed) in FUSION_CASES:
        A = gen.uniform_csr(m, k, r, seed, 0)
        B = gen.uniform_csr(k, n, r, seed, 1)
        dAc = DeviceDCSR.from_csr(ctx, A.transpose())   # A's columns: the stored outer level
        dB = DeviceDCSR.from_csr(ctx, B)
        R = spgemm_outer(ctx, dAc, dB, m, k, n)
        nnz = R.nnz
        products = ctx.stats()["products"]
        t_pos = DeviceArray(ctx, n + 1, np.int32)
        t_crd = DeviceArray(ctx, max(nnz, 1), np.int32)
        t_val = DeviceArray(ctx, max(nnz, 1), np.float64)
        held = {}

        def transpose():
            X = held["R"]
            check(lib().spws_csr64_transpose_dev(ctx.h, m, n, X.nnz, X.pos.ptr, X.crd.ptr, X.vals.ptr, t_pos.ptr,
                                                 t_crd.ptr, t_val.ptr))

        def nonfused():
            held["R"] = spgemm_outer(ctx, dAc, dB, m, k, n)
            transpose()

        def fused():
            held["F"] = spgemm_outer(ctx, dB, dAc, n, k, m)

        nonfused()
        fused()
        ctx.sync()
        if nnz <= 20_000_000:
            F = held["F"].download()
            assert F.nnz == nnz
            assert np.array_equal(np.asarray(F.pos, np.int64), t_pos.to_host().astype(np.int64))
            assert np.array_equal(F.crd, t_crd.to_host()[:nnz])
            tv = t_val.to_host()[:nnz]
            if ctx.stats()["outer_method"] == 4:  # atomic adds: equal within the reassociation bound
                assert np.max(np.abs(F.vals - tv), initial=0.0) <= 1e-12 * max(1.0, float(np.abs(tv).max(initial=0)))
            else:
                assert np.array_equal(F.vals, tv)
        t_f = t_n = float("inf")
        for _ in range(2):   # alternate the legs so that drift affects both
            t_n = min(t_n, _median_time(ctx, nonfused))
            t_f = min(t_f, _median_time(ctx, fused))
        nonfused()
        t_t = _median_time(ctx, transpose)
        case = {"m": m, "k": k, "n": n, "per_row": r, "seed": seed, "products": int(products), "nnz_C": int(nnz),
                "fused_ms": round(t_f * 1e3, 4), "nonfused_ms": round(t_n * 1e3, 4),
                "transpose_ms": round(t_t * 1e3, 4), "speedup": round(t_n / t_f, 3),
                "outer_method": ctx.stats()["outer_method"],
                "compared": "bit for bit" if ctx.stats()["outer_method"] != 4 else "within 1e-12",
                "checked": nnz <= 20_000_000}
        log(f"bench: fusion case {m}x{k}x{n} r={r}: nnz(C)={nnz} fused {t_f * 1e3:.3f} ms, "
            f"non-fused {t_n * 1e3:.3f} ms (transpose {t_t * 1e3:.3f})")
        if not args.no_cpu_baseline and products <= 2_000_000:
            import oracle
            Ad, Bd = formats.dcsr_from_csr(A.transpose()), formats.dcsr_from_csr(B)
            w_cap = 1 << max(10, int(products).bit_length() - 1)

            def cpu_fused():
                oracle.outer_coord(Bd, Ad, n, m, w_cap)

            def cpu_nonfused():
                pos, crd, vals = oracle.outer_coord(Ad, Bd, m, n, w_cap)
                oracle.csr_transpose(gen.CSR(m, n, pos.astype(np.int32), crd, vals))

            cf = min(_median_time(ctx, cpu_fused, 0.5, 1, 10), _median_time(ctx, cpu_fused, 0.5, 1, 10))
            cn = min(_median_time(ctx, cpu_nonfused, 0.5, 1, 10), _median_time(ctx, cpu_nonfused, 0.5, 1, 10))
            case.update({"cpu_fused_ms": round(cf * 1e3, 4), "cpu_nonfused_ms": round(cn * 1e3, 4),
                         "cpu_speedup": round(cn / cf, 3)})
        cases.append(case)
        if big is None or nnz > big[0]:
            big = (nnz, m, k, n, dAc, dB, A.nnz + B.nnz)
        else:
            for a in (dAc.pos, dAc.crd, dAc.vals, dB.pos, dB.crd, dB.vals):
                a.free()
        for a in (t_pos, t_crd, t_val):
            a.free()
    # roofline of the largest case's fused product: compulsory bytes are both
    # operands read once and the output's int64 row pointers and entries
    # written once
    nnz, m, k, n, dAc, dB, nin = big
    out = {}

    def step():
        out["F"] = spgemm_outer(ctx, dB, dAc, n, k, m)

    names = ("radix_scatter", "radix_hist", "outer_expand", "outer_reduce", "bk_expand", "bk_count", "bk_reduce",
             "fz_count", "fz_reduce")
    sec, dom, dom_times, bd = timed_steps(ctx, step, 3, 1,
                                          lambda b: max((x for x in names if x in b), key=lambda x: b[x][0]))
    compulsory = 4 * (2 * k + 2) + 12 * nin + 8 * (n + 1) + 12 * nnz
    roof = roofline(dom, dom_times, 3, compulsory, compulsory, "fusion")
    roof["kernel_io"] = None
    ctx.close()
    speed = [c["speedup"] for c in cases]
    xs = [c["nnz_C"] for c in cases]
    res = {
        "metric": "fused vs non-fused transposed-output speedup (geometric mean)",
        "value": round(float(np.exp(np.mean(np.log(speed)))), 3), "unit": "x",
        "ms_per_step": round(sum(c["fused_ms"] for c in cases), 4), "dtype": "f64",
        "config": {"workload": "csc_csr_transposed_output_sweep", "cases": len(cases),
                   "operands": "uniform random, r distinct columns per row, m = k = n",
                   "output_nnz_bins": FUSION_BINS},
        "bins": _bin_stats(xs, speed),
        "cases": cases,
        "roofline": roof,
        "roofline_case": {"m": m, "k": k, "n": n, "nnz_C": int(nnz), "ms_per_step": round(sec * 1e3, 4)},
    }
    cpu = [c for c in cases if "cpu_speedup" in c]
    if cpu:
        res["bins_cpu"] = _bin_stats([c["nnz_C"] for c in cpu], [c["cpu_speedup"] for c in cpu])
        res["cpu_baseline"] = {
            "value": round(float(np.exp(np.mean(np.log([c["cpu_speedup"] for c in cpu])))), 3), "unit": "x",
            "cores": 1, "kind": "port",
            "sample": f"{len(cpu)} cases up to 2e6 products: the oracle's coordinate workspace "
                      "(DCSC_DCSR_coord_index_chase.h:194) fused (B^T A^T) vs A B then transpose() "
                      "(nonfuse/CSC_CSR_T_hash.h:352)"}
    return res


WORKLOADS = {"rowwise": run_rowwise, "outer": run_outer, "mttkrp": run_mttkrp, "ttm": run_ttm, "serve": run_serve,
             "transpose": run_transpose, "fusion": run_fusion}
DEFAULT_STEPS = {"rowwise": (50, 5), "outer": (5, 1), "mttkrp": (20, 3), "ttm": (20, 3), "serve": (1, 1),
                 "transpose": (20, 3), "fusion": (3, 1)}
# the default run: the outer product (headline) and one record per other
# workload of the measured set, each with its own roofline and CPU baseline
ALL_HEAD = "outer"
ALL_SUB = ("rowwise", "mttkrp", "ttm", "serve")


def run_one(name, args):
    """One workload with its own step counts (those given on the command line
    win); returns its record with steps and warmup filled in."""
    import copy
    a = copy.copy(args)
    ds, dw = DEFAULT_STEPS[name]
    a.steps = ds if args.steps is None else args.steps
    a.warmup = dw if args.warmup is None else args.warmup
    t0 = time.perf_counter()
    res = WORKLOADS[name](a)
    log(f"bench: {name} done in {time.perf_counter() - t0:.0f} s")
    rec = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"),
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": res.pop("ms_per_step"), "dtype": res.pop("dtype"),
           "config": res.pop("config")}
    rec.update(res)
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="all", choices=sorted(WORKLOADS) + ["all"],
                    help="all (default): the outer product as the headline line with a record for each of "
                         + ", ".join(ALL_SUB) + " under 'workloads'")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rocsparse", action="store_true", help="row-wise workload: skip the rocSPARSE comparison")
    ap.add_argument("--cpu-seconds", type=float, default=None,
                    help="CPU baseline seconds per workload (default 10; 5 in the combined run)")
    ap.add_argument("--requests", type=int, default=4096, help="serve workload: number of requests")
    ap.add_argument("--serve-engines", type=int, default=None,
                    help="serve workload: batch engines (default: the tuning table's, else 4)")
    ap.add_argument("--serve-cap-log2", type=int, default=None,
                    help="serve workload: log2 of the products (output upper bound) per batch "
                         "(default: the tuning table's, else 26)")
    ap.add_argument("--serve-large-log2", type=int, default=None,
                    help="serve workload: requests of at least 2^this many products form the large lane "
                         "(default: the tuning table's, else 24)")
    ap.add_argument("--serve-large-engines", type=int, default=None,
                    help="serve workload: engines reserved for the large lane, 0 = one lane "
                         "(default: the tuning table's, else 1)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="serve workload: ignore sparsews_amd/tuning.json (the engine defaults)")
    ap.add_argument("--outer-method", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="outer workload: 0 automatic, 1 LSD radix sort, 2 row/window buckets, "
                         "3 fused row windows, 4 fused row windows with atomic adds")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after the timed steps, one plain device copy of this many bytes (the reference "
                         "point for rocprofv3 FETCH_SIZE / WRITE_SIZE runs)")
    ap.add_argument("--ttm-method", type=int, default=0, choices=[0, 1, 2],
                    help="ttm workload: 0 automatic, 1 one radix sort of the tensor, 2 roots and column windows")
    ap.add_argument("--outer-window", type=int, default=0,
                    help="outer workload, bucket method: log2 of the column window (0 = automatic)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0
    combined = args.workload == "all"
    if args.cpu_seconds is None:
        args.cpu_seconds = 5.0 if combined else 10.0
    head = run_one(ALL_HEAD if combined else args.workload, args)
    line = {"metric": head.pop("metric"), "value": head.pop("value"), "unit": head.pop("unit"), "n_gpus": 1,
            "steps": head.pop("steps"), "warmup": head.pop("warmup"), "ms_per_step": head.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": head.pop("dtype"),
            "data": "synthetic", "config": head.pop("config")}
    line.update(head)
    if combined:
        line["workloads"] = {name: run_one(name, args) for name in ALL_SUB}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
