#!/usr/bin/env python3
"""Benchmark: nonhydrostatic pressure Poisson solves/s on one MI355X, mixed FP32/FP64 vs FP64.

Workload (default, "regional"): the 1024 x 768 x 60 regional tile (dx = dy = 1.85 km, dz 5 -> 250 m
over 5000 m, 25 % land, partial bottom cells f ~ U(0.1, 1], RHS = divergence of a seeded N(0, 0.2 m/s)
velocity), solved to ||b - A p|| / ||b|| <= 1e-10 from a zero initial guess. One step = one solve.
The mixed solver (FP32 inner PCG with FP32 reductions to 1e-5, FP64 refinement) is the headline value;
the plain FP64 GPU solver is timed beside it on the same inputs.

Prints ONE JSON line. Inputs are resident in HBM (set once before the timed region); the timed region
is K solves on the device. The dominant kernel's HBM roofline comes from HIP events recorded around its
launches inside the captured iteration graph; the CPU baseline is the oracle's FP64 solve (C, OpenMP)
on the host cores, run once on the same inputs.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from orctm_amd import abi, synth  # noqa: E402
from orctm_amd.nh import NHSolver, options  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Poisson solves/s + HBM GB/s, mixed vs FP64 GPU at rel.res 1e-10; p50/p99 lat"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def time_solves(solver, steps, warmup):
    for _ in range(warmup):
        st = solver.solve_resident()
        if st["rc"] != abi.OK:
            raise RuntimeError(f"warmup solve failed: {st} {abi.last_error()}")
    solver.profile(reset=True)
    lat, stats = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        st = solver.solve_resident()
        lat.append((time.perf_counter() - a) * 1e3)
        stats.append(st)
        if st["rc"] != abi.OK:
            raise RuntimeError(f"solve failed: {st} {abi.last_error()}")
    wall = time.perf_counter() - t0
    return wall, lat, stats, solver.profile(reset=True)


def cpu_baseline(g, divg, seconds_hint=30.0):
    import oracle

    A, S = oracle.assemble(g)
    b = oracle.rhs(g, divg)
    t = time.perf_counter()
    p, st = oracle.solve(g, A, S, b, precision=0)
    dt = time.perf_counter() - t
    return {"value": round(1.0 / dt, 6), "unit": "solves/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"one FP64 PCG + line-preconditioner solve of the full {g.nx}x{g.ny}x{g.nz} tile by the "
                      f"C oracle (OpenMP), rel.res {st.rel_res:.1e} in {st.inner_iters} iterations, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="regional", choices=["regional"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink nx, ny (debug only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (systems would shard "
              "across GPUs with no exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)

    spec = synth.config2()
    if args.scale != 1.0:
        spec.nx = max(8, int(spec.nx * args.scale))
        spec.ny = max(8, int(spec.ny * args.scale))
    t = time.perf_counter()
    g, divg = synth.build(spec)
    log(f"[bench] generated {spec.nx}x{spec.ny}x{spec.nz} ({g.cells} cells) in {time.perf_counter() - t:.1f} s")

    results = {}
    sols = {}
    for name, prec in (("mixed", abi.MIXED), ("fp64", abi.FP64)):
        s = NHSolver(g, options(precision=prec, warm_start=False, profile=True))
        s.set_rhs(divg)
        log(f"[bench] {name}: device bytes {s.device_bytes() / 1e9:.2f} GB")
        wall, lat, stats, prof = time_solves(s, args.steps, args.warmup)
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        s.get_solution(pnh)
        sols[name] = g.interior(pnh)
        results[name] = dict(wall=wall, lat=lat, stats=stats, prof=prof)
        log(f"[bench] {name}: {args.steps / wall:.3f} solves/s, last {stats[-1]}, prof {prof}")
        s.close()

    mx, f64 = results["mixed"], results["fp64"]
    K = args.steps
    sps = K / mx["wall"]
    sps64 = K / f64["wall"]
    inner = statistics.mean(st["inner_iters"] for st in mx["stats"])
    inner64 = statistics.mean(st["inner_iters"] for st in f64["stats"])
    cells = g.cells
    # roofline of the fused matvec kernel (dominant per-iteration kernel), FP32:
    # reads 7 coefficients + z + p_old, writes p_new + q = 11 words = 44 B per cell
    prof = mx["prof"]
    mv_avg_ms = prof["matvec_ms"] / max(prof["matvec_launches"], 1)
    up_avg_ms = prof["update_ms"] / max(prof["update_launches"], 1)
    mv_bytes = 44.0 * cells
    up_bytes = 40.0 * cells  # x, p, r, q, l, invD, l(k+1) reads + x, r, z writes (10 words)
    achieved = mv_bytes / (mv_avg_ms * 1e-3) / 1e9
    prof64 = f64["prof"]
    mv64_ms = prof64["matvec_ms"] / max(prof64["matvec_launches"], 1)
    err = float(np.linalg.norm(sols["mixed"] - sols["fp64"]) / max(np.linalg.norm(sols["fp64"]), 1e-300))

    line = {
        "metric": METRIC,
        "value": round(sps, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mx["wall"] / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-inner/fp64-refinement",
        "data": "synthetic (seeded generator, regional-tile recipe)",
        "config": {"workload": f"regional tile {spec.nx}x{spec.ny}x{spec.nz}, 25% land, partial cells, "
                               "mixed FP32 PCG + line preconditioner in FP64 refinement to rel.res 1e-10",
                   "cells": cells, "seed": spec.seed, "rtol": 1e-10, "rtol_inner": 1e-5},
        "fp64_solves_per_s": round(sps64, 4),
        "mixed_over_fp64": round(sps / sps64, 3),
        "cell_iters_per_s": round(cells * inner * K / mx["wall"], 1),
        "inner_iters_mixed": inner,
        "outer_iters_mixed": mx["stats"][-1]["outer_iters"],
        "inner_iters_fp64": inner64,
        "rel_res_mixed": mx["stats"][-1]["rel_res"],
        "rel_res_fp64": f64["stats"][-1]["rel_res"],
        "rel_err_mixed_vs_fp64": err,
        "latency_ms_p50": round(pct(mx["lat"], 50), 3),
        "latency_ms_p99": round(pct(mx["lat"], 99), 3),
        "latency_ms_p50_fp64": round(pct(f64["lat"], 50), 3),
        "latency_ms_p99_fp64": round(pct(f64["lat"], 99), 3),
        "roofline": {"bound": "hbm", "kernel": "k_spmv<float> (fused p-update + 7-point matvec + p.q)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell": 44, "avg_launch_ms": round(mv_avg_ms, 5),
                     "launches": prof["matvec_launches"]},
        "other_kernels": {
            "k_column<float> (x/r update + line solve)": {
                "avg_launch_ms": round(up_avg_ms, 5), "bytes_per_cell": 40,
                "GBps": round(up_bytes / (up_avg_ms * 1e-3) / 1e9, 1)},
            "k_spmv<double>": {"avg_launch_ms": round(mv64_ms, 5), "bytes_per_cell": 88,
                               "GBps": round(88.0 * cells / (mv64_ms * 1e-3) / 1e9, 1)},
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle FP64 solve on the host cores) ...")
        line["cpu_baseline"] = cpu_baseline(g, divg)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
