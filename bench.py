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

# The roofline figure of the hot kernel is the bare matvec y = A x: 7 coefficients + x + y per cell of
# the box, 36 B in FP32 (72 B in FP64). k_bwd performs that matvec (and the back substitution, the
# search-direction update, the deferred x update and p.q) per launch.
MATVEC_BYTES = {4: 36, 8: 72}

# The bytes each kernel itself streams per launch, per cell and per (i, j) column, for word size W
# (4 FP32, 8 FP64) and line-preconditioner factor size F (2: fp16 in FP32 mode, 8 in FP64 mode).
# k_bwd reads w, p_old, x (3 words) and l (F) per cell and writes p, q, x (3 words); per column the 7
#   metric products. With the table form (stats zw_table = 1) it also reads per column the ZW/ZP window
#   (6 words + an int) and, per cell, the uint16 flags (FP32) or, per column, the wet depths the FP64
#   kernel derives the flag bits from (3 words of 4 B); otherwise, per cell, ZW, ZP (2 words) and the
#   flags, and the bottom-face ZW per column.
# k_fwd (r update + forward elimination + r.z, r.r): reads r, q (2 words), l, invD (2 F), writes r, w;
#   in FP32 mode also the per-column scale of invD (4 B per column).
def bwd_bytes(cells, columns, W, table=True):
    F = 2 if W == 4 else 8
    if table:
        if W == 4:
            return cells * (6 * W + F + 2) + columns * (13 * W + 4)
        return cells * (6 * W + F) + columns * (13 * W + 16)
    return cells * (8 * W + F + 2) + columns * 8 * W


def fwd_bytes(cells, columns, W):
    F = 2 if W == 4 else 8
    return cells * (4 * W + 2 * F) + (columns * 4 if W == 4 else 0)


def measured_traffic(name):
    """HBM bytes per launch of a kernel on this workload, from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary kept in profiles/<name>.txt (line "traffic_bytes_per_launch: N"), or None if absent."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return int(float(ln.split(":", 1)[1]))
    except OSError:
        pass
    return None
METRIC = "Poisson solves/s + HBM GB/s, mixed vs FP64 GPU at rel.res 1e-10; p50/p99 lat"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def time_solves(solver, steps, warmup, profiled=True):
    for _ in range(warmup):
        st = solver.solve_resident()
        if st["rc"] != abi.OK:
            raise RuntimeError(f"warmup solve failed: {st} {abi.last_error()}")
    if profiled:
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
    return wall, lat, stats, (solver.profile(reset=True) if profiled else None)


def host_cpu():
    """The host's CPU model (first "model name" of /proc/cpuinfo) and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return model, n


def device_state():
    """The GPU's clocks and power as rocm-smi reports them after the timed solves (card 0: the levels
    its sclk/mclk/fclk lines mark as current, the power drawn and the power cap), or None where the tool
    is absent. Rates of the same build differ from one MI355X to another; this records which one ran."""
    import shutil
    import subprocess

    # rocm-smi is a Python script: run it with this interpreter rather than through its
    # "#!/usr/bin/env" line, so that no process replaces itself with another (under a profiler whose
    # library is preloaded into every child, that child has already opened the GPU)
    tool = shutil.which("rocm-smi")
    if tool is None:
        return None
    try:
        r = subprocess.run([sys.executable, os.path.realpath(tool), "--showclocks", "--showpower", "--showmaxpower",
                            "--showproductname", "--json"], capture_output=True, text=True, timeout=20)
        doc = json.loads(r.stdout)
    except (OSError, ValueError, subprocess.SubprocessError):
        return None
    card = next((v for k, v in sorted(doc.items()) if k.startswith("card")), None)
    if not isinstance(card, dict):
        return None
    keep = ("sclk", "mclk", "fclk", "socclk", "power", "card series", "card model", "gfx version")
    return {k: v for k, v in card.items() if isinstance(v, str) and any(w in k.lower() for w in keep)}


def cpu_baseline(g, divg, reps=5):
    """The oracle's FP64 PCG + line preconditioner (C, OpenMP) on the host cores: one untimed solve, then
    `reps` solves from a zero guess, each timed over its iterations only (the set-up of the inner operator,
    the line factors and the work arrays excluded); value = 1 / median. The reference's own method,
    FGMRES(30) with the line preconditioner standing in for PCMG(1 level), is timed once beside it.
    Returns the baseline object and the oracle's FP64 solution."""
    import oracle

    A, S = oracle.assemble(g)
    b = oracle.rhs(g, divg)
    p, st = oracle.solve(g, A, S, b, precision=0)  # warm-up: first touch of every work array
    times = []
    for _ in range(reps):
        p, st = oracle.solve(g, A, S, b, precision=0)
        times.append(st.solve_ms)
    xf, sf = oracle.fgmres(g, A, S, b, restart=30, rtol=1e-10)
    med = statistics.median(times)
    model, ncpu = host_cpu()
    base = {"value": round(1e3 / med, 6), "unit": "solves/s", "cores": oracle.num_threads(), "kind": "port",
            "host_cpu": model, "host_cpus_available": ncpu,
            "sample": f"median of {reps} warm FP64 PCG + line-preconditioner solves of the full {g.nx}x{g.ny}x{g.nz} "
                      f"tile by the C oracle (OpenMP), set-up excluded: {', '.join(f'{t / 1e3:.2f}' for t in times)} s, "
                      f"{st.inner_iters} iterations to rel.res {st.rel_res:.1e}",
            "fgmres30": {"value": round(1e3 / sf.solve_ms, 6), "unit": "solves/s", "iterations": sf.inner_iters,
                         "rel_res": sf.rel_res, "seconds": round(sf.solve_ms / 1e3, 3),
                         "note": "the reference's KSP type, FGMRES(30), with the line preconditioner in place of "
                                 "PCMG(1 level); one warm solve, set-up excluded"}}
    return base, p, A, b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="regional", choices=["regional", "batch", "transect", "serve", "troneu",
                                                                     "bottom"])
    ap.add_argument("--requests", type=int, default=48, help="serve workload: arrivals of the burst phase")
    ap.add_argument("--requests2", type=int, default=40, help="serve workload: arrivals of the sustained phase")
    ap.add_argument("--pool", type=int, default=7, help="serve workload: grid sizes pre-assembled")
    ap.add_argument("--rate", type=float, default=200.0, help="serve workload: Poisson arrival rate (1/s)")
    ap.add_argument("--lanes", type=int, default=4, help="serve workload: server lanes")
    ap.add_argument("--cohort-graphs", action="store_true",
                    help="serve workload: cohorts replay captured hipGraphs instead of launching directly")
    ap.add_argument("--nsys", type=int, default=64, help="batch workload: number of systems")
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"],
                    help="batch workload: the precisions run (both by default; one for a short profile)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink nx, ny (debug only)")
    ap.add_argument("--tile-rows", type=int, default=0, help="inner-kernel tile height: 0 = tuned, or 4, 8, 12")
    ap.add_argument("--zw-table", type=int, default=0, choices=[-1, 0, 1],
                    help="where the iteration kernels read ZW/ZP: 0 = by size, 1 = table, -1 = per-cell planes")
    ap.add_argument("--buffer-addressing", type=int, default=0, choices=[-1, 0],
                    help="matvec memory addressing: 0 = buffer descriptors where the arrays fit, -1 = 64-bit pointers")
    ap.add_argument("--sor-depth", type=int, default=0,
                    help="troneu workload: half-sweeps per launch, 0 = tuned, or 4, 6, 7, 8, 12")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (independent systems "
              "would shard whole across GPUs with no exchange, one large system by x-slabs with a one-cell "
              "halo face per matvec; neither is built).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "transect":
        return bench_transect(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "troneu":
        return bench_troneu(args)
    if args.workload == "bottom":
        return bench_bottom(args)
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
        # timed solves: the production path (inner iterations replayed from the captured hipGraph)
        s = NHSolver(g, options(precision=prec, warm_start=False, tile_rows=args.tile_rows, zw_table=args.zw_table,
                                buffer_addressing=args.buffer_addressing))
        s.set_rhs(divg)
        log(f"[bench] {name}: device bytes {s.device_bytes() / 1e9:.2f} GB")
        wall, lat, stats, _ = time_solves(s, args.steps, args.warmup, profiled=False)
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        s.get_solution(pnh)
        sols[name] = g.interior(pnh)
        # kernel timing: the same solves with HIP events around every inner kernel (direct launches)
        s.set_profile(True)
        _, _, pstats, prof = time_solves(s, 2, 1)
        s.close()
        results[name] = dict(wall=wall, lat=lat, stats=stats, prof=prof)
        log(f"[bench] {name}: {args.steps / wall:.3f} solves/s, last {stats[-1]}, prof {prof}")

    mx, f64 = results["mixed"], results["fp64"]
    K = args.steps
    sps = K / mx["wall"]
    sps64 = K / f64["wall"]
    inner = statistics.mean(st["inner_iters"] for st in mx["stats"])
    inner64 = statistics.mean(st["inner_iters"] for st in f64["stats"])
    cells = g.cells
    columns = spec.nx * spec.ny
    prof = mx["prof"]
    mv_avg_ms = prof["matvec_ms"] / max(prof["matvec_launches"], 1)
    up_avg_ms = prof["update_ms"] / max(prof["update_launches"], 1)
    # the first iteration of each inner solve (one per refinement step) has no pending x update
    first_frac = mx["stats"][-1]["outer_iters"] / max(mx["stats"][-1]["inner_iters"], 1)
    table = mx["stats"][-1].get("zw_table", 0) == 1
    kern_bytes = int(bwd_bytes(cells, columns, 4, table) - first_frac * 8 * cells)
    mv_bytes = MATVEC_BYTES[4] * cells
    achieved = mv_bytes / (mv_avg_ms * 1e-3) / 1e9
    prof64 = f64["prof"]
    mv64_ms = prof64["matvec_ms"] / max(prof64["matvec_launches"], 1)
    up64_ms = prof64["update_ms"] / max(prof64["update_launches"], 1)
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
        "dtype": "fp32-inner (fp16 line factors)/fp64-refinement",
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
        "tile_rows_mixed": mx["stats"][-1]["tile_rows"],
        "tile_rows_fp64": f64["stats"][-1]["tile_rows"],
        "rel_res_mixed": mx["stats"][-1]["rel_res"],
        "rel_res_fp64": f64["stats"][-1]["rel_res"],
        "rel_err_mixed_vs_fp64": err,
        "latency_ms_p50": round(pct(mx["lat"], 50), 3),
        "latency_ms_p99": round(pct(mx["lat"], 99), 3),
        "latency_ms_p50_fp64": round(pct(f64["lat"], 50), 3),
        "latency_ms_p99_fp64": round(pct(f64["lat"], 99), 3),
        "roofline": {"bound": "hbm", "kernel": "k_bwd<float> (7-point matvec + back substitution + p update + p.q)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (measured_traffic("traffic_k_bwd_f32_regional_1024x768x60")
                                 if args.buffer_addressing == 0 else None),
                     "bytes_per_launch": mv_bytes, "bytes_per_cell": MATVEC_BYTES[4],
                     "bytes_note": "the bare matvec y = A x: 7 coefficients + x + y per cell (FP32)",
                     "kernel_bytes_per_launch": kern_bytes, "kernel_GBps": round(kern_bytes / (mv_avg_ms * 1e-3) / 1e9, 1),
                     "zw_table": int(table), "buffer_addressing": mx["stats"][-1].get("buffer_addressing", 0),
                     "avg_launch_ms": round(mv_avg_ms, 5), "launches": prof["matvec_launches"]},
        "other_kernels": {
            "k_fwd<float>": {"avg_launch_ms": round(up_avg_ms, 5), "bytes_per_cell": 20, "bytes_per_column": 4,
                             "GBps": round(fwd_bytes(cells, columns, 4) / (up_avg_ms * 1e-3) / 1e9, 1)},
            "k_bwd<double>": {"avg_launch_ms": round(mv64_ms, 5), "matvec_bytes_per_cell": MATVEC_BYTES[8],
                              "matvec_GBps": round(MATVEC_BYTES[8] * cells / (mv64_ms * 1e-3) / 1e9, 1),
                              "kernel_GBps": round(bwd_bytes(cells, columns, 8, f64["stats"][-1].get("zw_table", 0) == 1)
                                                   / (mv64_ms * 1e-3) / 1e9, 1)},
            "k_fwd<double>": {"avg_launch_ms": round(up64_ms, 5), "bytes_per_cell": 48,
                              "GBps": round(fwd_bytes(cells, columns, 8) / (up64_ms * 1e-3) / 1e9, 1)},
        },
        "copy_probe_GBps": round(abi.copy_probe(1 << 30, 20), 1),
        "device": device_state(),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle FP64 solves on the host cores) ...")
        import oracle

        base, p_oracle, A, b = cpu_baseline(g, divg)
        line["cpu_baseline"] = base
        # the device solutions against the oracle: its FP64 solution and its FP64 operator's residual
        nref = float(np.linalg.norm(p_oracle))
        line["vs_oracle"] = {
            "rel_err_mixed": float(np.linalg.norm(sols["mixed"] - p_oracle) / nref),
            "rel_err_fp64": float(np.linalg.norm(sols["fp64"] - p_oracle) / nref),
            "oracle_rel_res_mixed": oracle.rel_res(g, A, sols["mixed"], b),
            "oracle_rel_res_fp64": oracle.rel_res(g, A, sols["fp64"], b),
            "note": "||p - p_oracle|| / ||p_oracle|| with p_oracle the oracle's FP64 PCG solution to rel.res 1e-10; "
                    "rel.res of the device solutions by the oracle's FP64 operator"}
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """64 independent systems of uneven sizes (nx, ny ~ U{96..512}, nz ~ U{30..60}) solved together,
    one launch per solver phase for the whole batch, per-system convergence masks, hipGraph chunks."""
    from orctm_amd.nh import NHBatch

    specs = synth.config4_sizes(n=args.nsys)
    if args.scale != 1.0:
        for sp in specs:
            sp.nx = max(8, int(sp.nx * args.scale))
            sp.ny = max(8, int(sp.ny * args.scale))
    t = time.perf_counter()
    systems = []
    for sp in specs:
        g, divg = synth.build(sp)
        systems.append((g, divg))
    cells = sum(g.cells for g, _ in systems)
    log(f"[bench] batch: {len(systems)} systems, {cells} cells, generated in {time.perf_counter() - t:.1f} s")
    out = {}
    precs = (("mixed", abi.MIXED), ("fp64", abi.FP64))
    if args.precision != "both":
        precs = tuple(x for x in precs if x[0] == args.precision)
    for name, prec in precs:
        solvers = []
        for g, divg in systems:
            s = NHSolver(g, options(precision=prec, warm_start=False, tile_rows=args.tile_rows, zw_table=args.zw_table,
                                    buffer_addressing=args.buffer_addressing))
            s.set_rhs(divg)
            solvers.append(s)
        batch = NHBatch(solvers)
        for _ in range(args.warmup):
            rc, st = batch.solve()
            if rc != abi.OK:
                raise RuntimeError(f"batch warmup failed: {rc} {abi.last_error()}")
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a = time.perf_counter()
            rc, st = batch.solve()
            lat.append((time.perf_counter() - a) * 1e3)
            if rc != abi.OK:
                raise RuntimeError(f"batch solve failed: {rc}")
        wall = time.perf_counter() - t0
        batch.close()
        # kernel timing: the same batch with HIP events around every inner kernel (direct launches)
        for s in solvers:
            s.set_profile(True)
        batch = NHBatch(solvers)
        batch.solve()
        batch.profile(reset=True)
        rc, pst = batch.solve()
        prof = batch.profile(reset=True)
        for s in solvers:
            s.set_profile(False)
        inner = sum(x["inner_iters"] for x in st)
        # cell-iterations actually processed per solve (converged systems drop out of later launches)
        cell_iters = sum(x["inner_iters"] * g.cells for x, (g, _) in zip(st, systems))
        mv_bytes = sum(x["inner_iters"] * MATVEC_BYTES[4] * g.cells for x, (g, _) in zip(st, systems))
        # one-by-one reference timing of the same systems (sequential, same kernels)
        t1 = time.perf_counter()
        for s in solvers:
            r = s.solve_resident()
            if r["rc"] != abi.OK:
                raise RuntimeError(f"single solve failed: {r}")
        seq = time.perf_counter() - t1
        heights = {}
        for x in st:
            heights[x["tile_rows"]] = heights.get(x["tile_rows"], 0) + 1
        out[name] = dict(wall=wall, lat=lat, inner=inner, cell_iters=cell_iters, mv_bytes=mv_bytes, heights=heights,
                         max_inner=max(x["inner_iters"] for x in st),
                         rel=max(x["rel_res"] for x in st), prof=prof, seq=seq)
        log(f"[bench] batch {name}: {len(systems) * args.steps / wall:.2f} solves/s, one-by-one "
            f"{len(systems) / seq:.2f} solves/s, sum inner {inner}, max rel {out[name]['rel']:.2e}")
        batch.close()
        for s in solvers:
            s.close()
    mx, f64 = out.get("mixed"), out.get("fp64")
    if mx is None:  # FP64 alone: report it in the mixed fields' place
        mx, f64 = f64, None
    n = len(systems)
    prof = mx["prof"]
    # bytes moved by the matvec launches of one profiled batch solve / their summed device time
    mv_gbs = mx["mv_bytes"] / (prof["matvec_ms"] * 1e-3) / 1e9
    # the counted traffic of the 64-system mixed batch (rocprofv3 summary in profiles/)
    traffic = None
    if "mixed" in out and args.nsys == 64 and args.scale == 1.0 and args.tile_rows in (0, 12) and args.zw_table == 0 \
            and args.buffer_addressing == 0:
        traffic = measured_traffic("traffic_k_bwd_f32_batch_64systems")
    line = {
        "metric": METRIC, "value": round(n * args.steps / mx["wall"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * mx["wall"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-inner (fp16 line factors)/fp64-refinement",
        "data": "synthetic (seeded generator, regional-tile recipe per system)",
        "config": {"workload": f"batch of {n} systems, nx,ny~U{{96..512}}, nz~U{{30..60}}, mixed, rel.res 1e-10",
                   "cells": cells},
        "precisions_run": [name for name, _ in precs],
        "fp64_solves_per_s": round(n * args.steps / f64["wall"], 3) if f64 else None,
        "mixed_over_fp64": round(f64["wall"] / mx["wall"], 3) if f64 else None,
        "one_by_one_solves_per_s_mixed": round(n / mx["seq"], 3),
        "cell_iters_per_s": round(mx["cell_iters"] * args.steps / mx["wall"], 1),
        "max_rel_res_mixed": mx["rel"], "max_rel_res_fp64": f64["rel"] if f64 else None,
        "tile_rows_mixed": mx["heights"], "tile_rows_fp64": f64["heights"] if f64 else None,
        "latency_ms_p50": round(pct(mx["lat"], 50), 3), "latency_ms_p99": round(pct(mx["lat"], 99), 3),
        "device": device_state(),
        "roofline": {"bound": "hbm", "kernel": "k_bwd<float> (batched)", "achieved": round(mv_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(mv_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_launch": MATVEC_BYTES[4] * cells,
                     "note": "the bare matvec bytes (36 B per cell, FP32) per active cell-iteration over the "
                             "summed matvec launch time; bytes_per_launch and traffic (HBM bytes counted by "
                             "FETCH_SIZE/WRITE_SIZE) are those of one launch in which every system iterates"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_transect(args):
    """The ISW transect 512 x 4 (periodic y) x 100: the reference's CPU method (FGMRES(30), here the
    oracle's restatement with the line preconditioner, FP64) timed on the host beside the GPU solves."""
    import oracle

    spec = synth.config1()
    g, divg = synth.build(spec)
    A, S = oracle.assemble(g)
    b = oracle.rhs(g, divg)
    t = time.perf_counter()
    reps = 0
    while True:
        x, st = oracle.fgmres(g, A, S, b, restart=30, rtol=1e-10)
        reps += 1
        if time.perf_counter() - t > 5.0 or reps >= 20:
            break
    cpu_s = (time.perf_counter() - t) / reps
    res = {}
    for name, prec in (("mixed", abi.MIXED), ("fp64", abi.FP64)):
        s = NHSolver(g, options(precision=prec, warm_start=False, zw_table=args.zw_table,
                                buffer_addressing=args.buffer_addressing))
        s.set_rhs(divg)
        wall, lat, stats, _ = time_solves(s, args.steps, args.warmup, profiled=False)
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        s.get_solution(pnh)
        # kernel timing (HIP events around every inner kernel, direct launches): this workload's kernel
        # rate is what the small-launch kernel forms are tuned for
        s.set_profile(True)
        _, _, _, prof = time_solves(s, 2, 1)
        res[name] = dict(wall=wall, lat=lat, stats=stats, prof=prof,
                         err=float(np.linalg.norm(g.interior(pnh) - x) / np.linalg.norm(x)),
                         rel=oracle.rel_res(g, A, g.interior(pnh), b))
        s.close()
    mx = res["mixed"]
    line = {
        "metric": METRIC, "value": round(args.steps / mx["wall"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * mx["wall"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-inner (fp16 line factors)/fp64-refinement",
        "data": "synthetic (ISW transect recipe)",
        "config": {"workload": "ISW transect 512x4x100 periodic y, Gaussian ridge", "cells": g.cells},
        "fp64_solves_per_s": round(args.steps / res["fp64"]["wall"], 3),
        "inner_iters_mixed": mx["stats"][-1]["inner_iters"], "inner_iters_fp64": res["fp64"]["stats"][-1]["inner_iters"],
        "oracle_rel_res_mixed": mx["rel"], "oracle_rel_res_fp64": res["fp64"]["rel"],
        "zw_table": mx["stats"][-1]["zw_table"],
        "buffer_addressing": mx["stats"][-1].get("buffer_addressing", 0),
        "kernel_us": {k: {"matvec": round(1e3 * v["prof"]["matvec_ms"] / max(v["prof"]["matvec_launches"], 1), 2),
                          "update": round(1e3 * v["prof"]["update_ms"] / max(v["prof"]["update_launches"], 1), 2)}
                      for k, v in res.items()},
        "rel_err_vs_cpu_fgmres": {k: v["err"] for k, v in res.items()},
        "latency_ms_p50": round(pct(mx["lat"], 50), 3), "latency_ms_p99": round(pct(mx["lat"], 99), 3),
        "cpu_baseline": {"value": round(1.0 / cpu_s, 4), "unit": "solves/s", "cores": oracle.num_threads(),
                         "kind": "port", "sample": f"FGMRES(30) + line preconditioner, FP64, rel.res 1e-10, "
                                                   f"{st.inner_iters} iterations, {reps} solves"},
    }
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Serving: Poisson arrivals, sizes log-uniform from 128x128x30 to 2048x1536x80, 10 time steps per
    request (RHS perturbed by 1 % N(0,1) per step, warm start from the resident pressure), mixed precision
    to rel.res 1e-10, admission by free HBM into cohorts of one size class that run concurrently on the
    server's lanes. Sizes snap to a pool of `--pool` pre-assembled grids at t = 0, 1/(pool-1), ..., 1
    (both ends of the range included; a request clones its grid's model, the operator is shared).
    Two phases: the first `--requests` arrivals at `--rate` (200/s, far above capacity: a burst whose
    latency is queueing), then `--requests2` arrivals at 60 % of the capacity the burst measured (a load
    the GPU sustains). `value` is the burst's solves/s."""
    from orctm_amd.nh import NHServer

    P = max(args.pool, 2)
    ts = [q / (P - 1) for q in range(P)]
    models = []
    t = time.perf_counter()
    for q, tq in enumerate(ts):
        nx, ny, nz = synth.config5_size(tq)
        if args.scale != 1.0:
            nx, ny = max(8, int(nx * args.scale)), max(8, int(ny * args.scale))
        g, divg = synth.build(synth.GridSpec(nx, ny, nz, seed=synth.BASE_SEED + 500 + q))
        m = NHSolver(g, options(precision=abi.MIXED, warm_start=True, zw_table=args.zw_table,
                                buffer_addressing=args.buffer_addressing))
        m.set_rhs(divg)
        for a in ("dlxp", "dlyp", "dlxu", "dlyv", "ddpo", "weto"):
            setattr(g, a, None)
        del divg
        models.append(m)
        log(f"[bench] serve: model {q}: {nx}x{ny}x{nz} = {nx * ny * nz} cells")
    log(f"[bench] serve: {len(models)} models assembled in {time.perf_counter() - t:.1f} s")
    nsteps = 10

    def phase(srv, n, rate, stream):
        u = synth.uniform(synth.BASE_SEED, stream, 2 * n)
        arrivals = np.cumsum(-np.log(1.0 - u[:n]) / rate)
        pick = [int(round(x * (P - 1))) for x in u[n:]]
        tickets = []
        t0 = time.perf_counter()
        for k in range(n):
            delay = arrivals[k] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(models[pick[k]], nsteps, 0.01, 9000 + 100 * k + 7 * stream))
        stats = [srv.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        bad = [s for s in stats if s["status"] != "OK"]
        if bad:
            raise RuntimeError(f"serve: {len(bad)} requests failed: {bad[:3]}")
        lat = [s["t_done_ms"] - s["t_submit_ms"] for s in stats]
        svc = [(s["t_done_ms"] - s["t_admit_ms"]) / nsteps for s in stats]
        by_size = {}
        for p_, l_, v_, s_ in zip(pick, lat, svc, stats):
            by_size.setdefault(p_, ([], [], []))
            by_size[p_][0].append(l_)
            by_size[p_][1].append(v_)
            by_size[p_][2].append(s_["inner_iters"] / nsteps)
        classes = {}
        for p_ in sorted(by_size):
            l_, v_, it_ = by_size[p_]
            classes[f"{models[p_].grid.nx}x{models[p_].grid.ny}x{models[p_].grid.nz}"] = {
                "requests": len(l_), "latency_ms_p50": round(pct(l_, 50), 1), "latency_ms_p99": round(pct(l_, 99), 1),
                "service_ms_per_step_p50": round(pct(v_, 50), 2), "service_ms_per_step_p99": round(pct(v_, 99), 2),
                "inner_iters_per_step": round(float(np.mean(it_)), 1)}
        cells = sum(models[p_].grid.cells for p_ in pick)
        cell_iters = sum(models[p_].grid.cells * s["inner_iters"] for p_, s in zip(pick, stats))
        return dict(n=n, rate=rate, wall=wall, lat=lat, svc=svc, classes=classes, cells=cells, cell_iters=cell_iters,
                    inner=sum(s["inner_iters"] for s in stats), rel=max(s["rel_res"] for s in stats))

    with NHServer(lanes=args.lanes, cohort_graphs=args.cohort_graphs) as srv:
        burst = phase(srv, args.requests, args.rate, 505)
        log(f"[bench] serve burst: {args.requests} requests in {burst['wall']:.2f} s")
        capacity = args.requests / burst["wall"]
        steady = phase(srv, args.requests2, 0.6 * capacity, 506) if args.requests2 > 0 else None
        info = srv.info()
    solves = nsteps * args.requests
    # the inner matvec's bytes (36 B per cell, FP32) over every inner iteration of the burst, per second of
    # its wall time: the whole service (admission, cohort set-up, the other kernels, idle lanes) counted
    mv_gbs = 36.0 * burst["cell_iters"] / burst["wall"] / 1e9

    def summary(ph):
        return {"requests": ph["n"], "offered_per_s": round(ph["rate"], 3), "wall_s": round(ph["wall"], 3),
                "solves_per_s": round(nsteps * ph["n"] / ph["wall"], 3),
                "cell_solves_per_s": round(nsteps * ph["cells"] / ph["wall"], 1),
                "latency_ms_p50": round(pct(ph["lat"], 50), 1), "latency_ms_p99": round(pct(ph["lat"], 99), 1),
                "service_ms_per_step_p50": round(pct(ph["svc"], 50), 2),
                "service_ms_per_step_p99": round(pct(ph["svc"], 99), 2), "by_size": ph["classes"]}

    line = {
        "metric": METRIC, "value": round(solves / burst["wall"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": solves, "warmup": 0, "ms_per_step": round(1e3 * burst["wall"] / solves, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32-inner (fp16 line factors)/fp64-refinement",
        "data": "synthetic (seeded generator, regional-tile recipe per grid)",
        "config": {"workload": f"serving: Poisson arrivals, first {args.requests} requests at {args.rate:g}/s, then "
                               f"{args.requests2} at 60% of measured capacity; {nsteps} warm-started steps each, "
                               f"sizes log-uniform 128x128x30..2048x1536x80 on a {P}-grid pool, mixed to rel.res "
                               f"1e-10, {args.lanes} lanes of size-class cohorts"
                               + (", cohorts from captured hipGraphs" if args.cohort_graphs else ""),
                   "pool_cells": [m.grid.cells for m in models]},
        "requests": args.requests, "rounds": info["rounds"],
        "latency_ms_p50": round(pct(burst["lat"], 50), 1), "latency_ms_p99": round(pct(burst["lat"], 99), 1),
        "burst": summary(burst), "sustained": summary(steady) if steady else None,
        "cell_solves_per_s": round(nsteps * burst["cells"] / burst["wall"], 1),
        "mean_inner_iters_per_step": round(burst["inner"] / solves, 2),
        "max_rel_res": max(burst["rel"], steady["rel"] if steady else 0.0),
        "roofline": {"bound": "hbm", "kernel": "k_bwd<float> over the whole burst", "achieved": round(mv_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(mv_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "the inner matvec's bytes (36 B per cell, FP32) of every inner iteration served, over "
                             "the burst's wall time (admission, cohort set-up, the other kernels and idle lanes "
                             "included): a lower bound of the service's streaming rate, not one kernel's"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    for m in models:
        m.close()


def bench_troneu(args):
    """The barotropic free-surface solve of the regional tile's surface (1026 x 770 with halos):
    B1O + 250 red-black SOR sweeps (troneu.f90:48-131) per call. The entry point takes the caller's host
    arrays, so each call uploads the transports and B1O and downloads Z1O and B1O; `value` is the
    device-resident rate (HIP events from after the uploads to after the last sweep, summed over the K
    timed calls), and the PCIe-inclusive rate of the same calls is reported beside it. The CPU leg is the
    oracle's restatement (C, one thread) on the same inputs."""
    import oracle
    from orctm_amd.nh import NHBarotropic

    spec = synth.config2()
    if args.scale != 1.0:
        spec.nx = max(8, int(spec.nx * args.scale))
        spec.ny = max(8, int(spec.ny * args.scale))
    g, _ = synth.build(spec)
    b = synth.make_barotropic(g)
    b1o = np.zeros((b.IE, b.JE), order="F")
    z1o = np.zeros((b.IE, b.JE), order="F")
    with NHBarotropic(b, half_sweeps=args.sor_depth) as s:
        info = s.info()
        for _ in range(args.warmup):
            s.troneu(b.u1o, b.uzo, b.v1e, b.vze, b1o, z1o)
        dev, lat = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a = time.perf_counter()
            dev.append(s.troneu(b.u1o, b.uzo, b.v1e, b.vze, b1o, z1o))
            lat.append((time.perf_counter() - a) * 1e3)
        wall = time.perf_counter() - t0
    cells = (b.IE - 2) * (b.JE - 2)
    # per blocked launch and interior cell: Z, B1O, FF, UF, VF read once and Z written once = 6 words
    # (the region halos and the second UF/VF neighbour come from cache); the few small kernels around
    # the blocked launches are counted in the device time
    dev_ms = statistics.median(dev)
    achieved = 48.0 * cells * (info["launches"] - 1) / (dev_ms * 1e-3) / 1e9
    coef = oracle.itprep(b)
    t = time.perf_counter()
    z_ref, _ = oracle.troneu(b, coef)
    cpu_s = time.perf_counter() - t
    err = float(np.abs(z1o - z_ref).max())
    line = {
        "metric": "barotropic SOR solves/s (TRONEU: B1O + 250 red-black sweeps)",
        "value": round(args.steps / (sum(dev) * 1e-3), 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sum(dev) / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "pcie_inclusive_solves_per_s": round(args.steps / wall, 3),
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (regional-tile surface, seeded transports)",
        "config": {"workload": f"TRONEU on the {b.IE}x{b.JE} surface of the regional tile, SORPAR {b.sorpar}, "
                               f"{b.iterations} sweeps", "cells": cells},
        "device_ms_p50": round(dev_ms, 3), "latency_ms_p50": round(pct(lat, 50), 3),
        "latency_ms_p99": round(pct(lat, 99), 3), "max_abs_diff_vs_oracle": err,
        "half_sweeps_per_launch": info["half_sweeps"], "launches": info["launches"],
        "roofline": {"bound": "hbm", "kernel": "k_sor_tb (all launches) + k_b1o", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "48 B per interior cell and blocked launch (launches after the first half-sweep); the "
                             "working arrays (50 MB) stay in the Infinity Cache"},
        "cpu_baseline": {"value": round(1.0 / cpu_s, 4), "unit": "solves/s", "cores": 1, "kind": "port",
                         "sample": f"one TRONEU of the same surface by the C oracle, one thread, {cpu_s:.2f} s"},
    }
    print(json.dumps(line), flush=True)


def mg_iteration_bytes(nx, ny):
    """Bytes one iteration of the 2-D PCG streams (8-byte words). On the box: k2_dirmv reads z, d, the 5
    coefficients and S and writes d, q (10 words); k2_update reads x, d, r, q and writes x, r (6 words).
    Per cell of a multigrid level above the one-block tail: the way down (k_mg_down, one launch) reads r
    and K (6 words) and writes the smoothed iterate and a quarter of a coarse right-hand side (1.25); the
    way up reads the iterate, a quarter of a coarse value, r and K and writes the result (8.25, k_mg_up);
    on the box k_mg_dotm then reads r and z for the r.z block sums (2). The tail (the
    coarsest levels of at most 1024 cells each, at most 4 of them and 2048 cells together, never the box)
    reads K (5 words per cell) and its first level's right-hand side, and writes that level's result.
    Levels: the box, then 2 x 2 aggregates down to at most 64 unknowns."""
    a, b = nx - 2, ny - 2
    levels = [nx * ny]
    while a * b > 64:
        a, b = (a + 1) // 2, (b + 1) // 2
        levels.append(a * b)
    L = len(levels) - 1
    T, tail = L + 1, 0
    while T - 1 >= 1 and levels[T - 1] <= 1024 and L - T + 2 <= 4 and tail + levels[T - 1] <= 2048:
        tail += levels[T - 1]
        T -= 1
    sep = min(T, L)  # levels with launches of their own
    words = 16 * levels[0]
    if sep >= 1:
        words += 17.5 * levels[0] + sum(15.5 * levels[l] for l in range(1, sep))
    if T <= L:
        words += 5 * tail + 2 * levels[T]
    else:
        words += 2 * levels[L]  # the dense solve of the last level: its right-hand side in, its result out
    return int(8 * words)


def bench_bottom(args):
    """The 2-D bottom-pressure solve of ORCTM's CG2D4NONHY option (ocvtot_nh.f90:64-147, mo_petsc.f90:147-208,
    252-374) on the regional tile (1024 x 768 x 60; 1024 x 768 unknowns): per step the divergence, the
    bottom values, PNH_bot from the 2-D solve and the corrections of DIVG (orctm_nh_divergence). Steps
    alternate between two seeded velocity states 1 % apart, so each 2-D solve is warm-started from the
    other state's solution, as a time step would be; the first solve (from zero) is reported beside.
    `value` is 2-D solves/s on the device: HIP events from the bottom values to the exchange of PNH_bot,
    summed over the K timed steps (the velocity upload and the 3-D divergence are outside). The CPU leg is
    the oracle's restatement of the same V-cycle PCG (C, one thread) on a bounded number of iterations."""
    import oracle

    spec = synth.config2()
    if args.scale != 1.0:
        spec.nx = max(8, int(spec.nx * args.scale))
        spec.ny = max(8, int(spec.ny * args.scale))
    g, _ = synth.build(spec)
    v1 = synth.make_velocity(g)
    v2 = synth.make_velocity(g)
    rng = np.random.default_rng(20261202)
    for n in ("uoo", "voe", "wno"):
        a = getattr(v1, n)
        setattr(v2, n, np.asfortranarray(a * (1.0 + 0.01 * rng.standard_normal(a.shape))))
    log(f"[bench] bottom: {g.nx}x{g.ny}x{g.nz}")

    def step(s, v):
        s.divergence(v.uoo, v.uoo_1, v.voe, v.voe_1, v.wno, v.wno_1, v.zo)
        return s.get_bottom()

    with NHSolver(g, options(precision=abi.MIXED, bottom_solve=True)) as s:
        s.attach_vgrid(v1)
        first = step(s, v1)
        log(f"[bench] bottom: first solve {first['iterations']} iterations, {first['device_ms']:.1f} ms")
        for w in range(args.warmup):
            step(s, v2 if w % 2 == 0 else v1)
        outs = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            outs.append(step(s, v1 if (args.warmup + k) % 2 == 0 else v2))
        wall = time.perf_counter() - t0
    dev = [o["device_ms"] for o in outs]
    iters = [o["iterations"] for o in outs]
    last = outs[-1]
    n = g.nx * g.ny
    achieved = mg_iteration_bytes(g.nx, g.ny) * sum(iters) / (sum(dev) * 1e-3) / 1e9
    # the oracle's residual of the device's last PNH_bot
    vlast = v1 if (args.warmup + args.steps - 1) % 2 == 0 else v2
    c = oracle.op2d_coef(g)
    b = last["divg_bot_surf"][1:-1, 1:-1].ravel(order="F")
    x = last["pnh_bot"][1:-1, 1:-1].ravel(order="F")
    rr_oracle = float(np.linalg.norm(b - oracle.op2d_apply(g, c, x)) / np.linalg.norm(b))
    cpu_it = 512
    t = time.perf_counter()
    oracle.solve2d(g, c, b, rtol=1e-30, maxit=cpu_it)
    cpu_s = time.perf_counter() - t
    per_solve = statistics.mean(iters)
    line = {
        "metric": "2-D bottom-pressure solves/s (CG2D4NONHY: bottom values, PNH_bot to rel.res 1e-10, DIVG corrections)",
        "value": round(args.steps / (sum(dev) * 1e-3), 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sum(dev) / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (regional tile, two seeded velocity states 1 % apart)",
        "config": {"workload": f"2-D bottom solve on the {g.nx}x{g.ny} interior box of the regional tile "
                               f"{g.nx}x{g.ny}x{g.nz}, warm-started", "unknowns": n},
        "iterations_per_solve": per_solve, "rel_res": last["rel_res"], "rel_res_oracle_operator": rr_oracle,
        "first_solve": {"iterations": first["iterations"], "device_ms": round(first["device_ms"], 3)},
        "step_wall_ms_p50": round(1e3 * wall / args.steps, 3),
        "roofline": {"bound": "hbm", "kernel": "the 2-D PCG with its V-cycle (all kernels of a solve)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_iteration": mg_iteration_bytes(g.nx, g.ny),
                     "note": "what the kernels stream: the PCG's 128 B per box cell, the fused level kernels' 124 B "
                             "per level cell (140 on the box), the one-block tail's operator; about 16 launches per "
                             "iteration (one per level and direction down to 64 x 48 cells, one block below that) "
                             "make the solve launch-bound, "
                             "and its working set (about 130 MB) stays in the Infinity Cache"},
        "cpu_baseline": {"value": round((cpu_it / cpu_s) / per_solve, 4), "unit": "solves/s", "cores": 1,
                         "kind": "port",
                         "sample": f"{cpu_it} iterations of the same PCG by the C oracle on the same box, one "
                                   f"thread, {cpu_s:.2f} s, scaled by the {per_solve:.0f} iterations per solve"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
