#!/usr/bin/env python3
"""Benchmark: time-to-1e-10 Darcy head solve at 256^3 on one MI355X.

One step = one complete head solve to ||b - A h|| <= 1e-10 ||b|| (face
coefficients of every level, RHS, MG-PCG) plus the face velocities, on a seeded
log-normal K field (sigma2_lnK = 4, correlation length 8 cells) resident in HBM.

Headline method: FP64 PCG with the FP32 V-cycle and BF16 face coefficients
(the dots, the operator, the coarsest 8^3 solve and the true-residual check
stay FP64). The plain FP64 MG-PCG is timed on the same field as the
denominator, and the mixed head/velocity are compared with it.

Also measured here:
  * roofline of the dominant kernel (the finest level's GSRB colour sweep of
    the headline V-cycle): algorithmic bytes per launch / average launch time
    from HIP events on the solve stream; HBM traffic per launch from the
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries committed under profiles/;
  * CPU baseline: the CPU FP64 oracle (OpenMP) solving the same 256^3 K to
    1e-10 (set-up + solve, median of 3); the GPU FP64 head and velocity
    checksums against that same-algorithm solve, and both GPU heads against
    one more oracle solve to 1e-13;
  * 512^3 (sigma2_lnK = 4, correlation length 16, 7 levels): plain FP64 and
    mixed-precision solve times, head/velocity error of mixed vs FP64;
  * the uneven Monte-Carlo batch (32 x 48^3, 32 x 64^3, 32 x 96^3): solves/s
    of the batched solve and of one-after-another solves.

Prints one JSON line. `--summarize-pmc DIR COUNTER` instead prints the
per-dispatch counter totals of a rocprofv3 --pmc run of this script.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "time-to-1e-10 Darcy solve at 256^3, mixed+warm-start vs FP64; smoother HBM GB/s"
# HBM traffic of the dominant kernel, from the committed FETCH_SIZE/WRITE_SIZE summary
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_hbm_traffic_headsolve_256.md")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--sigma2", type=float, default=4.0)
    ap.add_argument("--corr", type=float, default=8.0)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--coef", choices=["bf16", "fp16", "fp32"], default="bf16")
    ap.add_argument("--smoother-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-reps", type=int, default=3, help="timed oracle solves (median)")
    ap.add_argument("--no-512", action="store_true", help="skip the 512^3 solves")
    ap.add_argument("--no-batch", action="store_true", help="skip the Monte-Carlo batch")
    ap.add_argument("--batch-per-size", type=int, default=32)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("DIR", "COUNTER"))
    ap.add_argument("--summarize-trace", metavar="DIR")
    ap.add_argument("--sweep-only", choices=["fp64", "mixed"],
                    help="set up one hierarchy and time only its finest-level half-sweeps "
                         "(--smoother-reps red-black iterations; for kernel counters)")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="headline from the linear head instead of the learned warm start")
    ap.add_argument("--weights", default=None, help="warm-start parameters (.npy)")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def summarize_pmc(d: str, counter: str) -> int:
    """Per-kernel, per-dispatch counter value (KiB) of a rocprofv3 --pmc run."""
    import glob
    import sqlite3
    dbs = glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True)
    if not dbs:
        print(f"no rocprofv3 results under {d}")
        return 1
    cur = sqlite3.connect(dbs[0]).cursor()
    rows = cur.execute(
        "select kernel_name, grid_size, count(distinct dispatch_id), sum(value) "
        "from counters_collection where counter_name = ? group by kernel_name, grid_size "
        "order by sum(value) desc", (counter,)).fetchall()
    for name, grid, nd, tot in rows:
        if name.startswith("mf3d::") or name.startswith("void mf3d::"):
            print(f"{counter}\t{tot / max(nd, 1):.1f} KiB/dispatch\tdispatches={nd}\t"
                  f"grid={grid}\t{name[:110]}")
    return 0


def summarize_trace(d: str, tail: int = 0) -> int:
    """Per-kernel, per-grid launch count and average duration of a rocprofv3
    --kernel-trace --output-format csv run of this script; with tail > 0 also
    the average of the finest half-sweep's last `tail` launches."""
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        print(f"no kernel_trace.csv under {d}")
        return 1
    acc = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            name = row["Kernel_Name"]
            if not (name.startswith("mf3d::") or name.startswith("void mf3d::")):
                continue
            grid = 1
            for ax in ("X", "Y", "Z"):
                grid *= int(row.get(f"Grid_Size_{ax}", 1) or 1)
            start = int(row["Start_Timestamp"])
            dur = int(row["End_Timestamp"]) - start
            c = acc.setdefault((name, grid), [0, 0, []])
            c[0] += 1
            c[1] += dur
            c[2].append((start, dur))
    for (name, grid), (cnt, tot, _) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"{tot / cnt / 1e3:10.2f} us avg\tcalls={cnt}\tgrid={grid}\ttotal={tot / 1e6:.3f} ms"
              f"\t{name[:110]}")
    # the finest-level half-sweeps' last launches at each grid are the timed
    # smoother repetitions (2 per mg_smooth iteration), the launches bench.py's
    # HIP events cover; the averages above also hold the in-solve launches.
    # Mixed V-cycle: colour_split4_kernel<., 0, false>; FP64: sweep64_kernel<0, ...>
    # (the FP64 march kernel launches the same grid on every level it runs
    # on, so its rows above mix levels; its last launches are the finest
    # level's timed repetitions)
    sweeps = [(k, v) for k, v in acc.items()
              if ("colour_split4_kernel<" in k[0] and ", 0, false>" in k[0]) or
              "march64_kernel<0, false, false>" in k[0]]
    if tail > 0:
        for kind in ("colour_split4_kernel<", "march64_kernel<"):
            own = [kv for kv in sweeps if kind in kv[0][0]]
            # the two largest grids: the finest levels of the 256^3 and 512^3 runs
            for (name, grid), (_, _, launches) in sorted(own, key=lambda kv: -kv[0][1])[:2]:
                last = [d for _, d in sorted(launches)[-tail:]]
                print(f"{sum(last) / len(last) / 1e3:10.2f} us avg\tcalls={len(last)}\tgrid={grid}"
                      f"\tlast {len(last)} launches (timed smoother repetitions)\t{name[:80]}")
    return 0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def checksums(a):
    """(sum, sum of squares, max |.|) in FP64: the velocity checksums the
    reference prints (velocity_from_head.cu:942-962)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    return np.array([a.sum(), float(np.dot(a, a)), float(np.abs(a).max())])


def cpu_baseline(args, K_host, gpu):
    """The oracle's plain FP64 MG-PCG (direct coarsest, ||b - A h|| <= rtol ||b||,
    linear initial head like the GPU's FP64 solve) on the same K: set-up +
    solve timed, median of --cpu-reps runs; then one oracle solve to 1e-13 as
    the accuracy reference of the GPU heads (gpu: name -> (head, (U, V, W)) as
    host arrays)."""
    import numpy as np

    import oracle as O
    threads = O.num_threads()
    g = O.grid(args.n, args.n, args.n)
    b = O.darcy_bc()
    p = O.mode_n(mg_levels=args.levels, cg_rtol=args.rtol, init_guess=0)
    times, its, h10 = [], 0, None
    for _ in range(max(1, args.cpu_reps)):
        t0 = time.perf_counter()
        h10, r = O.solve_head(g, b, p, K_host)
        times.append(time.perf_counter() - t0)
        its = r.num_iterations
        log(args, f"cpu oracle: {its} PCG iterations in {times[-1]:.2f} s")
    v10 = O.velocity_compact(g, b, h10, K_host)
    t0 = time.perf_counter()
    ho, _ = O.solve_head(g, b, O.mode_n(mg_levels=args.levels, cg_rtol=1e-13), K_host)
    vo = O.velocity_compact(g, b, ho, K_host)
    log(args, f"cpu oracle: 1e-13 reference solve + velocities in {time.perf_counter() - t0:.2f} s")

    def devs(vel, ref_uvw):
        # (sum, sum of squares, max |.|) checksums; the sum on the scale of
        # the component's L1 mass (V and W sum to ~0)
        worst = np.zeros(3)
        for comp, ref in zip(vel, ref_uvw):
            cg, co = checksums(comp), checksums(ref)
            scale = np.array([np.abs(ref).sum(), co[1], co[2]])
            worst = np.maximum(worst, np.abs(cg - co) / scale)
        return [float(v) for v in worst]

    def rel(a, r):
        return float(np.linalg.norm(a - r) / np.linalg.norm(r))

    parity = {}
    for name, (h, vel) in gpu.items():
        parity[name] = {"head_rel_l2_vs_oracle_1e-13": rel(h, ho),
                        "velocity_checksum_devs_vs_oracle_1e-13": devs(vel, vo)}
    # the GPU FP64 solve and the oracle's timed solve run the same algorithm
    # (linear head, direct coarsest solve, the same stopping test)
    h64, v64 = gpu["fp64"]
    parity["fp64_same_algorithm"] = {"oracle_iterations": its,
                                     "head_rel_l2_vs_oracle_1e-10": rel(h64, h10),
                                     "velocity_checksum_devs_vs_oracle_1e-10": devs(v64, v10)}
    ref_sums = {c: checksums(a).tolist() for c, a in zip("UVW", vo)}
    med = float(np.median(times))
    return ({"value": round(med * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
             "sample": (f"CPU FP64 oracle (OpenMP, {threads} threads, {cpu_model()}) solving the "
                        f"same {args.n}^3 K to ||b - A h|| <= {args.rtol:g} ||b|| from the linear "
                        f"head (set-up + solve, {its} PCG iterations), median of {len(times)}"),
             "runs_ms": [round(t * 1e3, 1) for t in times], "iterations": its},
            parity, ref_sums)


def committed_traffic(n: int, coef: str):
    """HBM bytes per launch of the dominant kernel from the committed PMC summary:
    the line `traffic_bytes_per_launch n=<n> coef=<coef>: <bytes>`."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                m = re.match(r"\s*traffic_bytes_per_launch n=(\d+) coef=(\w+):\s*([0-9.eE+]+)", line)
                if m and int(m.group(1)) == n and m.group(2) == coef:
                    return float(m.group(3))
    except OSError:
        pass
    return None


def time_smoother(ctx, ws, reps):
    """Average duration (ms) of one finest-level half-sweep of ws's V-cycle:
    reps red-black iterations in place (mf3d_mg_smooth), HIP events on the
    solve stream; one untimed iteration first."""
    import torch

    from macroflow3d_amd import multigrid as mg
    mg.smooth_level(ctx, ws, 0, 1)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    mg.smooth_level(ctx, ws, 0, reps)
    e1.record(ctx.stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / (2 * reps)


def sweep_only(args):
    """One hierarchy (FP64 or mixed) on the workload's K, then --smoother-reps
    timed red-black iterations of its finest level: the only launches of the
    dominant half-sweep kernel in the run, for kernel traces and counters."""
    import torch

    from macroflow3d_amd import multigrid as mg
    from macroflow3d_amd import physics as ph
    from macroflow3d_amd.core import BCSpec, Grid3D
    from macroflow3d_amd.kfield import lognormal_lnK
    from macroflow3d_amd.runtime import HipContext
    ctx = HipContext(0)
    n = args.n
    g = Grid3D(n, n, n)
    K = torch.exp(lognormal_lnK(n, n, n, args.sigma2, args.corr, args.seed, device=ctx.device)[0])
    cfg = (ph.HeadSolveConfig.fp64_direct(args.rtol, args.levels) if args.sweep_only == "fp64"
           else ph.HeadSolveConfig.mixed(args.rtol, args.levels))
    # counter calibration: a plain copy of 2^27 doubles (1 GiB read, 1 GiB write)
    from macroflow3d_amd import numerics as nm
    ca = torch.ones(1 << 27, dtype=torch.float64, device=ctx.device)
    cb = torch.empty_like(ca)
    nm.blas.copy(ctx, ca, cb)
    torch.cuda.synchronize()
    del ca, cb
    ws = ph.FlowWorkspace()
    mg.mg_setup(ctx, ws, g, BCSpec.darcy_x(), cfg, K)
    t = time_smoother(ctx, ws, args.smoother_reps)
    ws.close()
    N = g.num_cells()
    bpc = 20.0 if args.sweep_only == "fp64" else 12.0
    gbs = bpc * N / (t * 1e-3) / 1e9
    print(json.dumps({"sweep": args.sweep_only, "cells": N, "avg_launch_us": round(t * 1e3, 2),
                      "algorithmic_bytes_per_launch": bpc * N, "achieved_GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)
    return 0


def fp64_smoother_line(t_ms, n):
    """The FP64 V-cycle's finest half-sweep (colour-split, faces from
    h = 1/(2K)): 20 B/cell algorithmic (h 8, other-colour x 4, own b 4, own x
    out 4)."""
    alg = 20.0 * n
    gbs = alg / (t_ms * 1e-3) / 1e9
    return {"kernel": "march64_kernel<0> (FP64 colour-split half-sweep)",
            "avg_launch_us": round(t_ms * 1e3, 2), "algorithmic_bytes_per_launch": alg,
            "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def main():
    args = parse()
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc)
    if args.summarize_trace:
        return summarize_trace(args.summarize_trace, 2 * args.smoother_reps)
    if args.sweep_only:
        return sweep_only(args)
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU only: one realisation's head "
                                   "solve is not sharded across GPUs",
                          "n_gpus": args.gpus}))
        return 2
    import torch

    from macroflow3d_amd import multigrid as mg
    from macroflow3d_amd import numerics as nm
    from macroflow3d_amd import physics as ph
    from macroflow3d_amd.core import BCSpec, Grid3D
    from macroflow3d_amd.kfield import lognormal_lnK
    from macroflow3d_amd.runtime import HipContext

    ctx = HipContext(0)
    n = args.n
    g = Grid3D(n, n, n)
    N = g.num_cells()
    bc = BCSpec.darcy_x()
    t0 = time.perf_counter()
    Y, kst = lognormal_lnK(n, n, n, args.sigma2, args.corr, args.seed, device=ctx.device)
    K = torch.exp(Y)
    del Y
    torch.cuda.synchronize()
    log(args, f"K field {n}^3 generated in {time.perf_counter() - t0:.2f} s, embedding "
              f"{kst['embedding']}, min eig ratio {kst['min_eig_ratio']:.2e}")

    # counter calibration: a plain copy of 2^27 doubles (1 GiB read, 1 GiB write)
    ca = torch.ones(1 << 27, dtype=torch.float64, device=ctx.device)
    cb = torch.empty_like(ca)
    nm.blas.copy(ctx, ca, cb)
    torch.cuda.synchronize()
    del ca, cb

    coef = {"bf16": ph.CoefFormat.BF16, "fp16": ph.CoefFormat.FP16,
            "fp32": ph.CoefFormat.FP32}[args.coef]
    # plain FP64 MG-PCG starts, like the pipeline, from the linear head
    # (init_head_guess before solve_head, EnsembleRunner.cu:286-290)
    cfg64 = ph.HeadSolveConfig.fp64_direct(args.rtol, args.levels)
    cfg64.init_guess = ph.InitGuess.REFERENCE
    cfglin = ph.HeadSolveConfig.mixed(args.rtol, args.levels, coef)
    cfglin.init_guess = ph.InitGuess.REFERENCE
    cfgws = ph.HeadSolveConfig.mixed(args.rtol, args.levels, coef)
    cfgws.init_guess = ph.InitGuess.PROVIDED
    net = None
    if not args.no_warm_start:
        from macroflow3d_amd import warmstart as wsm
        net = wsm.WarmStartNet.load(args.weights)

    def step(cfg, h, vel, ws, warm):
        if warm:
            net.predict(ctx, K, g, bc, h)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        ph.compute_velocity_from_head(vel, h, K, g, bc, ctx)
        return res

    def timed(cfg, steps, warmup, h, vel, ws, warm=False):
        res = None
        for _ in range(warmup):
            res = step(cfg, h, vel, ws, warm)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            res = step(cfg, h, vel, ws, warm)
        torch.cuda.synchronize()
        return res, (time.perf_counter() - t) * 1e3 / max(1, steps)

    # plain FP64 MG-PCG (denominator and accuracy reference)
    ws64 = ph.FlowWorkspace()
    h64 = ctx.zeros(N)
    v64 = ph.VelocityField(g, ctx)
    if args.fp64_steps:
        r64, ms64 = timed(cfg64, args.fp64_steps, 1, h64, v64, ws64)
    else:
        r64, ms64 = step(cfg64, h64, v64, ws64, False), float("nan")
    t64 = time_smoother(ctx, ws64, args.smoother_reps)
    ws64.close()
    log(args, f"fp64: {r64.num_iterations} iterations, rel true residual "
              f"{r64.true_residual / r64.bnorm:.3e}, {ms64:.2f} ms per solve; half-sweep "
              f"{t64 * 1e3:.1f} us")

    # mixed precision from the linear head
    ws = ph.FlowWorkspace()
    h = ctx.zeros(N)
    vel = ph.VelocityField(g, ctx)
    rlin, mslin = timed(cfglin, args.steps, args.warmup, h, vel, ws)
    log(args, f"mixed ({args.coef}), linear guess: {rlin.num_iterations} iterations, "
              f"{mslin:.2f} ms per step")
    # headline: mixed precision from the learned warm start
    if net is not None:
        res, ms = timed(cfgws, args.steps, args.warmup, h, vel, ws, warm=True)
        log(args, f"mixed ({args.coef}) + warm start: {res.num_iterations} iterations, rel true "
                  f"residual {res.true_residual / res.bnorm:.3e}, setup {res.setup_ms:.2f} ms, "
                  f"iteration {res.solve_ms:.2f} ms, step {ms:.2f} ms")
    else:
        res, ms = rlin, mslin
        step(cfglin, h, vel, ws, False)
    # iterations the last timed solve enqueued (one preconditioner application
    # each, plus one per restart pass): those past convergence cost a V-cycle
    _, n_pc = ws.solve_counts()
    enqueued = n_pc - res.restarts - 1

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    head_err = rel(h, h64)
    vel_err = max(rel(vel.U, v64.U), rel(vel.V, v64.V), rel(vel.W, v64.W))
    # mass-conservation check of the headline head's padded velocity (outside
    # the timed step): the divergence is the solve's residual b - A h
    pv = ph.PaddedVelocityField(g, ctx)
    ph.compute_velocity_from_head(pv, h, K, g, bc, ctx)
    diag = ph.VelocityDiagnostics(g, ctx)
    ph.compute_velocity_diagnostics(pv, diag, g, ctx)
    div_l2 = float(torch.linalg.vector_norm(diag.divergence))
    diagnostics = {"div_min": float(diag.stats[0, 0]), "div_max": float(diag.stats[0, 1]),
                   "div_mean": float(diag.stats[0, 2]), "div_l2_over_bnorm": div_l2 / res.bnorm,
                   "vorticity_mag_mean": float(diag.stats[1, 2]),
                   "helicity_mean": float(diag.stats[2, 2])}
    log(args, f"velocity diagnostics: ||div|| / ||b|| = {div_l2 / res.bnorm:.3e}, "
              f"mean |curl| {diagnostics['vorticity_mag_mean']:.4e}")
    del pv, diag

    # dominant kernel: finest-level colour sweep of the headline V-cycle,
    # timed in place on the workspace's own level-0 buffers
    mg.smooth_level(ctx, ws, 0, 1)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    mg.smooth_level(ctx, ws, 0, args.smoother_reps)
    e1.record(ctx.stream)
    e1.synchronize()
    t_launch_ms = e0.elapsed_time(e1) / (2 * args.smoother_reps)
    # one colour (half-sweep) of the mixed V-cycle: 3 faces x 2 B + opposite-
    # colour x 2 B + own-colour x write 2 B + own-colour b 2 B per cell
    alg_bytes = {"bf16": 12.0, "fp16": 12.0, "fp32": 18.0}[args.coef] * N
    achieved = alg_bytes / (t_launch_ms * 1e-3) / 1e9
    log(args, f"smoother: {t_launch_ms * 1e3:.1f} us per colour sweep, {achieved:.0f} GB/s "
              f"algorithmic")
    traffic = committed_traffic(n, args.coef)

    out = {
        "metric": METRIC,
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 outer / f32 V-cycle / {args.coef} coefficients",
        "data": "synthetic (seeded log-normal K, FFT circulant embedding)",
        "config": {
            "workload": f"head solve + face velocities, {n}^3 cells, sigma2_lnK={args.sigma2:g}, "
                        f"corr_len={args.corr:g} cells, {args.levels} levels (coarsest "
                        f"{n >> (args.levels - 1)}^3 direct FP64), V(2,2) GSRB, "
                        f"||b - A h|| <= {args.rtol:g} ||b||",
            "cells": N, "seed": args.seed, "method": "mixed-precision MG-PCG",
            "warm_start": ("learned (3-D conv encoder-decoder, BF16 MFMA, inference inside "
                           "the timed step)" if net is not None else "linear head"),
        },
        "iterations": res.num_iterations,
        "iterations_enqueued": enqueued,
        "mixed_linear_guess_ms": round(mslin, 3),
        "mixed_linear_guess_iterations": rlin.num_iterations,
        "rel_true_residual": res.true_residual / res.bnorm,
        "setup_ms": round(res.setup_ms, 3),
        "iteration_ms": round(res.solve_ms, 3),
        "fp64_ms": round(ms64, 3),
        "fp64_guess": "linear head",
        "fp64_iterations": r64.num_iterations,
        "speedup_vs_fp64": round(ms64 / ms, 3) if ms > 0 else None,
        "head_rel_l2_vs_fp64": head_err,
        "velocity_rel_l2_vs_fp64": vel_err,
        "velocity_diagnostics": diagnostics,
        "fp64_smoother": fp64_smoother_line(t64, N),
        "roofline": {
            "kernel": f"colour_split4_kernel<{args.coef}> (finest level, one colour)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_us": round(t_launch_ms * 1e3, 2),
        },
    }
    ws.close()
    if not args.no_cpu:
        torch.cuda.synchronize()
        gpu_heads = {"fp64": (h64.cpu().numpy(), [v64.U.cpu().numpy(), v64.V.cpu().numpy(),
                                                  v64.W.cpu().numpy()]),
                     "headline": (h.cpu().numpy(), [vel.U.cpu().numpy(), vel.V.cpu().numpy(),
                                                    vel.W.cpu().numpy()])}
        out["cpu_baseline"], parity, sums = cpu_baseline(args, K.cpu().numpy(), gpu_heads)
        out["parity_vs_cpu_oracle"] = parity
        out["head_rel_l2_vs_oracle"] = parity["headline"]["head_rel_l2_vs_oracle_1e-13"]
        out["velocity_checksums_vs_oracle"] = {
            "headline_devs_sum_sumsq_max": parity["headline"]["velocity_checksum_devs_vs_oracle_1e-13"],
            "fp64_same_algorithm_devs_sum_sumsq_max":
                parity["fp64_same_algorithm"]["velocity_checksum_devs_vs_oracle_1e-10"]}
        out["oracle_velocity_checksums"] = sums
        log(args, "parity vs oracle: " + json.dumps(parity))
        del gpu_heads
    del K, h, h64, vel, v64
    torch.cuda.empty_cache()
    if not args.no_512:
        out.update(bench_512(args, ctx, coef))
        torch.cuda.empty_cache()
    if not args.no_batch:
        out.update(bench_batch(args, ctx))
    print(json.dumps(out), flush=True)
    return 0


def bench_512(args, ctx, coef):
    """512^3, sigma2_lnK = 4, correlation length 16, 7 levels: plain FP64 and
    mixed-precision solve + velocities from the linear head (K resident)."""
    import torch

    from macroflow3d_amd import physics as ph
    from macroflow3d_amd.core import BCSpec, Grid3D
    from macroflow3d_amd.kfield import lognormal_lnK
    n, levels = 512, 7
    g, bc = Grid3D(n, n, n), BCSpec.darcy_x()
    t0 = time.perf_counter()
    Y, _ = lognormal_lnK(n, n, n, 4.0, 16.0, 3031, device=ctx.device)
    K = torch.exp(Y)
    del Y
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(args, f"K field 512^3 generated in {time.perf_counter() - t0:.2f} s")
    out = {}
    heads = {}
    for name, cfg in (("fp64", ph.HeadSolveConfig.fp64_direct(args.rtol, levels)),
                      ("mixed", ph.HeadSolveConfig.mixed(args.rtol, levels, coef))):
        cfg.init_guess = ph.InitGuess.REFERENCE
        ws = ph.FlowWorkspace()
        h = ctx.zeros(g.num_cells())
        v = ph.VelocityField(g, ctx)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)  # warm-up
        ph.compute_velocity_from_head(v, h, K, g, bc, ctx)
        reps = 2 if name == "fp64" else 3
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
            ph.compute_velocity_from_head(v, h, K, g, bc, ctx)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / reps
        ts = time_smoother(ctx, ws, args.smoother_reps)
        if name == "fp64":
            out["fp64_smoother_512"] = fp64_smoother_line(ts, g.num_cells())
        else:  # 12 B/cell: BF16 faces 6, other-colour x 2, own b 2, own x out 2
            gbs = 12.0 * g.num_cells() / (ts * 1e-3) / 1e9
            out["smoother_512"] = {"kernel": "colour_split4_kernel<bf16> (finest level, one colour)",
                                   "avg_launch_us": round(ts * 1e3, 2),
                                   "achieved_GBs": round(gbs, 1),
                                   "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(args, f"512^3 {name} half-sweep {ts * 1e3:.1f} us")
        ws.close()
        out[f"{name}_ms_512"] = round(ms, 3)
        out[f"{name}_iterations_512"] = res.num_iterations
        out[f"{name}_rel_true_residual_512"] = res.true_residual / res.bnorm
        heads[name] = (h, v)
        log(args, f"512^3 {name}: {res.num_iterations} iterations, {ms:.2f} ms per solve")
        torch.cuda.empty_cache()

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    (h64, v64), (hmx, vmx) = heads["fp64"], heads["mixed"]
    out["head_rel_l2_vs_fp64_512"] = rel(hmx, h64)
    out["velocity_rel_l2_vs_fp64_512"] = max(rel(vmx.U, v64.U), rel(vmx.V, v64.V),
                                             rel(vmx.W, v64.W))
    out["speedup_vs_fp64_512"] = round(out["fp64_ms_512"] / out["mixed_ms_512"], 3)
    return out


def bench_batch(args, ctx):
    """The uneven Monte-Carlo batch (32 x 48^3, 32 x 64^3, 32 x 96^3; sigma2_lnK
    ~ U[0.5, 4], correlation length ~ U[2, 8]; mixed precision, 6 levels, K
    resident): one batched solve of all realisations vs one solve after another."""
    import numpy as np
    import torch

    from macroflow3d_amd import ensemble
    from macroflow3d_amd import physics as ph
    from macroflow3d_amd.core import BCSpec
    reals = ensemble.monte_carlo_set(4000, args.batch_per_size)
    Ks = ensemble.generate_fields(ctx, reals)
    grids = [re.grid for re in reals]
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(args.rtol, 6)
    ws = [ph.FlowWorkspace() for _ in reals]
    hs = [ctx.zeros(gr.num_cells()) for gr in grids]
    res = ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, ws)  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        res = ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, ws)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t) / reps
    for w in ws:
        w.close()
    ensemble.solve_sequential(ctx, reals[:4], Ks[:4], bc, cfg)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, rseq = ensemble.solve_sequential(ctx, reals, Ks, bc, cfg)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t
    its = [r.num_iterations for r in res]
    log(args, f"batch of {len(reals)}: {t_batch * 1e3:.1f} ms batched, {t_seq * 1e3:.1f} ms "
              f"sequential")
    return {"batch_workload": f"{len(reals)} realisations: " + ", ".join(
                f"{args.batch_per_size} x {n}^3" for n in (48, 64, 96)),
            "batched_solves_per_s": round(len(reals) / t_batch, 1),
            "sequential_solves_per_s": round(len(reals) / t_seq, 1),
            "batched_ms": round(t_batch * 1e3, 2), "sequential_ms": round(t_seq * 1e3, 2),
            "batch_iterations_min": int(min(its)), "batch_iterations_mean": float(np.mean(its)),
            "batch_iterations_max": int(max(its)),
            "batch_same_iterations_as_sequential": all(
                a.num_iterations == b.num_iterations for a, b in zip(res, rseq)),
            "batch_all_converged": all(r.converged and r.true_residual <= args.rtol * r.bnorm
                                       for r in res)}


if __name__ == "__main__":
    sys.exit(main())
