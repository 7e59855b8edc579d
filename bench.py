#!/usr/bin/env python3
"""Benchmark: time-to-1e-10 Darcy head solve at 256^3 on one MI355X.

One step = one complete head solve (hierarchy coefficients, RHS, MG-PCG to
||r||/||b|| <= 1e-10) plus the face velocities, on a seeded log-normal K field
(sigma2_lnK = 4, correlation length 8 cells) already resident in HBM.

Also measured here:
  * roofline of the dominant kernel (fine-level GSRB half-sweep): algorithmic
    bytes per launch / average launch time from HIP events on the solve stream;
  * CPU baseline: the CPU FP64 oracle (OpenMP) on the same K field, a bounded
    sample of PCG iterations extrapolated to the GPU's iteration count.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "time-to-1e-10 Darcy solve at 256^3, mixed+warm-start vs FP64; smoother HBM GB/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--sigma2", type=float, default=4.0)
    ap.add_argument("--corr", type=float, default=8.0)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--smoother-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def cpu_baseline(args, K_host, gpu_iters):
    """Oracle at the same size: setup + 1 PCG iteration, then setup + 1 + cpu_iters
    iterations; per-iteration time from the difference, extrapolated."""
    import oracle as O
    threads = O.num_threads()
    g = O.grid(args.n, args.n, args.n)
    b = O.darcy_bc()
    times = []
    for its in (1, 1 + args.cpu_iters):
        p = O.mode_n(mg_levels=args.levels, cg_rtol=args.rtol, cg_max_iter=its)
        t0 = time.perf_counter()
        O.solve_head(g, b, p, K_host)
        times.append(time.perf_counter() - t0)
        log(args, f"cpu oracle: {its} PCG iterations in {times[-1]:.2f} s")
    per_iter = max(times[1] - times[0], 1e-9) / args.cpu_iters
    est = times[0] + (gpu_iters - 1) * per_iter
    return {"value": est * 1e3, "unit": "ms", "cores": threads, "kind": "port",
            "sample": (f"CPU FP64 oracle (OpenMP, {threads} threads) on the same {args.n}^3 K: "
                       f"setup+1 and setup+{1 + args.cpu_iters} PCG iterations timed "
                       f"({times[0]:.2f} s, {times[1]:.2f} s), extrapolated to the GPU's "
                       f"{gpu_iters} iterations"),
            "per_iteration_ms": per_iter * 1e3, "setup_plus_first_iteration_ms": times[0] * 1e3}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the head solve of one "
                                   "realisation is not sharded (see DESIGN.md, sharding)",
                          "n_gpus": args.gpus}))
        return 2
    import torch

    from macroflow3d_amd import multigrid as mg
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

    cfg = ph.HeadSolveConfig.north_star(args.rtol, args.levels)
    ws = ph.FlowWorkspace()
    h = ctx.zeros(N)
    vel = ph.VelocityField(g, ctx)

    def step():
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        ph.compute_velocity_from_head(vel, h, K, g, bc, ctx)
        return res

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for _ in range(args.steps):
        results.append(step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / max(1, args.steps)
    res = results[-1]
    log(args, f"solve: {res.num_iterations} iterations, rel true residual "
              f"{res.true_residual / res.bnorm:.3e}, setup {res.setup_ms:.2f} ms, "
              f"iteration {res.solve_ms:.2f} ms, step {ms_per_step:.2f} ms")

    # dominant kernel: fine-level GSRB, one launch per colour (2 per iteration)
    x = h.clone()
    b = ctx.empty(N)
    ph.build_rhs_head(b, K, g, bc, ctx)
    mg.gsrb_smooth_3d(ctx, g, x, b, K, 1, bc)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    for _ in range(args.smoother_reps):
        mg.gsrb_smooth_3d(ctx, g, x, b, K, 1, bc)
    e1.record(ctx.stream)
    e1.synchronize()
    launches = 2 * args.smoother_reps
    t_launch_ms = e0.elapsed_time(e1) / launches
    alg_bytes = 20.0 * N  # K 8N + opposite-colour x 4N + own-colour x 4N + own-colour b 4N
    achieved = alg_bytes / (t_launch_ms * 1e-3) / 1e9
    log(args, f"smoother: {t_launch_ms * 1e3:.1f} us per half-sweep, {achieved:.0f} GB/s "
              f"algorithmic")

    out = {
        "metric": METRIC,
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded log-normal K, FFT circulant embedding)",
        "config": {
            "workload": f"plain FP64 MG-PCG head solve + face velocities, {n}^3 cells, "
                        f"sigma2_lnK={args.sigma2:g}, corr_len={args.corr:g} cells, "
                        f"{args.levels} levels (coarsest {n >> (args.levels - 1)}^3 direct), "
                        f"V(2,2) GSRB, ||r||/||b|| <= {args.rtol:g}",
            "cells": N, "seed": args.seed, "precision": "fp64", "warm_start": "zero guess",
        },
        "iterations": res.num_iterations,
        "converged": res.converged,
        "rel_true_residual": res.true_residual / res.bnorm,
        "setup_ms": round(res.setup_ms, 3),
        "iteration_ms": round(res.solve_ms, 3),
        "roofline": {
            "kernel": "gsrb_colour_kernel (fine level, one colour)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_us": round(t_launch_ms * 1e3, 2),
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, K.cpu().numpy(), res.num_iterations)
    print(json.dumps(out), flush=True)
    ws.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
