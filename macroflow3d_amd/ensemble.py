"""Monte-Carlo ensemble driver of the head solve.

Mirrors the realisation loop of MacroFlow3D's pipeline
(src/runtime/ensemble/EnsembleRunner.cu:259-311: per-realisation seed
``base + r`` (:268), K generation, head solve, face velocities) with the
solves of all realisations batched into one run
(``physics.solve_head_batched``) instead of one after another. Realisations
may have different grids and field statistics.

``python -m macroflow3d_amd.ensemble`` runs the uneven Monte-Carlo batch
(32 x 48^3, 32 x 64^3, 32 x 96^3; sigma2_lnK ~ U[0.5, 4], correlation length
~ U[2, 8] cells) batched and sequentially and prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import autotune
from . import physics as ph
from .core import BCSpec, Grid3D
from .kfield import lognormal_lnK


@dataclass
class Realisation:
    grid: Grid3D
    sigma2: float
    corr_len: float
    seed: int


def monte_carlo_set(base_seed: int = 4000, per_size: int = 32,
                    sizes=(48, 64, 96), sigma2=(0.5, 4.0), corr=(2.0, 8.0)):
    """Realisations of the uneven batch: field statistics drawn from a host
    generator seeded with base_seed; realisation r gets field seed base + r."""
    rng = np.random.default_rng(base_seed)
    out = []
    r = 0
    for n in sizes:
        for _ in range(per_size):
            out.append(Realisation(Grid3D(n, n, n), float(rng.uniform(*sigma2)),
                                   float(rng.uniform(*corr)), base_seed + r))
            r += 1
    return out


def generate_fields(ctx, reals):
    """K = exp(ln K) per realisation, device-resident."""
    return [torch.exp(lognormal_lnK(re.grid.nx, re.grid.ny, re.grid.nz, re.sigma2, re.corr_len,
                                    re.seed, device=ctx.device)[0]) for re in reals]


def solve_batched(ctx, reals, Ks, bc, cfg, workspaces=None, use_graph=True):
    ws = workspaces or [ph.FlowWorkspace() for _ in reals]
    hs = [ctx.zeros(re.grid.num_cells()) for re in reals]
    res = ph.solve_head_batched(hs, Ks, [re.grid for re in reals], bc, cfg, ctx, ws,
                                use_graph=use_graph)
    return hs, res, ws


def solve_sequential(ctx, reals, Ks, bc, cfg):
    ws = ph.FlowWorkspace()
    hs, res = [], []
    for re, K in zip(reals, Ks):
        h = ctx.zeros(re.grid.num_cells())
        res.append(ph.solve_head(h, K, re.grid, bc, cfg, ctx, ws))
        hs.append(h)
    ws.close()
    return hs, res


def velocities(ctx, reals, hs, Ks, bc):
    out = []
    for re, h, K in zip(reals, hs, Ks):
        v = ph.VelocityField(re.grid, ctx)
        ph.compute_velocity_from_head(v, h, K, re.grid, bc, ctx)
        out.append(v)
    return out


def _main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-size", type=int, default=32)
    ap.add_argument("--seed", type=int, default=4000)
    ap.add_argument("--levels", type=int, default=6,
                    help="hierarchy depth (6: coarsest 3^3 or 2^3 for the uneven batch)")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true")
    a = ap.parse_args(argv)
    from .runtime import HipContext
    ctx = HipContext(0)
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(a.rtol, a.levels)
    reals = monte_carlo_set(a.seed, a.per_size)
    Ks = generate_fields(ctx, reals)
    grids = [re.grid for re in reals]
    ws = [ph.FlowWorkspace() for _ in reals]
    hs = [ctx.zeros(g.num_cells()) for g in grids]
    tuned = None
    use_graph = True
    if not a.no_tune:
        tuned = autotune.tune_batch(ctx, hs, Ks, grids, bc, cfg, ws)
        autotune.apply(ws, tuned)
        use_graph = bool(tuned["graph"])
    ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, ws, use_graph=use_graph)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        res = ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, ws, use_graph=use_graph)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / a.reps
    solve_sequential(ctx, reals[:4], Ks[:4], bc, cfg)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hseq, rseq = solve_sequential(ctx, reals, Ks, bc, cfg)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    same = all(r.num_iterations == q.num_iterations for r, q in zip(res, rseq))
    dmax = max(float(torch.linalg.vector_norm(h - q) / torch.linalg.vector_norm(q))
               for h, q in zip(hs, hseq))
    out = {
        "workload": f"{len(reals)} realisations: " + ", ".join(
            f"{a.per_size} x {n}^3" for n in (48, 64, 96)),
        "cells": int(sum(g.num_cells() for g in grids)),
        "batched_ms": round(t_batch * 1e3, 2), "sequential_ms": round(t_seq * 1e3, 2),
        "batched_setup_ms": round(res[0].setup_ms, 2), "batched_iteration_ms": round(res[0].solve_ms, 2),
        "solves_per_s_batched": round(len(reals) / t_batch, 1),
        "solves_per_s_sequential": round(len(reals) / t_seq, 1),
        "speedup": round(t_seq / t_batch, 2),
        "iterations_mean": float(np.mean([r.num_iterations for r in res])),
        "iterations_min": int(min(r.num_iterations for r in res)),
        "iterations_max": int(max(r.num_iterations for r in res)),
        "same_iterations_as_sequential": same, "max_head_rel_diff_vs_sequential": dmax,
        "all_converged": all(r.converged and r.true_residual <= a.rtol * r.bnorm for r in res),
        "tuned": tuned, "graph": use_graph,
    }
    print(json.dumps(out), flush=True)
    for w in ws:
        w.close()
    return 0


if __name__ == "__main__":
    sys.exit(_main())
