#!/usr/bin/env python3
"""Writes the 32^3 head-solve fixtures (config_32cubed.npz / .json) with the CPU oracle.

Workload: 32^3 cells, unit spacing, ln K Gaussian (mean 0, sigma2 = 1,
exponential covariance, correlation length 4 cells, FFT circulant embedding,
seed 1), h = 1 at x = 0 and h = 0 at x = L, no-flow elsewhere; MG-PCG with
3 levels, V(2,2) GSRB, relative residual 1e-10, then face velocities. Two
solver semantics: the reference's (linear guess, 50 GSRB sweeps on the
coarsest level, ||r||/||r0|| tested every 10 iterations) and the extension's
(zero guess, dense coarsest solve, ||r||/||b|| every iteration).

The arrays are plain float64 (np.load(..., allow_pickle=False)); the JSON holds
iteration counts, residuals and velocity checksums (sum, sum of squares and
max |.| per component).

Usage: python tests/golden/make_golden.py   (needs oracle/libmf3d_oracle.so)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

N, SIGMA2, CORR, SEED, LEVELS, RTOL = 32, 1.0, 4.0, 1, 3, 1e-10


def checksums(U, V, W):
    return {c: {"sum": float(a.sum()), "sum_sq": float((a * a).sum()),
                "max_abs": float(np.abs(a).max())} for c, a in (("U", U), ("V", V), ("W", W))}


def main() -> int:
    g = O.grid(N, N, N)
    bc = O.darcy_bc()
    lnK, stats = O.lognormal_lnK(N, N, N, SIGMA2, CORR, SEED)
    K = np.exp(lnK)
    meta = {"n": N, "sigma2": SIGMA2, "corr_len": CORR, "seed": SEED, "mg_levels": LEVELS,
            "rtol": RTOL, "embedding_stats": [float(v) for v in stats]}
    out = {"K": K}
    for name, params in (("mode_r", O.mode_r(mg_levels=LEVELS, cg_rtol=RTOL)),
                         ("mode_n", O.mode_n(mg_levels=LEVELS, cg_rtol=RTOL))):
        h, res = O.solve_head(g, bc, params, K)
        U, V, W = O.velocity_compact(g, bc, h, K)
        out[f"h_{name}"] = h
        meta[name] = {"iterations": int(res.num_iterations), "converged": int(res.converged),
                      "initial_residual": res.initial_residual,
                      "final_residual": res.final_residual, "bnorm": res.bnorm,
                      "true_residual": res.true_residual,
                      "velocity_checksums": checksums(U, V, W)}
    np.savez(os.path.join(HERE, "config_32cubed.npz"), **out)
    with open(os.path.join(HERE, "config_32cubed.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps({k: meta[k] for k in ("mode_r", "mode_n")}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
