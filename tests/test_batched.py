"""Batched head solves (mf3d_solve_head_batched): every realisation of a
batch with uneven grids reproduces its own single solve (same iterations,
same head to rounding), with and without the captured hipGraph, and meets
||b - A h|| <= 1e-10 ||b||."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from macroflow3d_amd import _lib
from macroflow3d_amd import physics as ph
from macroflow3d_amd.core import BCSpec, Grid3D
from macroflow3d_amd.kfield import lognormal_lnK


def _realisations(ctx, shapes, seed0=500):
    rng = np.random.default_rng(seed0)
    out = []
    for r, sh in enumerate(shapes):
        s2, lam = rng.uniform(0.5, 4.0), rng.uniform(2.0, 8.0)
        lnK, _ = lognormal_lnK(*sh, s2, lam, seed0 + r, device=ctx.device)
        out.append((Grid3D(*sh), torch.exp(lnK)))
    return out


def test_batched_rejects_unsupported_config():
    cfg = ph.HeadSolveConfig.fp64_direct(1e-10, 4).c()  # FP64 preconditioner
    bc = BCSpec.darcy_x().c()
    g = (_lib.mf3d_grid * 1)(Grid3D(32, 32, 32).c())
    null = (C.c_void_p * 1)(None)
    n = (C.c_size_t * 1)(32 ** 3)
    res = (_lib.mf3d_head_solve_result * 1)()
    st = _lib.lib().mf3d_solve_head_batched(1, null, None, g, C.byref(bc), C.byref(cfg), null,
                                            null, n, 0, res)
    assert st == 7  # MF3D_ERR_UNSUPPORTED
    assert b"batched" in _lib.lib().mf3d_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [True, False])
def test_batched_matches_single_solves(ctx, use_graph):
    shapes = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (32, 48, 64), (48, 48, 48), (64, 32, 48)]
    reals = _realisations(ctx, shapes)
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(1e-10, 4)
    print(f"batched solve of {len(shapes)} realisations graph={use_graph}", flush=True)
    hs = [ctx.zeros(g.num_cells()) for g, _ in reals]
    wss = [ph.FlowWorkspace() for _ in reals]
    res = ph.solve_head_batched(hs, [K for _, K in reals], [g for g, _ in reals], bc, cfg, ctx, wss,
                                use_graph=use_graph)
    ws1 = ph.FlowWorkspace()
    for (g, K), h, r in zip(reals, hs, res):
        h1 = ctx.zeros(g.num_cells())
        r1 = ph.solve_head(h1, K, g, bc, cfg, ctx, ws1)
        assert r.converged and r.true_residual <= 1e-10 * r.bnorm
        assert r.num_iterations == r1.num_iterations
        d = float(torch.linalg.vector_norm(h - h1) / torch.linalg.vector_norm(h1))
        assert d <= 1e-12, d
    for w in wss + [ws1]:
        w.close()


@pytest.mark.gpu
def test_batched_second_call_reuses_workspaces(ctx):
    """A second batched call on the same workspaces (new fields) is as exact."""
    shapes = [(32, 32, 32), (64, 64, 64)]
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(1e-10, 4)
    wss = [ph.FlowWorkspace() for _ in shapes]
    for seed in (900, 901):
        reals = _realisations(ctx, shapes, seed)
        hs = [ctx.zeros(g.num_cells()) for g, _ in reals]
        res = ph.solve_head_batched(hs, [K for _, K in reals], [g for g, _ in reals], bc, cfg, ctx,
                                    wss)
        for (g, K), h, r in zip(reals, hs, res):
            ws1 = ph.FlowWorkspace()
            h1 = ctx.zeros(g.num_cells())
            r1 = ph.solve_head(h1, K, g, bc, cfg, ctx, ws1)
            ws1.close()
            assert r.num_iterations == r1.num_iterations
            assert float(torch.linalg.vector_norm(h - h1) / torch.linalg.vector_norm(h1)) <= 1e-12
    for w in wss:
        w.close()

