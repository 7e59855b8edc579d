"""The rows after the head solve, on the GPU: the FFT circulant-embedding K
generator run on the device (the same field as the oracle's C++ generator for
the same seed, and the exponential covariance it is meant to have), the
Monte-Carlo ensemble loop (seed base + r per realisation, EnsembleRunner.cu:268;
batched and one-after-another solves of uneven grids), and the autotuner
(its choices leave the head unchanged)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from helpers import bc_pair, grid_pair, random_K
from macroflow3d_amd import autotune, ensemble, kfield
from macroflow3d_amd import numerics as nm
from macroflow3d_amd import physics as ph
from macroflow3d_amd.core import BCSpec

pytestmark = pytest.mark.gpu


def test_device_generator_matches_oracle(ctx):
    n = 64
    Yd, st = kfield.lognormal_lnK(n, n, n, 2.0, 6.0, 1234, device=ctx.device)
    Yo, sto = O.lognormal_lnK(n, n, n, 2.0, 6.0, 1234)
    assert tuple(st["embedding"]) == tuple(int(v) for v in sto[:3])
    yd = Yd.cpu().numpy()
    assert np.abs(yd - Yo).max() < 1e-11 * np.abs(Yo).max()
    # and the host run of the same generator
    Yc, _ = kfield.lognormal_lnK(n, n, n, 2.0, 6.0, 1234)
    assert np.abs(yd - Yc.numpy()).max() < 1e-11 * np.abs(Yo).max()


def test_device_generator_covariance(ctx):
    sig2, lam, n = 2.0, 4.0, 64
    acc = {L: [] for L in (0, 1, 2, 4, 8)}
    for seed in range(4):
        Y, st = kfield.lognormal_lnK(n, n, n, sig2, lam, 77 + seed, device=ctx.device)
        assert st["clipped_mass"] == 0.0
        cov = kfield.empirical_covariance(Y, (n, n, n), lags=tuple(acc))
        for L in acc:
            acc[L].append(cov[L])
    for L, vals in acc.items():
        target = sig2 * np.exp(-L / lam)
        assert abs(np.mean(vals) - target) < 0.12 * sig2, (L, np.mean(vals), target)


def test_ensemble_loop_seed_rule_batched_and_sequential(ctx):
    reals = ensemble.monte_carlo_set(500, 2, sizes=(32, 48))
    assert [re.seed for re in reals] == [500, 501, 502, 503]
    Ks = ensemble.generate_fields(ctx, reals)
    re = reals[3]  # the field of realisation r is the generator's field for seed base + r
    Yo, _ = O.lognormal_lnK(re.grid.nx, re.grid.ny, re.grid.nz, re.sigma2, re.corr_len, 503)
    assert np.abs(np.log(Ks[3].cpu().numpy()) - Yo).max() < 1e-11 * np.abs(Yo).max()
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(1e-10, 4)
    hs, res, wss = ensemble.solve_batched(ctx, reals, Ks, bc, cfg)
    hq, rq = ensemble.solve_sequential(ctx, reals, Ks, bc, cfg)
    for a, b, ha, hb in zip(res, rq, hs, hq):
        assert a.converged and a.true_residual <= 1e-10 * a.bnorm
        assert a.num_iterations == b.num_iterations
        assert float(torch.linalg.vector_norm(ha - hb) / torch.linalg.vector_norm(hb)) <= 1e-12
    vel = ensemble.velocities(ctx, reals, hs, Ks, bc)
    for re, v in zip(reals, vel):  # equal total x-flux through every x-plane
        q = v.U.view(re.grid.nz, re.grid.ny, re.grid.nx + 1).sum(dim=(0, 1))
        assert float((q - q.mean()).abs().max() / q.mean().abs()) < 1e-7
    for w in wss:
        w.close()


def test_autotuner_choices_leave_the_head_unchanged(ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("MF3D_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    g, _ = grid_pair(64, 32, 48)
    bc, _ = bc_pair("periodic_yz")
    n = g.num_cells()
    K = torch.as_tensor(random_K(n, 63, 1.5)).to(ctx.device)
    cfg = ph.HeadSolveConfig.mixed(1e-10, 4)
    t1 = autotune.tune_sweep_block(ctx, g, bc, cfg, K, reps=2)
    assert t1["sweep_block"] in autotune.BLOCKS
    assert autotune.tune_sweep_block(ctx, g, bc, cfg, K, reps=2) == t1  # cache hit
    # the chosen sweep block gives the same preconditioned vector as the default
    r = torch.as_tensor(np.random.default_rng(64).standard_normal(n)).to(ctx.device)
    out = []
    for params in ({"sweep_block": 256}, t1):
        ws = ph.FlowWorkspace()
        autotune.apply([ws], params)
        M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, K)
        z = ctx.empty(n)
        M.apply(ctx, r, z)
        torch.cuda.synchronize()
        out.append(z.cpu().numpy())
        ws.close()
    assert np.array_equal(out[0], out[1])
    # whole-solve tuning: the tuned solve reaches the head of the default one
    t2 = autotune.tune_solve(ctx, g, bc, cfg, K, reps=1, blocks=(128, 256))
    assert {"sweep_block", "graph", "op_tile", "op_chunk", "ms"} <= set(t2)
    assert t2["op_chunk"] in autotune.OP_CHUNKS
    heads = []
    for params in ({}, t2):
        ws = ph.FlowWorkspace()
        autotune.apply([ws], params)
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        assert res.converged and res.true_residual <= 1e-10 * res.bnorm
        heads.append(h)
        ws.close()
    assert float(torch.linalg.vector_norm(heads[0] - heads[1]) /
                 torch.linalg.vector_norm(heads[0])) < 1e-10
