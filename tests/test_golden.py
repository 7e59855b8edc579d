"""The committed 32^3 head-solve fixtures (tests/golden/config_32cubed.*,
written by tests/golden/make_golden.py with the CPU oracle).

CPU tests: the oracle still reproduces the fixture (generator, both solver
semantics, velocity checksums), the fixture satisfies the discrete physics
(equal x-flux through every x-plane, no flow through the no-flow faces), and
the package's torch generator draws the same ln K. GPU tests: the HIP solve on
the fixture K reproduces the fixture head and velocity checksums."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest
import torch

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def golden():
    arrays = np.load(os.path.join(HERE, "config_32cubed.npz"), allow_pickle=False)
    with open(os.path.join(HERE, "config_32cubed.json")) as f:
        meta = json.load(f)
    return {k: arrays[k] for k in arrays.files}, meta


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _checksums_close(U, V, W, ref, tol):
    for c, a in (("U", U), ("V", V), ("W", W)):
        r = ref[c]
        assert abs(a.sum() - r["sum"]) <= tol * max(1.0, abs(r["sum"])), c
        assert abs((a * a).sum() - r["sum_sq"]) <= tol * r["sum_sq"], c
        assert abs(np.abs(a).max() - r["max_abs"]) <= tol * r["max_abs"], c


def test_oracle_generator_reproduces_fixture_K(golden):
    arrays, meta = golden
    lnK, _ = O.lognormal_lnK(meta["n"], meta["n"], meta["n"], meta["sigma2"], meta["corr_len"],
                             meta["seed"])
    assert _rel(np.exp(lnK), arrays["K"]) < 1e-13


def test_torch_generator_matches_fixture_K(golden):
    from macroflow3d_amd import kfield
    arrays, meta = golden
    n = meta["n"]
    lnK = kfield.lognormal_lnK(n, n, n, meta["sigma2"], meta["corr_len"], meta["seed"])[0]
    assert _rel(np.exp(lnK.numpy()), arrays["K"]) < 1e-12


@pytest.mark.parametrize("mode", ["mode_r", "mode_n"])
def test_oracle_reproduces_fixture_solve(golden, mode):
    arrays, meta = golden
    n = meta["n"]
    g, bc = O.grid(n, n, n), O.darcy_bc()
    params = (O.mode_r if mode == "mode_r" else O.mode_n)(mg_levels=meta["mg_levels"],
                                                          cg_rtol=meta["rtol"])
    h, res = O.solve_head(g, bc, params, arrays["K"])
    assert res.num_iterations == meta[mode]["iterations"] and res.converged
    assert _rel(h, arrays[f"h_{mode}"]) < 1e-12
    U, V, W = O.velocity_compact(g, bc, h, arrays["K"])
    _checksums_close(U, V, W, meta[mode]["velocity_checksums"], 1e-10)


@pytest.mark.parametrize("mode", ["mode_r", "mode_n"])
def test_fixture_conserves_mass(golden, mode):
    """Converged to 1e-10: the x-flux through every x-plane is the same and the
    no-flow faces carry none."""
    arrays, meta = golden
    n = meta["n"]
    g, bc = O.grid(n, n, n), O.darcy_bc()
    U, V, W = O.velocity_compact(g, bc, arrays[f"h_{mode}"], arrays["K"])
    plane = U.reshape(n, n, n + 1).sum(axis=(0, 1))
    assert np.abs(plane - plane.mean()).max() < 1e-7 * abs(plane.mean())
    V3, W3 = V.reshape(n, n + 1, n), W.reshape(n + 1, n, n)
    assert np.abs(V3[:, 0, :]).max() == 0 and np.abs(V3[:, n, :]).max() == 0
    assert np.abs(W3[0]).max() == 0 and np.abs(W3[n]).max() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["mode_r", "mode_n"])
def test_gpu_solve_reproduces_fixture(ctx, golden, mode):
    from macroflow3d_amd import physics as ph
    from macroflow3d_amd.core import BCSpec, Grid3D
    arrays, meta = golden
    n = meta["n"]
    g, bc = Grid3D(n, n, n), BCSpec.darcy_x()
    if mode == "mode_r":
        cfg = ph.HeadSolveConfig(mg_levels=meta["mg_levels"], cg_rtol=meta["rtol"])
    else:
        cfg = ph.HeadSolveConfig.fp64_direct(meta["rtol"], mg_levels=meta["mg_levels"])
    K = torch.as_tensor(arrays["K"]).to(ctx.device)
    h = ctx.zeros(g.num_cells())
    ws = ph.FlowWorkspace()
    res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
    vel = ph.VelocityField(g, ctx)
    ph.compute_velocity_from_head(vel, h, K, g, bc, ctx)
    torch.cuda.synchronize()
    assert res.converged and res.num_iterations == meta[mode]["iterations"]
    assert _rel(h.cpu().numpy(), arrays[f"h_{mode}"]) < 1e-10
    _checksums_close(vel.U.cpu().numpy(), vel.V.cpu().numpy(), vel.W.cpu().numpy(),
                     meta[mode]["velocity_checksums"], 1e-9)
    ws.close()
