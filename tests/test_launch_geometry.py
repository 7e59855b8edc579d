"""Launch-geometry tuning keys change which workgroup computes which planes,
not what is computed: the FP64 V-cycle's sweep workgroup target leaves the
preconditioned vector bitwise unchanged (each half-sweep updates one colour
from the other, so chunk boundaries do not matter) and only regroups the
(r, z) partials; the operator's chunk depth regroups the (p, Ap) partials,
so solves reach the same head to rounding."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import bc_pair, grid_pair, random_K, rel_l2
from macroflow3d_amd import numerics as nm
from macroflow3d_amd import physics as ph

pytestmark = pytest.mark.gpu


def dev(a, ctx):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(ctx.device)


@pytest.mark.parametrize("mix", ["darcy_x", "all_periodic"])
def test_march_target_leaves_fp64_vcycle_bitwise(ctx, mix):
    g, _ = grid_pair(64, 48, 96)
    bc, _ = bc_pair(mix)
    n = g.num_cells()
    K = dev(random_K(n, 71, 1.5), ctx)
    r = dev(np.random.default_rng(72).standard_normal(n), ctx)
    cfg = ph.HeadSolveConfig(mg_levels=4, coarse_solver=ph.CoarseSolver.DIRECT)
    out = []
    for blocks in (2048, 256, 1024, 8192):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_MARCH_BLOCKS, blocks)
        M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, K)
        z = ctx.empty(n)
        M.apply(ctx, r, z)
        torch.cuda.synchronize()
        out.append(z.cpu().numpy())
        ws.close()
    assert np.all(np.isfinite(out[0]))
    for o in out[1:]:
        assert np.array_equal(out[0], o), mix


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_geometry_keys_keep_the_solve(ctx, precision):
    g, _ = grid_pair(96, 64, 64)
    bc, _ = bc_pair("darcy_x")
    n = g.num_cells()
    K = dev(random_K(n, 73, 1.5), ctx)
    cfg = (ph.HeadSolveConfig.fp64_direct(1e-10, 4) if precision == "fp64"
           else ph.HeadSolveConfig.mixed(1e-10, 4))
    out = []
    for chunk, blocks in ((16, 2048), (1, 2048), (5, 512), (64, 8192)):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_OP_CHUNK, chunk)
        ws.set_tuning(ph.FlowWorkspace.TUNE_MARCH_BLOCKS, blocks)
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        assert res.converged and res.true_residual <= 1e-10 * res.bnorm, (chunk, blocks)
        torch.cuda.synchronize()
        out.append((h.cpu().numpy(), res.num_iterations))
        ws.close()
    for h, it in out[1:]:
        assert abs(it - out[0][1]) <= 1
        assert rel_l2(h, out[0][0]) < 1e-9
