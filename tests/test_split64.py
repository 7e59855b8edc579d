"""The FP64 V-cycle on colour-split levels (x, b and h = 1 / (2 K) stored
colour by colour, faces formed from h) against the natural-order FP64 V-cycle
with stored faces and against the oracle: every boundary mix, uneven and
one-sided smoothing counts (including no pre- or no post-smoothing, which take
the conversion and prolongation kernels instead of the folded ones), a split
coarse level below a split level and a natural one below it, every sweep
workgroup size, solves in both solver semantics and the MG-only solver."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from helpers import BC_MIXES, bc_pair, grid_pair, random_K, rel_err, rel_l2
from macroflow3d_amd import numerics as nm
from macroflow3d_amd import physics as ph

pytestmark = pytest.mark.gpu


def dev(a, ctx):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(ctx.device)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _precondition(ctx, g, bc, cfg, K, r, split, tb=256):
    ws = ph.FlowWorkspace()
    ws.set_tuning(ph.FlowWorkspace.TUNE_SPLIT, split)
    ws.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, tb)
    M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, dev(K, ctx))
    z = ctx.empty(g.num_cells())
    M.apply(ctx, dev(r, ctx), z)
    out = host(z)
    ws.close()
    return out


@pytest.mark.parametrize("mix", list(BC_MIXES))
@pytest.mark.parametrize("shape,levels,pre_post", [((32, 16, 24), 3, (2, 2)),
                                                   ((64, 64, 64), 4, (1, 3)),
                                                   ((48, 32, 32), 4, (0, 2)),
                                                   ((32, 32, 16), 3, (2, 0)),
                                                   ((128, 32, 32), 5, (3, 1))])
def test_split_fp64_vcycle_matches_natural_and_oracle(ctx, mix, shape, levels, pre_post):
    g, og = grid_pair(*shape)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    K = random_K(n, 31, 1.5)
    r = np.random.default_rng(32).standard_normal(n)
    pre, post = pre_post
    cfg = ph.HeadSolveConfig(mg_levels=levels, mg_pre_smooth=pre, mg_post_smooth=post,
                             coarse_solver=ph.CoarseSolver.DIRECT)
    zs = _precondition(ctx, g, bc, cfg, K, r, 1)
    zn = _precondition(ctx, g, bc, cfg, K, r, 0)
    assert np.all(np.isfinite(zs))
    # faces 1 / (h_a + h_b) vs stored 2 / (1/K_a + 1/K_b): rounding only; the
    # red-only residual drops black residuals that vanish up to rounding
    assert rel_err(zs, zn) < 1e-11, (mix, shape, pre_post)
    p = O.SolveParams(mg_levels=levels, mg_pre_smooth=pre, mg_post_smooth=post,
                      coarse_solver=1, pin_mode=0)
    zo = O.v_cycle_precond(og, obc, p, K, r)
    assert rel_err(zs, zo) < 1e-10, (mix, shape, pre_post)


@pytest.mark.parametrize("tb", [64, 128, 512])
def test_split_fp64_sweep_block_is_result_invariant(ctx, tb):
    g, _ = grid_pair(64, 32, 48)
    bc, _ = bc_pair("periodic_yz")
    n = g.num_cells()
    K = random_K(n, 33, 1.5)
    r = np.random.default_rng(34).standard_normal(n)
    cfg = ph.HeadSolveConfig(mg_levels=4, coarse_solver=ph.CoarseSolver.DIRECT)
    assert np.array_equal(_precondition(ctx, g, bc, cfg, K, r, 1, 256),
                          _precondition(ctx, g, bc, cfg, K, r, 1, tb))


@pytest.mark.parametrize("mix", ["darcy_x", "all_dirichlet", "periodic_x_dirichlet_z"])
@pytest.mark.parametrize("solver", [ph.HeadSolverType.PCG_MG, ph.HeadSolverType.MG])
def test_split_fp64_solves_match_oracle_reference_semantics(ctx, mix, solver):
    """Reference semantics (GSRB coarsest, ||r|| / ||r0|| every 10 iterations):
    the same iteration count as the oracle and the same head."""
    g, og = grid_pair(64, 32, 32)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    K = random_K(n, 35, 1.2)
    cfg = ph.HeadSolveConfig(solver_type=solver, mg_levels=4, cg_rtol=1e-10, rtol=1e-9,
                             mg_max_cycles=60)
    ws = ph.FlowWorkspace()
    h = ctx.zeros(n)
    res = ph.solve_head(h, dev(K, ctx), g, bc, cfg, ctx, ws)
    ws.close()
    p = O.mode_r(solver_type=int(solver), mg_levels=4, cg_rtol=1e-10, rtol=1e-9,
                 mg_max_cycles=60)
    ho, ro = O.solve_head(og, obc, p, K)
    assert res.converged and ro.converged
    assert res.num_iterations == ro.num_iterations, (res.num_iterations, ro.num_iterations)
    assert rel_l2(host(h), ho) < 1e-10


def test_split_fp64_256cubed_matches_natural_order(ctx):
    """The 256^3 FP64 solve (sigma2 = 4, correlation length 8, 6 levels):
    same iterations, same head as the natural-order FP64 V-cycle."""
    from macroflow3d_amd import kfield
    n = 256
    print("split fp64: 256^3 K field", flush=True)
    K = torch.exp(kfield.lognormal_lnK(n, n, n, 4.0, 8.0, 2024, device=ctx.device)[0])
    g = grid_pair(n, n, n)[0]
    bc = bc_pair("darcy_x")[0]
    out = []
    for split in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_SPLIT, split)
        h = ctx.zeros(g.num_cells())
        cfg = ph.HeadSolveConfig.fp64_direct(1e-10, 6)
        cfg.init_guess = ph.InitGuess.REFERENCE
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        assert res.converged and res.true_residual <= 1e-10 * res.bnorm
        out.append((h, res.num_iterations))
        ws.close()
        print(f"split fp64: split={split} {res.num_iterations} iterations", flush=True)
    assert out[0][1] == out[1][1]
    assert float(torch.linalg.vector_norm(out[0][0] - out[1][0]) /
                 torch.linalg.vector_norm(out[1][0])) < 1e-12


@pytest.mark.parametrize("mix", ["darcy_x", "all_periodic", "periodic_x_dirichlet_z"])
@pytest.mark.parametrize("semantics", ["reference", "fp64_direct"])
def test_split_fp64_folded_r_update_equals_separate_kernels(ctx, mix, semantics):
    """The FP64 PCG with r -= alpha Ap in the V-cycle's first half-sweep and
    x, p updated together after it takes the same iterations to bitwise the
    same head as the separate r/x and p kernels: every per-cell formula is
    the same, only the (r, r) sum is grouped by march workgroup."""
    g, _ = grid_pair(64, 48, 32)
    bc, _ = bc_pair(mix)
    n = g.num_cells()
    K = dev(random_K(n, 37, 1.5), ctx)
    cfg = (ph.HeadSolveConfig(mg_levels=4, cg_rtol=1e-10) if semantics == "reference"
           else ph.HeadSolveConfig.fp64_direct(1e-10, 4))
    out = []
    for fold in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_FOLD_R, fold)
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        assert res.converged, (mix, semantics, fold)
        out.append((host(h), res.num_iterations, res.restarts))
        ws.close()
    assert out[0][1:] == out[1][1:], (out[0][1:], out[1][1:])
    assert np.array_equal(out[0][0], out[1][0])
