"""Dense coarsest solve: the band factor of the coarsest operator (host
thread for a single solve, device kernel for a batched one) and its explicit
inverse -L^{-T} L^{-1} on the GPU.

A one-level hierarchy makes the V-cycle exactly the coarsest solve, so the
preconditioned vector z must solve A z = r: checked by the oracle's residual
(||r - A z|| <= 1e-13 ||r||, backward stability of a Cholesky solve; measured
2e-14 at most) and, up to 1024 cells, against the oracle's own dense coarsest
solve. The shapes cover both inverse variants (band width w <= 64 in blocks
staged in LDS, wider bands row by row), the largest dense level (16^3 = 4096
cells), periodic wraps whose neighbours coincide (a dimension of 2) or are the
cell itself (a dimension of 1), and the pinned singular mixes. The batched
solve's device factor is checked against single solves, in its global-memory
variant (band 256) and its LDS window at its largest (band 88). An operator
that is not definite (negative K) is reported as MF3D_ERR_SINGULAR by the
solve, by the first mf3d_mg_precondition and by the batched solve."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from helpers import bc_pair, grid_pair, random_K
from macroflow3d_amd import _lib
from macroflow3d_amd import numerics as nm
from macroflow3d_amd import physics as ph
from macroflow3d_amd.core import BCSpec, Grid3D

pytestmark = pytest.mark.gpu

CASES = [
    ((8, 8, 8), "darcy_x"),                 # w = 64: inverse in LDS-staged blocks
    ((8, 8, 8), "periodic_x_dirichlet_z"),  # w = 64
    ((2, 3, 4), "all_neumann"),             # pinned, w = 6
    ((2, 2, 2), "all_periodic"),            # pinned, both x-neighbours the same cell
    ((5, 4, 1), "all_periodic"),            # z-faces wrap onto the cell itself
    ((12, 10, 9), "all_dirichlet"),         # w = 120: inverse row by row
    ((6, 5, 7), "periodic_yz"),             # w = 180 (z wrap)
    ((16, 16, 16), "darcy_x"),              # 4096 cells, w = 256
]


def dev(a, ctx):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(ctx.device)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def pin_for(name):
    return name in ("all_neumann", "all_periodic")


def direct_cfg(levels=1):
    return ph.HeadSolveConfig(mg_levels=levels, coarse_solver=ph.CoarseSolver.DIRECT)


@pytest.mark.parametrize("shape,mix", CASES)
def test_coarsest_direct_solve_solves_the_operator(ctx, shape, mix):
    g, og = grid_pair(*shape, dx=0.5)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    K = random_K(n, 71, 1.5)
    r = np.random.default_rng(72).standard_normal(n)
    print(f"coarsest direct solve {shape} {mix}", flush=True)
    ws = ph.FlowWorkspace()
    M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, direct_cfg(), dev(K, ctx))
    z = ctx.empty(n)
    M.apply(ctx, dev(r, ctx), z)
    zg = host(z)
    res = O.compute_residual_3d(og, obc, pin_for(mix), zg, r, K)
    assert np.linalg.norm(res) <= 1e-13 * np.linalg.norm(r), (shape, mix)
    if n <= 1024:  # the oracle's dense solve is O(n^3) on one core
        zo = O.v_cycle_precond(og, obc, O.SolveParams(mg_levels=1, coarse_solver=1), K, r)
        assert np.linalg.norm(zg - zo) <= 1e-10 * np.linalg.norm(zo), (shape, mix)
    # a second set-up on the same workspace with another K
    K2 = random_K(n, 73, 1.0)
    M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, direct_cfg(), dev(K2, ctx))
    M.apply(ctx, dev(r, ctx), z)
    res = O.compute_residual_3d(og, obc, pin_for(mix), host(z), r, K2)
    assert np.linalg.norm(res) <= 1e-13 * np.linalg.norm(r), (shape, mix)
    ws.close()


def test_setup_whose_inverse_is_never_used_then_a_new_setup(ctx):
    """A set-up whose coarsest inverse no application consumes, then another
    set-up on the same workspace: the same operator size, a wider band on the
    same grid (a z wrap: the pinned band staging grows while the unused
    inverse's upload may still be in flight), a larger grid and a smaller one.
    The inverse then applied solves the new operator (residual <= 1e-13)."""
    ws = ph.FlowWorkspace()
    steps = [((8, 8, 8), "darcy_x"), ((8, 8, 8), "darcy_x"), ((8, 8, 8), "periodic_yz"),
             ((16, 16, 16), "darcy_x"), ((6, 5, 7), "periodic_yz")]
    for t, ((shape, mix), nxt) in enumerate(zip(steps, steps[1:])):
        print(f"unused set-up {shape} {mix}, then {nxt}", flush=True)
        g, _ = grid_pair(*shape, dx=0.5)
        bc, _ = bc_pair(mix)
        nm.MultigridPreconditioner.build(ctx, ws, g, bc, direct_cfg(),
                                         dev(random_K(g.num_cells(), 90 + t, 1.5), ctx))
        g, og = grid_pair(*nxt[0], dx=0.5)
        bc, obc = bc_pair(nxt[1])
        n = g.num_cells()
        K = random_K(n, 95 + t, 1.5)
        M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, direct_cfg(), dev(K, ctx))
        r = np.random.default_rng(96 + t).standard_normal(n)
        z = ctx.empty(n)
        M.apply(ctx, dev(r, ctx), z)
        res = O.compute_residual_3d(og, obc, pin_for(nxt[1]), host(z), r, K)
        assert np.linalg.norm(res) <= 1e-13 * np.linalg.norm(r), nxt
    ws.close()


def test_operator_that_is_not_definite_is_reported(ctx):
    g = Grid3D(8, 8, 8)
    bc = BCSpec.darcy_x()
    n = g.num_cells()
    Kneg = dev(-random_K(n, 74), ctx)
    print("negative K: solve, precondition, batch", flush=True)
    ws = ph.FlowWorkspace()
    h = ctx.zeros(n)
    cfg = ph.HeadSolveConfig(mg_levels=2, coarse_solver=ph.CoarseSolver.DIRECT, cg_max_iter=5)
    with pytest.raises(_lib.MF3DError) as e:
        ph.solve_head(h, Kneg, g, bc, cfg, ctx, ws)
    assert e.value.status == 6 and "pivot 0" in str(e.value)
    M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, direct_cfg(), Kneg)
    z = ctx.empty(n)
    with pytest.raises(_lib.MF3DError) as e:
        M.apply(ctx, dev(np.ones(n), ctx), z)
    assert e.value.status == 6
    # the same workspace then solves a definite problem
    Kpos = dev(random_K(n, 75), ctx)
    res = ph.solve_head(h, Kpos, g, bc, ph.HeadSolveConfig.fp64_direct(1e-10, 2), ctx, ws)
    assert res.converged
    # batched: the realisation that is not definite is named
    grids = [Grid3D(32, 32, 32), Grid3D(32, 32, 32)]
    Ks = [dev(random_K(32 ** 3, 76), ctx), dev(-random_K(32 ** 3, 77), ctx)]
    hs = [ctx.zeros(32 ** 3) for _ in grids]
    wss = [ph.FlowWorkspace() for _ in grids]
    bcfg = ph.HeadSolveConfig.mixed(1e-10, 4)
    bcfg.cg_max_iter = 5
    with pytest.raises(_lib.MF3DError) as e:
        ph.solve_head_batched(hs, Ks, grids, bc, bcfg, ctx, wss)
    assert e.value.status == 6 and "realisation 1" in str(e.value)
    for w in wss + [ws]:
        w.close()


@pytest.mark.parametrize("shape,levels", [((64, 64, 64), 3), ((64, 88, 24), 4)],
                         ids=["coarsest16cubed_band256", "coarsest8x11x3_band88"])
def test_batched_coarsest_factor_matches_single_solves(ctx, shape, levels):
    """Batched solves factor their coarsest operators on the device: a 16^3
    coarsest level (band width 256) takes the variant working in global
    memory, an 8 x 11 x 3 one (band 88) the LDS window at its largest
    (w + 1 = 89 of at most 90, 64,080 bytes). Each realisation reproduces its
    single solve, whose factor is computed on the host: same iterations,
    head within 1e-12."""
    from macroflow3d_amd.kfield import lognormal_lnK
    from macroflow3d_amd.multigrid import hierarchy_grids
    grids = [Grid3D(*shape), Grid3D(*shape)]
    cg = hierarchy_grids(grids[0], levels)[-1]
    assert cg.nx * cg.ny == (256 if levels == 3 else 88)  # the band width (no wraps)
    Ks = [torch.exp(lognormal_lnK(*shape, 2.0, 6.0, 91 + r, device=ctx.device)[0])
          for r in range(2)]
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(1e-10, levels)
    print(f"batched 2 x {shape}, {levels} levels", flush=True)
    hs = [ctx.zeros(g.num_cells()) for g in grids]
    wss = [ph.FlowWorkspace() for _ in grids]
    res = ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, wss)
    ws1 = ph.FlowWorkspace()
    for g, K, h, r in zip(grids, Ks, hs, res):
        h1 = ctx.zeros(g.num_cells())
        r1 = ph.solve_head(h1, K, g, bc, cfg, ctx, ws1)
        assert r.converged and r.true_residual <= 1e-10 * r.bnorm
        assert r.num_iterations == r1.num_iterations
        d = float(torch.linalg.vector_norm(h - h1) / torch.linalg.vector_norm(h1))
        assert d <= 1e-12, d
    for w in wss + [ws1]:
        w.close()
