"""GPU parity: every HIP kernel of the path against the CPU FP64 oracle,
called through the C ABI (libmf3d_hip.so via ctypes).

Tolerances: the kernels evaluate the oracle's formulas term by term in FP64;
the GPU may fuse multiply-adds, so single operations agree to a few ulp
(relative 1e-12 of the field's max); solves compare heads in relative L2
(north-star bar 1e-8)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from helpers import BC_MIXES, bc_pair, grid_pair, random_K, rel_err, rel_l2
from macroflow3d_amd import multigrid as mg
from macroflow3d_amd import numerics as nm
from macroflow3d_amd import physics as ph
from macroflow3d_amd.core import PinSpec

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (16, 12, 10), (2, 6, 4), (32, 32, 32)]
MIXES = list(BC_MIXES)
OP_TOL = 1e-12


def dev(a, ctx):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(ctx.device)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def pin_for(name):
    return name in ("all_neumann", "all_periodic")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mix", MIXES)
def test_gsrb_matches_oracle(ctx, shape, mix):
    g, og = grid_pair(*shape, dx=0.5)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    rng = np.random.default_rng(1)
    K, x0, b = random_K(n, 2), rng.standard_normal(n), rng.standard_normal(n)
    pin = pin_for(mix)
    for iters in (1, 3):
        xg = dev(x0, ctx)
        mg.gsrb_smooth_3d(ctx, g, xg, dev(b, ctx), dev(K, ctx), iters, bc, PinSpec(pin))
        xo = O.gsrb_smooth_3d(og, obc, pin, x0, b, K, iters)
        assert rel_err(host(xg), xo) < OP_TOL, (mix, shape, iters)


@pytest.mark.parametrize("pre_post", [(2, 2), (1, 3), (3, 1), (0, 2)])
def test_vcycle_fast_levels_match_oracle(ctx, workspace, pre_post):
    """64^3 FP64 V-cycle with uneven and one-sided smoothing counts against the
    oracle's."""
    g, og = grid_pair(64, 64, 64)
    bc, obc = bc_pair("darcy_x")
    n = g.num_cells()
    K = random_K(n, 23, 1.5)
    r = np.random.default_rng(24).standard_normal(n)
    pre, post = pre_post
    cfg = ph.HeadSolveConfig(mg_levels=4, mg_pre_smooth=pre, mg_post_smooth=post,
                             coarse_solver=ph.CoarseSolver.DIRECT)
    ws = ph.FlowWorkspace()
    M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, dev(K, ctx))
    z = ctx.empty(n)
    M.apply(ctx, dev(r, ctx), z)
    p = O.SolveParams(mg_levels=4, mg_pre_smooth=pre, mg_post_smooth=post, coarse_solver=1)
    zo = O.v_cycle_precond(og, obc, p, K, r)
    assert rel_err(host(z), zo) < 1e-10, pre_post


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mix", MIXES)
def test_residual_and_operator_match_oracle(ctx, shape, mix):
    g, og = grid_pair(*shape, dx=2.0)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    rng = np.random.default_rng(3)
    K, x, b = random_K(n, 4), rng.standard_normal(n), rng.standard_normal(n)
    pin = pin_for(mix)
    r = ctx.empty(n)
    mg.compute_residual_3d(ctx, g, dev(x, ctx), dev(b, ctx), dev(K, ctx), r, bc, PinSpec(pin))
    assert rel_err(host(r), O.compute_residual_3d(og, obc, pin, x, b, K)) < OP_TOL
    y = ctx.empty(n)
    nm.VarCoeffLaplacian(g, dev(K, ctx), bc, PinSpec(pin)).apply(ctx, dev(x, ctx), y)
    assert rel_err(host(y), O.varcoeff_apply(og, obc, pin, K, x)) < OP_TOL


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (64, 32, 16)])
def test_transfers_and_coarsening_match_oracle(ctx, shape):
    g, og = grid_pair(*shape)
    c, oc = grid_pair(shape[0] // 2, shape[1] // 2, shape[2] // 2, 2.0)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(g.num_cells())
    bcoarse = ctx.empty(c.num_cells())
    mg.restrict_3d(ctx, g, c, dev(f, ctx), bcoarse)
    assert np.array_equal(host(bcoarse), O.restrict_3d(oc, f))
    xc = rng.standard_normal(c.num_cells())
    xf = dev(f, ctx)
    mg.prolong_3d_add(ctx, c, g, dev(xc, ctx), xf)
    assert np.array_equal(host(xf), O.prolong_3d_add(oc, xc, f))
    K = random_K(g.num_cells(), 6, 2.0)
    Kc = ctx.empty(c.num_cells())
    ph.coarsen_K(ctx, c, g, dev(K, ctx), Kc)
    assert rel_err(host(Kc), O.coarsen_K(oc, K)) < 1e-14


@pytest.mark.parametrize("mix", MIXES)
def test_rhs_and_guess_match_oracle(ctx, mix):
    g, og = grid_pair(12, 10, 8, dx=0.25)
    bc, obc = bc_pair(mix)
    K = random_K(g.num_cells(), 7)
    rhs = ctx.empty(g.num_cells())
    ph.build_rhs_head(rhs, dev(K, ctx), g, bc, ctx)
    assert rel_err(host(rhs), O.build_rhs_head(og, obc, K)) < 1e-15 or \
        np.array_equal(host(rhs), O.build_rhs_head(og, obc, K))
    h = ctx.empty(g.num_cells())
    ph.init_head_guess(h, g, bc, ctx)
    assert np.array_equal(host(h), O.init_head_guess(og, obc))


@pytest.mark.parametrize("mix", MIXES)
@pytest.mark.parametrize("shape", [(9, 7, 5), (16, 16, 16)])
def test_velocity_matches_oracle(ctx, mix, shape):
    g, og = grid_pair(*shape, dx=0.5)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    rng = np.random.default_rng(8)
    K, h = random_K(n, 9), rng.standard_normal(n)
    v = ph.VelocityField(g, ctx)
    ph.compute_velocity_from_head(v, dev(h, ctx), dev(K, ctx), g, bc, ctx)
    U, V, W = O.velocity_compact(og, obc, h, K)
    for a, b_ in ((v.U, U), (v.V, V), (v.W, W)):
        assert rel_err(host(a), b_) < OP_TOL
    pv = ph.PaddedVelocityField(g, ctx)
    ph.compute_velocity_from_head(pv, dev(h, ctx), dev(K, ctx), g, bc, ctx)
    U, V, W = O.velocity_padded(og, obc, h, K)
    for a, b_ in ((pv.U, U), (pv.V, V), (pv.W, W)):
        assert rel_err(host(a), b_) < OP_TOL


def test_dot_nrm2_and_vector_ops(ctx):
    rng = np.random.default_rng(10)
    for n in (1, 255, 256, 257, 100_003):
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        red = nm.ReductionWorkspace()
        d = nm.blas.dot_host(ctx, dev(x, ctx), dev(y, ctx), red)
        exact = O.dot(x, y)
        assert abs(d - exact) <= 1e-13 * (np.abs(x) @ np.abs(y))
        assert abs(nm.blas.nrm2_host(ctx, dev(x, ctx), red) - O.nrm2(x)) <= 1e-14 * O.nrm2(x)
        yy = dev(y, ctx)
        nm.blas.axpy(ctx, 0.5, dev(x, ctx), yy)
        assert rel_err(host(yy), 0.5 * x + y) < 1e-15
        yy = dev(y, ctx)
        nm.blas.axpby(ctx, 0.5, dev(x, ctx), -2.0, yy)
        assert rel_err(host(yy), 0.5 * x - 2.0 * y) < 1e-15
        yy = dev(y, ctx)
        nm.blas.scal(ctx, yy, 3.0)
        assert np.array_equal(host(yy), 3.0 * y)
        nm.blas.fill(ctx, yy, 7.0)
        assert np.all(host(yy) == 7.0)
        nm.blas.copy(ctx, dev(x, ctx), yy)
        assert np.array_equal(host(yy), x)
    # empty vector: dot is 0 (dot.cu:13-18)
    red = nm.ReductionWorkspace()
    e = torch.empty(0, dtype=torch.float64, device=ctx.device)
    d = red.scalar(ctx)
    nm.blas.dot_device(ctx, e, e, d, red)
    assert host(d)[0] == 0.0


@pytest.mark.parametrize("coarse", [0, 1])
@pytest.mark.parametrize("mix", ["darcy_x", "all_dirichlet", "periodic_yz", "all_neumann"])
def test_vcycle_matches_oracle(ctx, workspace, coarse, mix):
    g, og = grid_pair(16, 16, 16)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    K = random_K(n, 11)
    r = np.random.default_rng(12).standard_normal(n)
    cfg = ph.HeadSolveConfig(mg_levels=3, coarse_solver=ph.CoarseSolver(coarse))
    Kd = dev(K, ctx)
    M = nm.MultigridPreconditioner.build(ctx, workspace, g, bc, cfg, Kd)
    z = ctx.empty(n)
    M.apply(ctx, dev(r, ctx), z)
    p = O.SolveParams(mg_levels=3, coarse_solver=coarse)
    zo = O.v_cycle_precond(og, obc, p, K, r)
    assert rel_err(host(z), zo) < 1e-10, mix


def _solve_pair(ctx, workspace, shape, mix, cfg, params, K, h0=None):
    g, og = grid_pair(*shape)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    h = ctx.zeros(n) if h0 is None else dev(h0, ctx)
    res = ph.solve_head(h, dev(K, ctx), g, bc, cfg, ctx, workspace)
    ho, ro = O.solve_head(og, obc, params, K, h0)
    return host(h), res, ho, ro


def test_solve_head_reference_semantics_32cubed(ctx, workspace):
    """32^3, sigma2 = 1, lambda = 4, V(2,2), 3 levels, 50-GSRB coarsest,
    ||r||/||r0|| < 1e-10 checked every 10 iterations (mode R)."""
    from macroflow3d_amd import kfield
    K = np.exp(kfield.lognormal_lnK(32, 32, 32, 1.0, 4.0, 1234)[0].numpy())
    cfg = ph.HeadSolveConfig(mg_levels=3, cg_rtol=1e-10)
    p = O.mode_r(mg_levels=3, cg_rtol=1e-10)
    h, res, ho, ro = _solve_pair(ctx, workspace, (32, 32, 32), "darcy_x", cfg, p, K)
    assert res.converged and ro.converged
    assert res.num_iterations == ro.num_iterations
    assert res.num_iterations % 10 == 0
    assert rel_l2(h, ho) < 1e-10
    assert abs(res.initial_residual - ro.initial_residual) <= 1e-12 * ro.initial_residual


def test_solve_head_extension_semantics_32cubed(ctx, workspace):
    """Direct coarsest solve, ||r||/||b|| <= 1e-10 every iteration (mode N),
    zero guess; head within 1e-8 relative L2 of an oracle solve taken to 1e-13."""
    from macroflow3d_amd import kfield
    K = np.exp(kfield.lognormal_lnK(32, 32, 32, 1.0, 4.0, 1234)[0].numpy())
    cfg = ph.HeadSolveConfig.fp64_direct(1e-10, mg_levels=3)
    p = O.mode_n(mg_levels=3)
    h, res, ho, ro = _solve_pair(ctx, workspace, (32, 32, 32), "darcy_x", cfg, p, K)
    assert res.converged and res.num_iterations == ro.num_iterations
    assert res.true_residual / res.bnorm <= 1.5e-10
    hdeep, _ = O.solve_head(O.grid(32, 32, 32), O.darcy_bc(), O.mode_n(mg_levels=3, cg_rtol=1e-13),
                            K)
    assert rel_l2(h, hdeep) < 1e-8


@pytest.mark.parametrize("solver", [0, 1])
def test_cg_and_mg_solvers_match_oracle(ctx, workspace, solver):
    K = random_K(16 ** 3, 13, 0.7)
    cfg = ph.HeadSolveConfig(solver_type=ph.HeadSolverType(solver), mg_levels=3, cg_rtol=1e-9,
                             rtol=1e-9, mg_max_cycles=60)
    p = O.SolveParams(solver_type=solver, mg_levels=3, cg_rtol=1e-9, rtol=1e-9, mg_max_cycles=60)
    h, res, ho, ro = _solve_pair(ctx, workspace, (16, 16, 16), "darcy_x", cfg, p, K)
    assert res.converged and ro.converged
    assert res.num_iterations == ro.num_iterations
    assert rel_l2(h, ho) < 1e-9


@pytest.mark.parametrize("mix", ["periodic_x_dirichlet_z", "periodic_yz", "all_dirichlet"])
def test_periodic_and_dirichlet_solves(ctx, workspace, mix):
    """Periodic wraps and all-Dirichlet boundaries through the whole PCG
    (the pinned singular case is covered by the V-cycle test)."""
    K = random_K(16 * 8 * 8, 14, 0.5)
    cfg = ph.HeadSolveConfig(mg_levels=3, cg_rtol=1e-10, coarse_solver=ph.CoarseSolver.DIRECT,
                             stop_norm=ph.StopNorm.RHS_NORM, cg_check_every=1)
    p = O.SolveParams(mg_levels=3, cg_rtol=1e-10, coarse_solver=1, stop_norm=1, cg_check_every=1)
    h, res, ho, ro = _solve_pair(ctx, workspace, (16, 8, 8), mix, cfg, p, K)
    assert res.converged and res.num_iterations == ro.num_iterations
    assert rel_l2(h, ho) < 1e-9


@pytest.mark.parametrize("mix", MIXES)
@pytest.mark.parametrize("shape", [(15, 8, 6), (2, 6, 4), (40, 6, 4), (130, 4, 4)])
def test_solve_operator_variants_match_oracle(ctx, workspace, mix, shape):
    """The FP64 operator of the solve in both forms: single cells (odd nx, a
    one-level direct-preconditioned solve) and cell pairs (even nx; rows of 20
    and 65 pairs cross wave boundaries, 2 x 1 pairs is the smallest), every
    boundary mix including the pinned singular ones. FP64 PCG to 1e-11 against the oracle's solve (faces 2/(1/a + 1/b)
    there, 1/(h_a + h_b) here, h = 1/(2K))."""
    K = random_K(int(np.prod(shape)), 21, 0.8)
    print(f"operator variant {shape} {mix}", flush=True)
    lv = 1 if shape[0] % 2 else 2  # odd sizes cannot be restricted: one (direct) level
    cfg = ph.HeadSolveConfig(mg_levels=lv, cg_rtol=1e-11, coarse_solver=ph.CoarseSolver.DIRECT,
                             stop_norm=ph.StopNorm.RHS_NORM, cg_check_every=1)
    p = O.SolveParams(mg_levels=lv, cg_rtol=1e-11, coarse_solver=1, stop_norm=1, cg_check_every=1)
    h, res, ho, ro = _solve_pair(ctx, workspace, shape, mix, cfg, p, K)
    assert res.converged and ro.converged
    assert abs(res.num_iterations - ro.num_iterations) <= 1
    assert rel_l2(h, ho) < 1e-9
    assert abs(res.initial_residual - ro.initial_residual) <= 1e-12 * ro.initial_residual


def test_misaligned_head_uses_single_cell_operator(ctx, workspace):
    """A head vector that is not 16-byte aligned (a view at an odd offset)
    runs the single-cell operator; the solve matches the aligned one."""
    shape = (32, 16, 16)
    n = int(np.prod(shape))
    K = random_K(n, 22, 0.8)
    g, _ = grid_pair(*shape)
    bc, _ = bc_pair("darcy_x")
    cfg = ph.HeadSolveConfig.fp64_direct(1e-11, mg_levels=3)
    h = ctx.zeros(n)
    r1 = ph.solve_head(h, dev(K, ctx), g, bc, cfg, ctx, workspace)
    buf = ctx.zeros(n + 1)
    hv = buf[1:]
    assert hv.data_ptr() % 16 == 8
    r2 = ph.solve_head(hv, dev(K, ctx), g, bc, cfg, ctx, workspace)
    assert r1.converged and r2.converged
    assert abs(r1.num_iterations - r2.num_iterations) <= 1
    assert rel_l2(host(hv), host(h)) < 1e-10


def test_provided_initial_guess_keeps_caller_head(ctx, workspace):
    """InitGuess.PROVIDED: the solve starts from the caller's h (warm start)."""
    K = random_K(16 ** 3, 16, 0.8)
    g, og = grid_pair(16, 16, 16)
    bc, obc = bc_pair("darcy_x")
    ho, _ = O.solve_head(og, obc, O.mode_n(mg_levels=3, cg_rtol=1e-12), K)
    h0 = ho + 1e-6 * np.random.default_rng(17).standard_normal(ho.size)
    cfg = ph.HeadSolveConfig.fp64_direct(1e-10, mg_levels=3)
    cfg.init_guess = ph.InitGuess.PROVIDED
    h = dev(h0, ctx)
    res = ph.solve_head(h, dev(K, ctx), g, bc, cfg, ctx, workspace)
    cfg0 = ph.HeadSolveConfig.fp64_direct(1e-10, mg_levels=3)
    h2 = ctx.zeros(K.size)
    res0 = ph.solve_head(h2, dev(K, ctx), g, bc, cfg0, ctx, workspace)
    assert res.converged and res0.converged
    assert res.num_iterations < res0.num_iterations
    assert rel_l2(host(h), ho) < 1e-8


@pytest.mark.parametrize("shape,levels,coef", [((16, 16, 16), 3, ph.CoefFormat.BF16),
                                               ((12, 10, 8), 1, ph.CoefFormat.FP32),
                                               ((16, 16, 16), 1, ph.CoefFormat.FP16),
                                               ((24, 16, 16), 2, ph.CoefFormat.BF16)])
def test_mixed_solve_with_natural_finest_level_matches_oracle(ctx, shape, levels, coef):
    """FP32 V-cycles whose finest level is not colour-split (a level of <= 4096
    cells is smoothed in one workgroup; one level is the coarsest itself) still
    give the FP64 operator its h = 1 / (2 K): the head matches the FP64 oracle
    and the true residual, recomputed here from the returned head with the
    oracle's operator, meets the tolerance."""
    g, og = grid_pair(*shape)
    bc, obc = bc_pair("darcy_x")
    n = g.num_cells()
    K = random_K(n, 70 + levels, 1.5)
    ws = ph.FlowWorkspace()
    h = ctx.zeros(n)
    res = ph.solve_head(h, dev(K, ctx), g, bc, ph.HeadSolveConfig.mixed(1e-10, levels, coef), ctx,
                        ws)
    ws.close()
    hg = host(h)
    rhs = O.build_rhs_head(og, obc, K)
    tr = np.linalg.norm(rhs - O.varcoeff_apply(og, obc, False, K, hg)) / np.linalg.norm(rhs)
    # the reference head to 1e-13 from the oracle's 3-level solve (its dense
    # coarsest inverse is O(cells^3): a one-level oracle solve of 16^3 would
    # invert a 4096 x 4096 matrix)
    ho, _ = O.solve_head(og, obc, O.mode_n(mg_levels=3, cg_rtol=1e-13), K)
    assert res.converged and tr <= 1.05e-10, (shape, levels, tr)
    assert rel_l2(hg, ho) < 1e-8, (shape, levels)


@pytest.mark.parametrize("precision", [ph.PrecondPrecision.FP32, ph.PrecondPrecision.FP64])
@pytest.mark.parametrize("n", [16, 32])
def test_workspace_reuse_with_a_new_K_field(ctx, precision, n):
    """Two different K fields solved one after the other on the same workspace:
    the second solve works on the second field's operator (no coefficient of
    the first survives), in the FP64 and the mixed V-cycle."""
    g, og = grid_pair(n, n, n)
    bc, obc = bc_pair("darcy_x")
    cfg = ph.HeadSolveConfig.mixed(1e-10, 3) if precision == ph.PrecondPrecision.FP32 \
        else ph.HeadSolveConfig.fp64_direct(1e-10, 3)
    ws = ph.FlowWorkspace()
    for seed in (81, 82):
        K = random_K(g.num_cells(), seed, 1.5)
        h = ctx.zeros(g.num_cells())
        res = ph.solve_head(h, dev(K, ctx), g, bc, cfg, ctx, ws)
        ho, _ = O.solve_head(og, obc, O.mode_n(mg_levels=3, cg_rtol=1e-13), K)
        assert res.converged and rel_l2(host(h), ho) < 1e-8, (n, seed)
    ws.close()


def test_odd_coarsest_level_and_two_cell_levels(ctx, workspace):
    """48^3 coarsens to 3^3 (odd coarsest) and 8x4x2 has a 2-cell dimension."""
    for shape, levels in (((48, 48, 48), 6), ((8, 4, 2), 4)):
        K = random_K(int(np.prod(shape)), 18, 0.5)
        cfg = ph.HeadSolveConfig.fp64_direct(1e-10, mg_levels=levels)
        p = O.mode_n(mg_levels=levels)
        h, res, ho, ro = _solve_pair(ctx, workspace, shape, "darcy_x", cfg, p, K)
        assert res.converged and res.num_iterations == ro.num_iterations, shape
        assert rel_l2(h, ho) < 1e-9


def test_errors_are_raised(ctx, workspace):
    from macroflow3d_amd._lib import MF3DError
    from macroflow3d_amd.core import BCFace, BCSpec, BCType, Grid3D
    g = Grid3D(8, 8, 8)
    K = ctx.zeros(10)
    h = ctx.zeros(8 ** 3)
    with pytest.raises(RuntimeError):
        ph.solve_head(h, K, g, BCSpec.darcy_x(), ph.HeadSolveConfig(), ctx, workspace)
    bad = BCSpec.darcy_x()
    bad.ymin = BCFace(BCType.Periodic, 0.0)
    with pytest.raises(RuntimeError, match="Periodic"):
        ph.solve_head(h, ctx.zeros(8 ** 3), g, bad, ph.HeadSolveConfig(), ctx, workspace)
    # odd non-coarsest level
    g2 = Grid3D(10, 10, 10)
    with pytest.raises(MF3DError):
        ph.solve_head(ctx.zeros(1000), ctx.zeros(1000) + 1, g2, BCSpec.darcy_x(),
                      ph.HeadSolveConfig(mg_levels=3), ctx, workspace)


def test_non_convergence_is_a_flag(ctx, workspace):
    from macroflow3d_amd.core import BCSpec
    K = random_K(16 ** 3, 19, 1.0)
    g, _ = grid_pair(16, 16, 16)
    cfg = ph.HeadSolveConfig(mg_levels=3, cg_max_iter=3, cg_rtol=1e-14, cg_check_every=1)
    res = ph.solve_head(ctx.zeros(K.size), dev(K, ctx), g, BCSpec.darcy_x(), cfg, ctx, workspace)
    assert not res.converged and res.num_iterations == 3


def test_large_grid_properties_256cubed(ctx):
    """Full-size 256^3, sigma2 = 4, lambda = 8: true residual <= 1e-10 ||b||,
    and Darcy flux conservation (equal total x-flux through every x-plane,
    zero flux through the no-flow faces)."""
    from macroflow3d_amd import kfield
    from macroflow3d_amd.core import BCSpec, Grid3D
    n = 256
    Y, _ = kfield.lognormal_lnK(n, n, n, 4.0, 8.0, 2024, device=ctx.device)
    K = torch.exp(Y)
    g = Grid3D(n, n, n)
    bc = BCSpec.darcy_x()
    ws = ph.FlowWorkspace()
    h = ctx.zeros(g.num_cells())
    res = ph.solve_head(h, K, g, bc, ph.HeadSolveConfig.fp64_direct(1e-10, 6), ctx, ws)
    assert res.converged
    assert res.true_residual / res.bnorm <= 2e-10
    v = ph.VelocityField(g, ctx)
    ph.compute_velocity_from_head(v, h, K, g, bc, ctx)
    U = v.U.view(n, n, n + 1)
    plane_flux = U.sum(dim=(0, 1))
    q = plane_flux.mean()
    assert float((plane_flux - q).abs().max() / q.abs()) < 1e-4
    V = v.V.view(n, n + 1, n)
    assert float(V[:, 0, :].abs().max()) == 0.0 and float(V[:, n, :].abs().max()) == 0.0
    ws.close()


@pytest.mark.parametrize("coef,tol", [(ph.CoefFormat.FP32, 1e-5), (ph.CoefFormat.BF16, 2e-2),
                                      (ph.CoefFormat.FP16, 5e-3)])
def test_mixed_vcycle_close_to_fp64_oracle(ctx, coef, tol):
    """FP32 V-cycle with 32/16-bit face coefficients: the preconditioned vector
    stays within the coefficient format's precision of the FP64 oracle V-cycle."""
    g, og = grid_pair(64, 64, 32)
    bc, obc = bc_pair("darcy_x")
    n = g.num_cells()
    K = random_K(n, 31, 2.0)
    r = np.random.default_rng(32).standard_normal(n)
    cfg = ph.HeadSolveConfig(mg_levels=4, coarse_solver=ph.CoarseSolver.DIRECT,
                             precond_precision=ph.PrecondPrecision.FP32, coef_format=coef)
    ws = ph.FlowWorkspace()
    M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, dev(K, ctx))
    z = ctx.empty(n)
    M.apply(ctx, dev(r, ctx), z)
    zo = O.v_cycle_precond(og, obc, O.SolveParams(mg_levels=4, coarse_solver=1), K, r)
    assert rel_l2(host(z), zo) < tol


@pytest.mark.parametrize("coef", [ph.CoefFormat.BF16, ph.CoefFormat.FP16, ph.CoefFormat.FP32])
@pytest.mark.parametrize("flexible", [False, True])
def test_mixed_precision_solve_meets_fp64_tolerance(ctx, coef, flexible):
    """Mixed-precision MG-PCG reaches ||b - A h|| <= 1e-10 ||b|| (true residual,
    FP64) and the head matches the FP64 oracle to 1e-8 relative L2."""
    from macroflow3d_amd import kfield
    n = 64
    K = np.exp(kfield.lognormal_lnK(n, n, n, 4.0, 8.0, 41)[0].numpy())
    g, og = grid_pair(n, n, n)
    bc, obc = bc_pair("darcy_x")
    ws = ph.FlowWorkspace()
    h = ctx.zeros(g.num_cells())
    res = ph.solve_head(h, dev(K, ctx), g, bc, ph.HeadSolveConfig.mixed(1e-10, 4, coef, flexible),
                        ctx, ws)
    assert res.converged and res.true_residual / res.bnorm <= 1e-10
    hd, _ = O.solve_head(og, obc, O.mode_n(mg_levels=4, cg_rtol=1e-13), K)
    assert rel_l2(host(h), hd) < 1e-8


@pytest.mark.parametrize("mix", ["darcy_x", "all_neumann", "periodic_yz", "all_periodic",
                                 "periodic_x_dirichlet_z"])
@pytest.mark.parametrize("shape,levels,pre_post", [((32, 16, 24), 4, (2, 2)),
                                                   ((64, 32, 32), 4, (1, 3)),
                                                   ((32, 32, 32), 3, (2, 0)),
                                                   ((64, 64, 32), 4, (2, 0)),
                                                   ((20, 16, 16), 2, (2, 2)),
                                                   ((36, 16, 12), 3, (1, 2))])
@pytest.mark.parametrize("coef", [ph.CoefFormat.BF16, ph.CoefFormat.FP16])
def test_colour_split_vcycle_equals_natural_order(ctx, mix, shape, levels, pre_post, coef):
    """Storing the FP32 V-cycle vectors and faces colour by colour (and folding
    the prolongation into the first post-smoothing half-sweep) gives the same
    preconditioned vector as the natural-order kernels; nx a multiple of 8 runs
    the 4-cells-per-thread kernel, other even nx the one-cell kernel."""
    g, _ = grid_pair(*shape)
    bc, _ = bc_pair(mix)
    n = g.num_cells()
    K = random_K(n, 51, 1.5)
    r = np.random.default_rng(52).standard_normal(n)
    cfg = ph.HeadSolveConfig(mg_levels=levels, coarse_solver=ph.CoarseSolver.DIRECT,
                             mg_pre_smooth=pre_post[0], mg_post_smooth=pre_post[1],
                             precond_precision=ph.PrecondPrecision.FP32, coef_format=coef)
    out = []
    for split in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_SPLIT, split)
        M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, dev(K, ctx))
        z = ctx.empty(n)
        M.apply(ctx, dev(r, ctx), z)
        out.append(host(z))
        ws.close()
    assert np.all(np.isfinite(out[0]))
    assert rel_l2(out[0], out[1]) < 1e-6


def test_colour_split_solve_matches_natural_order(ctx):
    """The mixed solve takes the same iterations with either storage order."""
    from macroflow3d_amd import kfield
    n = 64
    K = torch.exp(kfield.lognormal_lnK(n, n, n, 4.0, 8.0, 47)[0]).to(ctx.device)
    g = grid_pair(n, n, n)[0]
    bc = bc_pair("darcy_x")[0]
    out = []
    for split in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_SPLIT, split)
        h = ctx.zeros(g.num_cells())
        res = ph.solve_head(h, K, g, bc, ph.HeadSolveConfig.mixed(1e-10, 4), ctx, ws)
        assert res.converged and res.true_residual / res.bnorm <= 1e-10
        out.append((host(h), res.num_iterations))
        ws.close()
    assert out[0][1] == out[1][1]
    assert rel_l2(out[0][0], out[1][0]) < 1e-10


@pytest.mark.parametrize("block", [64, 128, 512])
def test_sweep_block_size_does_not_change_results(ctx, block):
    """The colour-split sweep's workgroup size is a tuning choice only: the
    preconditioned vector is bitwise the same as with the default 256."""
    g, _ = grid_pair(64, 32, 48)
    bc, _ = bc_pair("periodic_yz")
    n = g.num_cells()
    K = random_K(n, 61, 1.5)
    r = np.random.default_rng(62).standard_normal(n)
    cfg = ph.HeadSolveConfig(mg_levels=4, coarse_solver=ph.CoarseSolver.DIRECT,
                             precond_precision=ph.PrecondPrecision.FP32,
                             coef_format=ph.CoefFormat.BF16)
    out = []
    for tb in (256, block):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, tb)
        M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, dev(K, ctx))
        z = ctx.empty(n)
        M.apply(ctx, dev(r, ctx), z)
        out.append(host(z))
        ws.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("mix", ["darcy_x", "all_neumann", "periodic_yz", "all_periodic"])
@pytest.mark.parametrize("shape,levels,pre_post", [((64, 64, 64), 4, (2, 2)),
                                                   ((64, 32, 48), 4, (1, 3)),
                                                   ((128, 64, 64), 5, (3, 1))])
def test_folded_first_coarse_sweep_matches_separate(ctx, mix, shape, levels, pre_post):
    """The first red half-sweep of a colour-split coarse level, computed by the
    restriction kernel that produces its b (x starts at zero, so a red cell's
    value needs only its b and faces), gives the preconditioned vector of the
    separate zero-start half-sweep."""
    g, _ = grid_pair(*shape)
    bc, _ = bc_pair(mix)
    n = g.num_cells()
    K = random_K(n, 71, 1.2)
    r = np.random.default_rng(72).standard_normal(n)
    cfg = ph.HeadSolveConfig(mg_levels=levels, coarse_solver=ph.CoarseSolver.DIRECT,
                             mg_pre_smooth=pre_post[0], mg_post_smooth=pre_post[1],
                             precond_precision=ph.PrecondPrecision.FP32,
                             coef_format=ph.CoefFormat.BF16)
    print(f"folded first sweep {shape} {mix} {pre_post}", flush=True)
    out = []
    for fold in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_FOLD_FIRST, fold)
        M = nm.MultigridPreconditioner.build(ctx, ws, g, bc, cfg, dev(K, ctx))
        z = ctx.empty(n)
        M.apply(ctx, dev(r, ctx), z)
        out.append(host(z))
        ws.close()
    assert np.all(np.isfinite(out[0]))
    assert rel_l2(out[0], out[1]) < 1e-6


@pytest.mark.parametrize("mix", ["darcy_x", "all_periodic", "all_neumann", "all_dirichlet"])
@pytest.mark.parametrize("shape", [(32, 16, 16), (64, 24, 8)])
def test_fused_residual_update_matches_separate(ctx, mix, shape):
    """The PCG residual update r -= alpha Ap done by the first finest-level
    half-sweep of the FP32 V-cycle (with its (r, r) partials, one per sweep
    workgroup), and the (r, z) partials formed by the last one, give the
    solve of the separate kernels, in every combination: same stopping
    iteration (within one: the sums are added in another order) and the same
    head."""
    g, _ = grid_pair(*shape)
    bc, _ = bc_pair(mix)
    n = g.num_cells()
    K = dev(random_K(n, 73, 1.5), ctx)
    cfg = ph.HeadSolveConfig.mixed(1e-10, mg_levels=3)
    print(f"fused r update / (r, z) {shape} {mix}", flush=True)
    out = {}
    for fr, frz in ((0, 0), (1, 0), (0, 1), (1, 1)):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_FOLD_R, fr)
        ws.set_tuning(ph.FlowWorkspace.TUNE_FOLD_RZ, frz)
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        out[(fr, frz)] = (host(h), res)
        ws.close()
    h0, r0 = out[(0, 0)]
    assert r0.converged
    for key, (h1, r1) in out.items():
        assert r1.converged, key
        assert abs(r1.num_iterations - r0.num_iterations) <= 1, key
        assert rel_l2(h1, h0) < 1e-9, key


@pytest.mark.parametrize("mix", MIXES)
@pytest.mark.parametrize("shape", [(64, 40, 6), (512, 4, 4), (34, 34, 4), (2, 6, 4)])
def test_plane_tile_operator_matches_paired(ctx, mix, shape):
    """The FP64 operator with plane tiles and their y neighbours staged in LDS
    (1024 cells per workgroup: several tiles per plane with a partial last
    one, nx at its 512 limit, rows straddling tiles, a plane smaller than a
    tile) solves as the paired kernel does: the same faces and per-cell sums,
    so the same iterations and the head to rounding of the reductions."""
    g, og = grid_pair(*shape)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    Kh = random_K(n, 74, 1.0)
    K = dev(Kh, ctx)
    print(f"plane-tile operator {shape} {mix}", flush=True)
    cfg = ph.HeadSolveConfig(mg_levels=2, cg_rtol=1e-11, coarse_solver=ph.CoarseSolver.DIRECT,
                             stop_norm=ph.StopNorm.RHS_NORM, cg_check_every=1)
    out = []
    for tile in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_OP_TILE, tile)
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        out.append((host(h), res))
        ws.close()
    (h1, r1), (h0, r0) = out
    assert r1.converged and r0.converged
    assert abs(r1.num_iterations - r0.num_iterations) <= 1  # reduction order only
    assert abs(r1.initial_residual - r0.initial_residual) <= 1e-13 * r0.initial_residual
    assert rel_l2(h1, h0) < 1e-10
    p = O.SolveParams(mg_levels=2, cg_rtol=1e-11, coarse_solver=1, stop_norm=1, cg_check_every=1)
    ho, ro = O.solve_head(og, obc, p, Kh)
    assert rel_l2(h1, ho) < 1e-9


@pytest.mark.parametrize("tb,pre,post", [(64, 2, 2), (512, 2, 2), (128, 3, 1), (256, 2, 1),
                                         (256, 2, 3)])
def test_fused_pcg_steps_across_sweep_blocks_and_smoothing(ctx, tb, pre, post):
    """The r update in the first finest half-sweep and the (r, z) partials in
    the last one write one partial per sweep workgroup: every workgroup size
    (64..512 threads) and smoothing count (the first half-sweep red with x = 0,
    the last black after the prolongation's half-sweep) solves as the separate
    kernels do. (With a single pre-smoothing iteration this red-black V-cycle
    is too far from symmetric for CG on this field: V(1, k) stagnates near
    1e-4 in FP64 too, fused or not, so the cases keep pre >= 2.)"""
    g, _ = grid_pair(64, 32, 48)
    bc, _ = bc_pair("periodic_yz")
    n = g.num_cells()
    K = dev(random_K(n, 75, 1.5), ctx)
    cfg = ph.HeadSolveConfig.mixed(1e-10, mg_levels=4)
    cfg.mg_pre_smooth, cfg.mg_post_smooth = pre, post
    print(f"fused PCG steps tb={tb} V({pre},{post})", flush=True)
    out = []
    for fold in (1, 0):
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, tb)
        ws.set_tuning(ph.FlowWorkspace.TUNE_FOLD_R, fold)
        ws.set_tuning(ph.FlowWorkspace.TUNE_FOLD_RZ, fold)
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        out.append((host(h), res))
        ws.close()
    (h1, r1), (h0, r0) = out
    assert r1.converged and r0.converged
    assert abs(r1.num_iterations - r0.num_iterations) <= 1
    assert rel_l2(h1, h0) < 1e-9


@pytest.mark.parametrize("mix", ["darcy_x", "periodic_yz", "all_dirichlet"])
def test_reported_true_residual_is_that_of_the_returned_head(ctx, mix):
    """result.true_residual is ||b - A h|| of the head the solve returns,
    evaluated again by the oracle: for solves that stop on a passing
    true-residual check and report that check's sums (the mixed V-cycle, FP64
    stopping on ||b||, a batch of two) and for one that evaluates it after the
    solve (the reference's FP64 semantics, stopping on ||r0||). A residual near
    1e-10 ||b|| keeps only a few digits through the cancellation in b - A h,
    hence 1e-3 relative; the residual of the iterate one step earlier differs
    by far more."""
    g, og = grid_pair(32, 32, 32)
    bc, obc = bc_pair(mix)
    n = g.num_cells()
    pin = pin_for(mix)
    Ks = [random_K(n, 91, 1.5), random_K(n, 92, 1.5)]

    def oracle_norm(h, K):
        rhs = O.build_rhs_head(og, obc, K)
        return np.linalg.norm(O.compute_residual_3d(og, obc, pin, h, rhs, K))

    print(f"reported true residual {mix}", flush=True)
    cfgs = {"mixed": ph.HeadSolveConfig.mixed(1e-10, 4),
            "fp64_bnorm": ph.HeadSolveConfig.fp64_direct(1e-10, 4),
            "fp64_r0": ph.HeadSolveConfig(mg_levels=4)}
    for name, cfg in cfgs.items():
        ws = ph.FlowWorkspace()
        h = ctx.zeros(n)
        res = ph.solve_head(h, dev(Ks[0], ctx), g, bc, cfg, ctx, ws)
        ws.close()
        want = oracle_norm(host(h), Ks[0])
        assert res.converged, name
        assert abs(res.true_residual - want) <= 1e-3 * want, (name, res.true_residual, want)
    hs = [ctx.zeros(n) for _ in Ks]
    wss = [ph.FlowWorkspace() for _ in Ks]
    rs = ph.solve_head_batched(hs, [dev(K, ctx) for K in Ks], [g, g], bc,
                               ph.HeadSolveConfig.mixed(1e-10, 4), ctx, wss)
    for i, (r, hb, K) in enumerate(zip(rs, hs, Ks)):
        want = oracle_norm(host(hb), K)
        assert r.converged and abs(r.true_residual - want) <= 1e-3 * want, (i, r.true_residual, want)
    for w in wss:
        w.close()
