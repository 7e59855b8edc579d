"""Pins the CPU FP64 oracle (no GPU).

The reference path is CUDA-only and ships no golden vectors, so the oracle is
pinned by what the mathematics fixes, each check chosen so that a plausible
mistake (wrong sign, index, harmonic mean, BC term, colour order) fails:
  * closed forms: homogeneous K (linear head), K layered across the flow
    (series law), K layered along the flow (parallel layers);
  * per-cell Darcy flux balance at convergence, computed by the velocity
    routine (a separate code path with the 2ab/(a+b) harmonic mean);
  * operator symmetry and definiteness, smoother/operator consistency;
  * transfer identities R(1) = 1, R(P c) = c, geometric-mean coarsening;
  * red-black Gauss-Seidel on a 1-cell-thick slab equals textbook 1-D
    red-black Gauss-Seidel; GSRB and the solvers converge to the dense solve;
  * the reference's stopping-rule quirks (checks every 10 iterations, early
    exit when the initial residual vanishes).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import BC_MIXES, bc_pair, random_K, rel_err

D, N, P = 0, 1, 2


def dense_matrix(g, b, pin, K):
    """Matrix of the oracle operator by applying it to unit vectors."""
    n = O.ncells(g)
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = O.varcoeff_apply(g, b, pin, K, e)
        e[j] = 0.0
    return A


def test_homogeneous_closed_form():
    g = O.grid(16, 8, 8, 0.5)
    b = O.darcy_bc(2.0, -1.0)
    K = np.full(O.ncells(g), 3.0)
    h_lin = O.init_head_guess(g, b)
    x = (np.arange(16) + 0.5) / 16
    assert np.allclose(h_lin.reshape(8, 8, 16), 2.0 - 3.0 * x, atol=1e-15)
    rhs = O.build_rhs_head(g, b, K)
    # the linear profile is the exact discrete solution: A h = b
    assert rel_err(O.varcoeff_apply(g, b, False, K, h_lin), rhs) < 1e-13
    h, r = O.solve_head(g, b, O.mode_n(mg_levels=3, cg_rtol=1e-12), K)
    assert rel_err(h, h_lin) < 1e-11
    U, V, W = O.velocity_compact(g, b, h, K)
    # uniform Darcy flux K (h_W - h_E) / (nx dx), no transverse flow
    assert np.allclose(U, 3.0 * 3.0 / (16 * 0.5), rtol=1e-10)
    assert np.abs(V).max() < 1e-10 and np.abs(W).max() < 1e-10


def test_reference_quirks_early_exit_and_check_cadence():
    g = O.grid(16, 8, 8)
    b = O.darcy_bc()
    # the reference's linear guess is exact for homogeneous K: ||r0|| < 1e-15
    _, r = O.solve_head(g, b, O.mode_r(mg_levels=3, cg_rtol=1e-10), np.ones(O.ncells(g)))
    assert r.converged and r.num_iterations == 0
    # convergence is only tested every cg_check_every = 10 iterations
    _, r = O.solve_head(g, b, O.mode_r(mg_levels=3, cg_rtol=1e-10), random_K(O.ncells(g), 1))
    assert r.converged and r.num_iterations % 10 == 0 and r.final_residual / r.initial_residual < 1e-10


@pytest.mark.parametrize("dx", [1.0, 0.5])
def test_series_layers_closed_form(dx):
    """K depends on x only: series law, U = (h_W - h_E) / (dx sum 1/K_i) on every x-face."""
    nx, ny, nz = 24, 4, 4
    g = O.grid(nx, ny, nz, dx)
    b = O.darcy_bc(1.0, 0.0)
    k1 = np.exp(np.random.default_rng(0).standard_normal(nx))
    K = np.tile(k1, ny * nz)
    q = 1.0 / (dx * np.sum(1.0 / k1))
    # head from the series resistances (cell-centre to cell-centre)
    h1 = np.empty(nx)
    h1[0] = 1.0 - q * dx / (2 * k1[0])
    for i in range(nx - 1):
        h1[i + 1] = h1[i] - q * dx * 0.5 * (1 / k1[i] + 1 / k1[i + 1])
    assert abs(h1[-1] - q * dx / (2 * k1[-1])) < 1e-12
    rhs = O.build_rhs_head(g, b, K)
    assert rel_err(O.varcoeff_apply(g, b, False, K, np.tile(h1, ny * nz)), rhs) < 1e-11
    h, _ = O.solve_head(g, b, O.mode_n(mg_levels=2, cg_rtol=1e-13), K)
    assert rel_err(h, np.tile(h1, ny * nz)) < 1e-10
    U, V, W = O.velocity_compact(g, b, h, K)
    assert np.allclose(U, q, rtol=1e-9)


def test_parallel_layers_closed_form():
    """K depends on y only, no-flow sides: head linear in x, U_j = K_j dh / L."""
    nx, ny, nz = 16, 8, 4
    g = O.grid(nx, ny, nz)
    b = O.darcy_bc(1.0, 0.0)
    kj = np.exp(np.random.default_rng(1).standard_normal(ny))
    K = np.repeat(np.tile(kj, nz), nx)
    h, _ = O.solve_head(g, b, O.mode_n(mg_levels=2, cg_rtol=1e-13), K)
    assert rel_err(h, O.init_head_guess(g, b)) < 1e-10
    U, V, W = O.velocity_compact(g, b, h, K)
    U = U.reshape(nz, ny, nx + 1)
    assert np.allclose(U, (kj / nx)[None, :, None], rtol=1e-9)
    assert np.abs(V).max() < 1e-10 and np.abs(W).max() < 1e-10


@pytest.mark.parametrize("mix", ["darcy_x", "all_dirichlet", "periodic_yz", "periodic_x_dirichlet_z"])
def test_flux_balance_at_convergence(mix):
    """At the solution every cell's net Darcy flux (from the velocity routine) vanishes."""
    g = O.grid(16, 12, 8, 0.5)
    _, b = bc_pair(mix)
    K = random_K(O.ncells(g), 2, 1.5)
    h, r = O.solve_head(g, b, O.mode_n(mg_levels=2, cg_rtol=1e-13), K)
    U, V, W = O.velocity_compact(g, b, h, K)
    U = U.reshape(8, 12, 17)
    V = V.reshape(8, 13, 16)
    W = W.reshape(9, 12, 16)
    div = (U[:, :, 1:] - U[:, :, :-1]) + (V[:, 1:, :] - V[:, :-1, :]) + (W[1:] - W[:-1])
    assert np.abs(div).max() < 1e-9 * np.abs(U).max()


@pytest.mark.parametrize("mix", list(BC_MIXES))
def test_operator_symmetric_negative_definite(mix):
    g = O.grid(5, 4, 6, 0.7)
    _, b = bc_pair(mix)
    pin = mix in ("all_neumann", "all_periodic")
    K = random_K(O.ncells(g), 3, 1.0)
    A = dense_matrix(g, b, pin, K)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 0
    # every row is a 7-point stencil at most
    assert ((np.abs(A) > 0).sum(axis=1) <= 7).all()


@pytest.mark.parametrize("mix", ["darcy_x", "all_dirichlet", "periodic_yz", "periodic_x_dirichlet_z"])
def test_residual_matches_operator_without_pin(mix):
    g = O.grid(6, 5, 4, 2.0)
    _, b = bc_pair(mix)
    K = random_K(O.ncells(g), 4)
    rng = np.random.default_rng(5)
    x, f = rng.standard_normal(O.ncells(g)), rng.standard_normal(O.ncells(g))
    r = O.compute_residual_3d(g, b, False, x, f, K)
    assert rel_err(r, f - O.varcoeff_apply(g, b, False, K, x)) < 1e-12


def test_transfer_identities():
    c = O.grid(4, 3, 5, 2.0)
    f_shape = 8 * 6 * 10
    assert np.allclose(O.restrict_3d(c, np.ones(f_shape)), 1.0, atol=0, rtol=1e-15)
    xc = np.random.default_rng(6).standard_normal(O.ncells(c))
    fine = O.prolong_3d_add(c, xc, np.zeros(f_shape))
    assert rel_err(O.restrict_3d(c, fine), xc) < 1e-15  # (8 c) / 8 up to rounding
    fine3 = fine.reshape(10, 6, 8)
    assert fine3[3, 5, 7] == xc.reshape(5, 3, 4)[1, 2, 3]


def test_coarsening_is_geometric_mean():
    c = O.grid(4, 4, 4, 2.0)
    assert np.allclose(O.coarsen_K(c, np.full(512, 7.5)), 7.5, rtol=1e-15)
    z, y, x = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    Y = 0.3 * x - 0.2 * y + 0.1 * z
    Kc = O.coarsen_K(c, np.exp(Y).ravel())
    Z, Yc, X = np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij")
    expect = np.exp(0.3 * (2 * X + 0.5) - 0.2 * (2 * Yc + 0.5) + 0.1 * (2 * Z + 0.5))
    assert rel_err(Kc, expect.ravel()) < 1e-14


def test_gsrb_on_slab_is_1d_red_black():
    """ny = nz = 1 with no-flow sides: 3-D GSRB is 1-D red-black Gauss-Seidel."""
    nx = 11
    g = O.grid(nx, 1, 1, 0.25)
    b = O.darcy_bc()
    rng = np.random.default_rng(7)
    K, x0, f = np.exp(rng.standard_normal(nx)), rng.standard_normal(nx), rng.standard_normal(nx)
    x = x0.copy()
    dx2 = 0.0625
    kf = 2.0 / (1.0 / K[:-1] + 1.0 / K[1:])  # interior face conductances
    for colour in (0, 1):
        for i in range(colour, nx, 2):
            w = kf[i - 1] if i > 0 else 2.0 * K[0]          # Dirichlet: 2 K, ghost 0
            e = kf[i] if i < nx - 1 else 2.0 * K[-1]
            xw = x[i - 1] if i > 0 else 0.0
            xe = x[i + 1] if i < nx - 1 else 0.0
            x[i] = (w * xw + e * xe - f[i] * dx2) / (w + e)
    assert rel_err(O.gsrb_smooth_3d(g, b, False, x0, f, K, 1), x) < 1e-14


@pytest.mark.parametrize("mix", ["darcy_x", "all_neumann", "periodic_x_dirichlet_z"])
def test_gsrb_and_solvers_converge_to_dense_solution(mix):
    g = O.grid(6, 4, 4)
    _, b = bc_pair(mix)
    pin = mix == "all_neumann"
    n = O.ncells(g)
    K = random_K(n, 8, 0.8)
    f = np.random.default_rng(9).standard_normal(n)
    A = dense_matrix(g, b, pin, K)
    x_exact = np.linalg.solve(A, f)
    x = O.gsrb_smooth_3d(g, b, pin, np.zeros(n), f, K, 3000)
    # the smoother's and the operator's pin conventions differ (aC with or
    # without boundary terms), so compare GSRB with its own fixed point only
    # when no pin is involved
    if not pin:
        assert rel_err(x, x_exact) < 1e-9
    Ainv = O.dense_inverse(g, b, pin, K)
    r = O.compute_residual_3d(g, b, pin, Ainv @ f, f, K)
    assert np.abs(r).max() < 1e-10 * np.abs(f).max()


def test_pcg_mg_cg_agree_with_dense_solve():
    g = O.grid(8, 8, 8)
    b = O.darcy_bc(1.0, 0.0)
    n = O.ncells(g)
    K = random_K(n, 10, 1.0)
    rhs = O.build_rhs_head(g, b, K)
    h_exact = np.linalg.solve(dense_matrix(g, b, False, K), rhs)
    for params in (O.mode_n(mg_levels=3, cg_rtol=1e-13),
                   O.SolveParams(solver_type=0, cg_rtol=1e-13, stop_norm=1, cg_check_every=1),
                   O.SolveParams(solver_type=1, mg_levels=3, rtol=1e-13, mg_max_cycles=200),
                   O.mode_r(mg_levels=3, cg_rtol=1e-12)):
        h, r = O.solve_head(g, b, params, K)
        assert r.converged, params
        assert rel_err(h, h_exact) < 1e-9, params


def test_padded_layout_matches_compact():
    g = O.grid(5, 4, 3)
    b = O.darcy_bc()
    rng = np.random.default_rng(11)
    K, h = random_K(60, 12), rng.standard_normal(60)
    U, V, W = O.velocity_compact(g, b, h, K)
    Up, Vp, Wp = O.velocity_padded(g, b, h, K)
    Up, Vp, Wp = (a.reshape(4, 5, 6) for a in (Up, Vp, Wp))
    assert np.array_equal(Up[:3, :4, :6], U.reshape(3, 4, 6))
    assert np.array_equal(Vp[:3, :5, :5], V.reshape(3, 5, 5))
    assert np.array_equal(Wp[:4, :4, :5], W.reshape(4, 4, 5))
    assert (Up[3] == 0).all() and (Up[:, 4] == 0).all() and (Vp[:, :, 5] == 0).all()


def test_reductions_extended_precision():
    x = np.array([1e16, 1.0, -1e16, 1.0])
    assert O.dot(x, np.ones(4)) == 2.0  # naive double summation gives 1.0 or 0.0
    assert O.nrm2(np.array([3.0, 4.0])) == 5.0
