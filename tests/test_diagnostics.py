"""Velocity diagnostics (divergence, |curl|, helicity; velocity_diagnostics.cu)
and field statistics (compute_field_stats, stochastic.cu:369-414).

CPU: the oracle is pinned by closed forms on linear fields (every derivative
term distinct, anisotropic spacing so that a swapped axis or spacing fails),
by the divergence of the Darcy velocity of any head being the head system's
residual b - A h (a separate code path: the operator and the velocity
routine), and by numpy for the statistics.
GPU: the HIP kernel against the oracle through the C ABI on ragged shapes,
the fused statistics against numpy on the kernel's own output, and at 256^3
the divergence of a converged solve's velocity against the solve's residual.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from helpers import bc_pair, random_K

# ---- oracle pinning (CPU) ---------------------------------------------------


def padded_linear(nx, ny, nz, dx, dy, dz, fu, fv, fw):
    """Padded components with U(i,j,k) = fu(x_face_i, y_c, z_c) etc. on the
    physical faces (padding left at zero)."""
    U = np.zeros((nz + 1, ny + 1, nx + 1))
    V = np.zeros_like(U)
    W = np.zeros_like(U)
    i = np.arange(nx + 1)[None, None, :]
    j = np.arange(ny + 1)[None, :, None]
    k = np.arange(nz + 1)[:, None, None]
    xc, yc, zc = (i + 0.5) * dx, (j + 0.5) * dy, (k + 0.5) * dz
    xf, yf, zf = i * dx, j * dy, k * dz
    U[:nz, :ny, :] = np.broadcast_to(fu(xf, yc, zc), U.shape)[:nz, :ny, :]
    V[:nz, :, :nx] = np.broadcast_to(fv(xc, yf, zc), U.shape)[:nz, :, :nx]
    W[:, :ny, :nx] = np.broadcast_to(fw(xc, yc, zf), U.shape)[:, :ny, :nx]
    return U.ravel(), V.ravel(), W.ravel()


def interior_mask(nx, ny, nz):
    m = np.zeros((nz, ny, nx), dtype=bool)
    if min(nx, ny, nz) >= 3:
        m[1:-1, 1:-1, 1:-1] = True
    return m


def test_oracle_diagnostics_rotational_closed_form():
    """u = (a y, c z, b x): div 0, curl (-c, -b, -a), helicity -(a c y + b c z + a b x)."""
    nx, ny, nz, dx, dy, dz = 7, 6, 5, 0.5, 0.7, 1.3
    a, b, c = 1.7, -0.6, 2.3
    U, V, W = padded_linear(nx, ny, nz, dx, dy, dz, lambda x, y, z: a * y + 0 * x * z,
                            lambda x, y, z: c * z + 0 * x * y, lambda x, y, z: b * x + 0 * y * z)
    div, om, hel = (f.reshape(nz, ny, nx)
                    for f in O.velocity_diagnostics(O.grid(nx, ny, nz, dx, dy, dz), U, V, W))
    m = interior_mask(nx, ny, nz)
    assert np.abs(div).max() < 1e-12
    assert np.allclose(om[m], np.sqrt(a * a + b * b + c * c), rtol=1e-12)
    xc = ((np.arange(nx) + 0.5) * dx)[None, None, :]
    yc = ((np.arange(ny) + 0.5) * dy)[None, :, None]
    zc = ((np.arange(nz) + 0.5) * dz)[:, None, None]
    h_exact = np.broadcast_to(-(a * c * yc + b * c * zc + a * b * xc), (nz, ny, nx))
    assert np.allclose(hel[m], h_exact[m], rtol=1e-12, atol=1e-12)
    assert (om[~m] == 0).all() and (hel[~m] == 0).all()


def test_oracle_diagnostics_gradient_field():
    """u = (p x, q y, r z) = grad of a quadratic: div = p + q + r, no curl."""
    nx, ny, nz, dx, dy, dz = 6, 5, 7, 0.9, 0.4, 1.1
    p, q, r = 0.8, -1.9, 0.3
    U, V, W = padded_linear(nx, ny, nz, dx, dy, dz, lambda x, y, z: p * x + 0 * y * z,
                            lambda x, y, z: q * y + 0 * x * z, lambda x, y, z: r * z + 0 * x * y)
    div, om, hel = O.velocity_diagnostics(O.grid(nx, ny, nz, dx, dy, dz), U, V, W)
    assert np.allclose(div, p + q + r, rtol=0, atol=1e-12)
    assert np.abs(om).max() < 1e-12 and np.abs(hel).max() < 1e-12


def test_oracle_diagnostics_no_curl_below_three():
    g = O.grid(2, 5, 4)
    rng = np.random.default_rng(3)
    n = 3 * 6 * 5
    U, V, W = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    div, om, hel = O.velocity_diagnostics(g, U, V, W)
    Up, Vp, Wp = (a.reshape(5, 6, 3) for a in (U, V, W))
    ref = (Up[:4, :5, 1:] - Up[:4, :5, :-1]) + (Vp[:4, 1:, :2] - Vp[:4, :-1, :2]) + \
        (Wp[1:, :5, :2] - Wp[:-1, :5, :2])
    assert np.allclose(div.reshape(4, 5, 2), ref, rtol=0, atol=1e-13)
    assert (om == 0).all() and (hel == 0).all()


@pytest.mark.parametrize("mix", ["darcy_x", "all_dirichlet", "periodic_yz",
                                 "periodic_x_dirichlet_z"])
def test_oracle_divergence_is_head_residual(mix):
    """For any head h, the divergence of its Darcy velocity is b - A h."""
    g = O.grid(9, 7, 6, 0.5)
    _, ob = bc_pair(mix)
    n = O.ncells(g)
    K = random_K(n, 5, 1.2)
    h = np.random.default_rng(6).standard_normal(n)
    rhs = O.build_rhs_head(g, ob, K)
    r = O.compute_residual_3d(g, ob, False, h, rhs, K)
    U, V, W = O.velocity_padded(g, ob, h, K)
    div, _, _ = O.velocity_diagnostics(g, U, V, W)
    assert np.abs(div - r).max() <= 1e-12 * np.abs(r).max()


def test_oracle_field_stats():
    x = np.random.default_rng(8).standard_normal(1001) * 3 + 0.5
    lo, hi, mean = O.field_stats(x)
    assert lo == x.min() and hi == x.max()
    assert abs(mean - x.mean()) < 1e-14
    assert O.field_stats(np.zeros(0)) == (0.0, 0.0, 0.0)


# ---- GPU parity (C ABI) -----------------------------------------------------

def _dev(a, ctx):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(ctx.device)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 5, 4), (3, 3, 3), (37, 19, 11), (64, 4, 17),
                                   (65, 9, 40), (32, 32, 32), (130, 70, 36)])
def test_diagnostics_match_oracle(ctx, shape):
    from macroflow3d_amd import physics as ph
    nx, ny, nz = shape
    print(f"velocity diagnostics {shape}", flush=True)
    g = ph.Grid3D(nx, ny, nz, 0.5, 0.7, 1.3)
    og = O.grid(nx, ny, nz, 0.5, 0.7, 1.3)
    vel = ph.PaddedVelocityField(g, ctx)
    s = vel.field_size()
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz)
    comps = [rng.standard_normal(s) for _ in range(3)]
    vel.U, vel.V, vel.W = (_dev(a, ctx) for a in comps)
    diag = ph.VelocityDiagnostics(g, ctx)
    ph.compute_velocity_diagnostics(vel, diag, g, ctx)
    ref = O.velocity_diagnostics(og, *comps)
    got = [t.cpu().numpy() for t in (diag.divergence, diag.vorticity_mag, diag.helicity)]
    for name, a, b in zip(("div", "omag", "hel"), got, ref):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), (shape, name)
    # the statistics fused into the kernel, against numpy on the kernel's output
    st = diag.stats.numpy()
    for q, a in enumerate(got):
        assert st[q, 0] == a.min() and st[q, 1] == a.max(), (shape, q)
        assert abs(st[q, 2] - a.mean()) <= 1e-13 * max(1.0, np.abs(a).max()), (shape, q)
    # compute_field_stats on the same fields
    for a, t in zip(got, (diag.divergence, diag.vorticity_mag, diag.helicity)):
        lo, hi, mean = ph.compute_field_stats(t, ctx)
        assert lo == a.min() and hi == a.max() and abs(mean - a.mean()) <= 1e-13 * max(
            1.0, np.abs(a).max())


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 1 << 20])
def test_field_stats_sizes(ctx, n):
    from macroflow3d_amd import physics as ph
    x = np.random.default_rng(n).standard_normal(n) * 2 - 0.3
    lo, hi, mean = ph.compute_field_stats(_dev(x, ctx), ctx)
    assert lo == x.min() and hi == x.max()
    assert abs(mean - x.mean()) <= 1e-14 * max(1.0, np.abs(x).max())
    assert ph.compute_field_stats(_dev(np.zeros(0), ctx), ctx) == (0.0, 0.0, 0.0)


@pytest.mark.gpu
def test_diagnostics_abi_errors(ctx):
    from macroflow3d_amd import _lib
    from macroflow3d_amd import physics as ph
    g = ph.Grid3D(4, 4, 4)
    buf = ctx.zeros(125)
    out = ctx.zeros(64)
    p = C.c_void_p(buf.data_ptr())
    q = C.c_void_p(out.data_ptr())
    lib = _lib.lib()
    # padded buffers too short
    st = lib.mf3d_velocity_diagnostics(None, ctx.handle, C.byref(g.c()), p, p, p, 124, q, q, q,
                                       64, None)
    assert _lib.STATUS_NAMES[st] == "SIZE_MISMATCH"
    # stats without a workspace
    st = lib.mf3d_velocity_diagnostics(None, ctx.handle, C.byref(g.c()), p, p, p, 125, q, q, q,
                                       64, q)
    assert _lib.STATUS_NAMES[st] == "INVALID_ARGUMENT"


@pytest.mark.gpu
def test_divergence_of_solved_velocity_is_residual_256(ctx, workspace):
    """At the full 256^3 size: after an FP64 solve to 1e-10, the divergence of
    the padded velocity equals the residual b - A h (size-independent
    identity), and its statistics bound the mass imbalance."""
    from macroflow3d_amd import kfield
    from macroflow3d_amd import multigrid as mg
    from macroflow3d_amd import physics as ph
    from macroflow3d_amd.core import PinSpec
    n = 256
    print("divergence vs residual 256^3", flush=True)
    g = ph.Grid3D(n, n, n, 1.0)
    bc, _ = bc_pair("darcy_x")
    K, _ = kfield.lognormal_K(n, n, n, 4.0, 8.0, 7, device=ctx.device)
    K = K.to(torch.float64).contiguous()
    h = ctx.zeros(g.num_cells())
    res = ph.solve_head(h, K, g, bc, ph.HeadSolveConfig.fp64_direct(1e-10, 6), ctx, workspace)
    assert res.converged
    vel = ph.PaddedVelocityField(g, ctx)
    ph.compute_velocity_from_head(vel, h, K, g, bc, ctx)
    diag = ph.VelocityDiagnostics(g, ctx)
    ph.compute_velocity_diagnostics(vel, diag, g, ctx)
    rhs = ctx.zeros(g.num_cells())
    ph.build_rhs_head(rhs, K, g, bc, ctx)
    r = ctx.zeros(g.num_cells())
    mg.compute_residual_3d(ctx, g, h, rhs, K, r, bc, PinSpec(False))
    torch.cuda.synchronize()
    d = diag.divergence
    scale = float(torch.abs(r).max())
    assert float(torch.abs(d - r).max()) <= 1e-9 * max(scale, 1e-300) + 1e-12
    # mass balance: the solve's true residual bounds the divergence
    assert float(torch.linalg.vector_norm(d)) <= 2e-10 * res.bnorm
    lo, hi, mean = (float(v) for v in diag.stats[0])
    assert lo == float(d.min()) and hi == float(d.max())
