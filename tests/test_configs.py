"""The workloads at their full sizes, on the GPU.

* 256^3 plain FP64 MG-PCG (sigma2_lnK = 4, correlation length 8, 6 levels)
  against the CPU oracle on the same K: head relative L2 and the velocity
  checksums (sum, sum of squares, max |.| per component) the reference prints
  (velocity_from_head.cu:942-962), for the FP64 and the mixed-precision head.
* 512^3 mixed precision (sigma2 = 4, correlation length 16, 7 levels) against
  the plain FP64 GPU solve of the same K, with Darcy flux conservation.
* The uneven Monte-Carlo batch (32 x 48^3, 32 x 64^3, 32 x 96^3) through the
  batched solve: every realisation converged on its true residual, and one
  realisation of each size equal to its single solve.

The oracle reference head is solved to ||b - A h|| <= 1e-13 ||b|| (Mode N:
direct coarsest solve), so its own error is far below the 1e-8 bar.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from macroflow3d_amd import ensemble
from macroflow3d_amd import physics as ph
from macroflow3d_amd.core import BCSpec, Grid3D
from macroflow3d_amd.kfield import lognormal_lnK

pytestmark = pytest.mark.gpu


def _rel(a, b) -> float:
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def checksums(U: np.ndarray) -> np.ndarray:
    """(sum, sum of squares, max |.|) of one velocity component, in FP64."""
    U = np.asarray(U, dtype=np.float64)
    return np.array([U.sum(), np.dot(U, U), np.abs(U).max()])


def _solve(ctx, K, g, bc, cfg):
    ws = ph.FlowWorkspace()
    h = ctx.zeros(g.num_cells())
    res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
    ws.close()
    v = ph.VelocityField(g, ctx)
    ph.compute_velocity_from_head(v, h, K, g, bc, ctx)
    return h, v, res


def _checksum_dev(v, ref_uvw):
    """Largest deviation of the (sum, sum of squares, max |.|) checksums of the
    three components; the sum is taken on the scale of the component's L1 mass
    (the V and W sums are ~0: no net transverse flux)."""
    dev = np.zeros(3)
    for comp, ref in zip(("U", "V", "W"), ref_uvw):
        cg, co = checksums(getattr(v, comp).cpu().numpy()), checksums(ref)
        scale = np.array([np.abs(ref).sum(), co[1], co[2]])
        dev = np.maximum(dev, np.abs(cg - co) / scale)
    return dev


def test_config2_fp64_and_mixed_heads_match_cpu_oracle_256cubed(ctx):
    n = 256
    print(f"config 2: {n}^3 K field", flush=True)
    Y, _ = lognormal_lnK(n, n, n, 4.0, 8.0, 2024, device=ctx.device)
    K = torch.exp(Y)
    del Y
    g, bc = Grid3D(n, n, n), BCSpec.darcy_x()
    cfg64 = ph.HeadSolveConfig.fp64_direct(1e-10, 6)
    cfg64.init_guess = ph.InitGuess.REFERENCE
    cfgmx = ph.HeadSolveConfig.mixed(1e-10, 6)
    cfgmx.init_guess = ph.InitGuess.REFERENCE
    h64, v64, r64 = _solve(ctx, K, g, bc, cfg64)
    hmx, vmx, rmx = _solve(ctx, K, g, bc, cfgmx)
    assert r64.converged and r64.true_residual <= 1e-10 * r64.bnorm
    assert rmx.converged and rmx.true_residual <= 1e-10 * rmx.bnorm
    print(f"config 2: GPU solves done ({r64.num_iterations} / {rmx.num_iterations} iterations); "
          f"CPU oracle solves to 1e-10 and 1e-13", flush=True)
    Kh = K.cpu().numpy()
    og, obc = O.grid(n, n, n), O.darcy_bc()
    # the same algorithm as the GPU FP64 solve: linear head, direct 8^3
    # coarsest solve, ||b - A h|| <= 1e-10 ||b|| checked every iteration
    h10, r10 = O.solve_head(og, obc, O.mode_n(mg_levels=6, cg_rtol=1e-10, init_guess=0), Kh)
    assert r10.converged
    v10 = O.velocity_compact(og, obc, h10, Kh)
    e_same = _rel(h64, torch.as_tensor(h10).to(ctx.device))
    dev_same = _checksum_dev(v64, v10)
    print(f"config 2: GPU FP64 vs oracle, same algorithm: {r64.num_iterations} vs "
          f"{r10.num_iterations} iterations, head rel L2 {e_same:.3e}, velocity checksums "
          f"(sum, sum of squares, max) max rel dev {dev_same}", flush=True)
    assert r64.num_iterations == r10.num_iterations
    assert e_same < 1e-10
    assert np.all(dev_same < 1e-8), dev_same
    # accuracy reference: the oracle solved to 1e-13
    ho, ro = O.solve_head(og, obc, O.mode_n(mg_levels=6, cg_rtol=1e-13), Kh)
    assert ro.converged
    vo = O.velocity_compact(og, obc, ho, Kh)
    hot = torch.as_tensor(ho).to(ctx.device)
    e64, emx = _rel(h64, hot), _rel(hmx, hot)
    d64, dmx = _checksum_dev(v64, vo), _checksum_dev(vmx, vo)
    print(f"config 2: vs the 1e-13 oracle head: fp64 {e64:.3e}, mixed {emx:.3e}; velocity "
          f"checksum devs fp64 {d64}, mixed {dmx}", flush=True)
    assert e64 < 1e-8 and emx < 1e-8
    # sums and sums of squares to 1e-8; the max |.| is one face value, whose
    # error is the 1e-10 solve's head error times the largest face K
    # (sigma2 = 4: K up to ~e^8), so it is held to 1e-7
    for d in (d64, dmx):
        assert d[0] < 1e-8 and d[1] < 1e-8 and d[2] < 1e-7, d


def test_config3_mixed_512cubed_matches_fp64(ctx):
    n = 512
    print(f"config 3: {n}^3 K field", flush=True)
    Y, _ = lognormal_lnK(n, n, n, 4.0, 16.0, 3031, device=ctx.device)
    K = torch.exp(Y)
    del Y
    torch.cuda.empty_cache()
    g, bc = Grid3D(n, n, n), BCSpec.darcy_x()
    cfg64 = ph.HeadSolveConfig.fp64_direct(1e-10, 7)
    cfg64.init_guess = ph.InitGuess.REFERENCE
    cfgmx = ph.HeadSolveConfig.mixed(1e-10, 7)
    cfgmx.init_guess = ph.InitGuess.REFERENCE
    h64, v64, r64 = _solve(ctx, K, g, bc, cfg64)
    torch.cuda.empty_cache()
    hmx, vmx, rmx = _solve(ctx, K, g, bc, cfgmx)
    print(f"config 3: fp64 {r64.num_iterations} it, mixed {rmx.num_iterations} it", flush=True)
    assert r64.converged and r64.true_residual <= 1e-10 * r64.bnorm
    assert rmx.converged and rmx.true_residual <= 1e-10 * rmx.bnorm
    eh = _rel(hmx, h64)
    ev = max(_rel(vmx.U, v64.U), _rel(vmx.V, v64.V), _rel(vmx.W, v64.W))
    print(f"config 3: mixed vs fp64: head {eh:.3e}, velocity {ev:.3e}", flush=True)
    assert eh < 1e-8
    assert ev < 1e-6
    # every x-plane carries the same total flux; no flow through y and z faces
    for v in (v64, vmx):
        q = v.U.view(n, n, n + 1).sum(dim=(0, 1))
        assert float((q - q.mean()).abs().max() / q.mean().abs()) < 1e-6
        Vf = v.V.view(n, n + 1, n)
        assert float(Vf[:, 0, :].abs().max()) == 0.0 and float(Vf[:, n, :].abs().max()) == 0.0


def test_config4_uneven_monte_carlo_batch(ctx):
    reals = ensemble.monte_carlo_set(4000, 32)
    assert [r.grid.nx for r in reals].count(96) == 32 and len(reals) == 96
    print("config 4: 96 K fields", flush=True)
    Ks = ensemble.generate_fields(ctx, reals)
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.mixed(1e-10, 6)
    hs, res, wss = ensemble.solve_batched(ctx, reals, Ks, bc, cfg)
    print(f"config 4: batched iterations {min(r.num_iterations for r in res)}.."
          f"{max(r.num_iterations for r in res)}", flush=True)
    for r in res:
        assert r.converged and r.true_residual <= 1e-10 * r.bnorm
    # one realisation of each size solved alone: same iterations, same head
    ws1 = ph.FlowWorkspace()
    for i in (0, 40, 95):
        re = reals[i]
        h1 = ctx.zeros(re.grid.num_cells())
        r1 = ph.solve_head(h1, Ks[i], re.grid, bc, cfg, ctx, ws1)
        assert res[i].num_iterations == r1.num_iterations, i
        assert res[i].restarts == r1.restarts, i
        assert _rel(hs[i], h1) <= 1e-12, i
    ws1.close()
    for w in wss:
        w.close()
