"""Numerics layer: operator, preconditioner, BLAS-1, PCG.

Mirrors MacroFlow3D's ``src/numerics``: ``VarCoeffLaplacian``
(operators/varcoeff_laplacian.cuh:43-67), ``MultigridPreconditioner``
(solvers/mg_preconditioner.cuh:34-81), ``PCGConfig``/``PCGResult``
(solvers/pcg.cuh:29-43), ``pcg_solve`` (pcg.cuh:95-185) and ``blas::*``
(blas/*.cuh). The vector kernels and reductions are libmf3d_hip.so's; this
``pcg_solve`` is the generic host loop of the reference (any operator and
preconditioner with ``apply``), while ``physics.solve_head`` runs the fused,
device-resident PCG.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from .core import BCSpec, Grid3D, PinSpec
from .multigrid import mg_setup, v_cycle
from .runtime import HipContext, dptr


class ReductionWorkspace:
    """Scratch for the two-stage reductions (blas/reduction_workspace.cuh:9-31)."""

    def __init__(self):
        from .physics import FlowWorkspace
        self.ws = FlowWorkspace()
        self.d_scalar = None

    def scalar(self, ctx: HipContext) -> torch.Tensor:
        if self.d_scalar is None or self.d_scalar.device != ctx.device:
            self.d_scalar = ctx.empty(1)
        return self.d_scalar


class blas:  # noqa: N801 -- namespace, as blas:: in the reference
    @staticmethod
    def dot_device(ctx, x, y, d_result, red: ReductionWorkspace) -> None:
        n = min(x.numel(), y.numel())
        _lib.call("mf3d_dot", red.ws.handle, ctx.handle, dptr(x, n, "x", ctx),
                  dptr(y, n, "y", ctx), n, dptr(d_result, 1, "d_result", ctx))

    @staticmethod
    def dot_host(ctx, x, y, red: ReductionWorkspace) -> float:
        d = red.scalar(ctx)
        blas.dot_device(ctx, x, y, d, red)
        ctx.synchronize()
        return float(d.item())

    @staticmethod
    def nrm2_device(ctx, x, d_result, red: ReductionWorkspace) -> None:
        n = x.numel()
        _lib.call("mf3d_nrm2", red.ws.handle, ctx.handle, dptr(x, n, "x", ctx), n,
                  dptr(d_result, 1, "d_result", ctx))

    @staticmethod
    def nrm2_host(ctx, x, red: ReductionWorkspace) -> float:
        d = red.scalar(ctx)
        blas.nrm2_device(ctx, x, d, red)
        ctx.synchronize()
        return float(d.item())

    @staticmethod
    def axpy(ctx, a, x, y) -> None:
        n = min(x.numel(), y.numel())
        _lib.call("mf3d_axpy", ctx.handle, float(a), dptr(x, n, "x", ctx), dptr(y, n, "y", ctx), n)

    @staticmethod
    def axpby(ctx, a, x, b, y) -> None:
        n = min(x.numel(), y.numel())
        _lib.call("mf3d_axpby", ctx.handle, float(a), dptr(x, n, "x", ctx), float(b),
                  dptr(y, n, "y", ctx), n)

    @staticmethod
    def copy(ctx, x, y) -> None:
        n = min(x.numel(), y.numel())
        _lib.call("mf3d_copy", ctx.handle, dptr(x, n, "x", ctx), dptr(y, n, "y", ctx), n)

    @staticmethod
    def fill(ctx, x, value) -> None:
        _lib.call("mf3d_fill", ctx.handle, dptr(x, x.numel(), "x", ctx), x.numel(), float(value))

    @staticmethod
    def scal(ctx, x, a) -> None:
        _lib.call("mf3d_scal", ctx.handle, dptr(x, x.numel(), "x", ctx), x.numel(), float(a))


class VarCoeffLaplacian:
    """y = A x = -sum_f K_f (x_C - x_nb) / dx^2, K_f harmonic mean; Dirichlet
    faces add 2 K_C x_C, Neumann faces add nothing (varcoeff_laplacian.cu:38-239)."""

    def __init__(self, grid: Grid3D, K: torch.Tensor, bc: BCSpec, pin: PinSpec = PinSpec()):
        if K.numel() != grid.num_cells():
            raise RuntimeError("K field size must match grid")
        bc.validate()
        self.grid, self.K, self.bc, self.pin = grid, K, bc, pin

    def apply(self, ctx: HipContext, x: torch.Tensor, y: torch.Tensor) -> None:
        n = self.grid.num_cells()
        _lib.call("mf3d_varcoeff_apply", ctx.handle, C.byref(self.grid.c()), C.byref(self.bc.c()),
                  1 if self.pin.enabled else 0, dptr(self.K, n, "K", ctx), dptr(x, n, "x", ctx),
                  dptr(y, n, "y", ctx), n)


class NegatedOperator:
    """y = -A x (operators/negated_operator.cuh:34-61)."""

    def __init__(self, op):
        self.op = op

    def apply(self, ctx, x, y):
        self.op.apply(ctx, x, y)
        blas.scal(ctx, y, -1.0)


class IdentityPreconditioner:
    """z = r (solvers/preconditioner.cu:12-16)."""

    def apply(self, ctx, r, z):
        blas.copy(ctx, r, z)


class MultigridPreconditioner:
    """z = one V-cycle with b = r and x = 0 (mg_preconditioner.cu:16-45).

    ``hierarchy`` is a FlowWorkspace whose hierarchy was built by
    ``multigrid.mg_setup`` (or a ``solve_head`` call) for the operator's K.
    """

    def __init__(self, hierarchy, bc: BCSpec = None, config=None, pin: PinSpec = PinSpec()):
        self.ws = hierarchy

    @staticmethod
    def build(ctx, workspace, grid, bc, head_cfg, K) -> "MultigridPreconditioner":
        mg_setup(ctx, workspace, grid, bc, head_cfg, K)
        return MultigridPreconditioner(workspace)

    def apply(self, ctx: HipContext, r: torch.Tensor, z: torch.Tensor) -> None:
        v_cycle(ctx, self.ws, r, z)


@dataclass
class PCGConfig:
    max_iter: int = 1000
    check_every: int = 10
    rtol: float = 1e-6
    verbose: bool = False


@dataclass
class PCGResult:
    iterations: int = 0
    initial_residual: float = 0.0
    final_residual: float = 0.0
    converged: bool = False


class PCGWorkspace:
    def __init__(self):
        self.r = self.z = self.p = self.Ap = None
        self.red = ReductionWorkspace()

    def ensure(self, n: int, ctx: HipContext) -> None:
        if self.r is None or self.r.numel() != n:
            self.r, self.z, self.p, self.Ap = (ctx.empty(n) for _ in range(4))


def pcg_solve(ctx: HipContext, A, M, b: torch.Tensor, x: torch.Tensor, cfg: PCGConfig,
              ws: PCGWorkspace) -> PCGResult:
    """Generic PCG host loop (pcg.cuh:95-185), same operations and stopping rule."""
    res = PCGResult()
    n = x.numel()
    ws.ensure(n, ctx)
    r, z, p, Ap, red = ws.r, ws.z, ws.p, ws.Ap, ws.red
    A.apply(ctx, x, Ap)
    blas.copy(ctx, b, r)
    blas.axpy(ctx, -1.0, Ap, r)
    res.initial_residual = blas.nrm2_host(ctx, r, red)
    if res.initial_residual < 1e-15:
        res.converged = True
        res.final_residual = res.initial_residual
        return res
    M.apply(ctx, r, z)
    blas.copy(ctx, z, p)
    rz_old = blas.dot_host(ctx, r, z, red)
    for it in range(cfg.max_iter):
        res.iterations = it + 1
        A.apply(ctx, p, Ap)
        pAp = blas.dot_host(ctx, p, Ap, red)
        alpha = rz_old / pAp
        blas.axpy(ctx, alpha, p, x)
        blas.axpy(ctx, -alpha, Ap, r)
        if (it + 1) % cfg.check_every == 0 or it == cfg.max_iter - 1:
            res.final_residual = blas.nrm2_host(ctx, r, red)
            if res.final_residual / res.initial_residual < cfg.rtol:
                res.converged = True
                return res
        M.apply(ctx, r, z)
        rz_new = blas.dot_host(ctx, r, z, red)
        beta = rz_new / rz_old
        blas.axpby(ctx, 1.0, z, beta, p)
        rz_old = rz_new
    res.final_residual = blas.nrm2_host(ctx, r, red)
    return res
