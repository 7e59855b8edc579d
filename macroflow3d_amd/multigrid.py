"""Multigrid layer: smoother, residual, transfers, hierarchy, V-cycle.

Mirrors MacroFlow3D's ``src/multigrid``: ``gsrb_smooth_3d`` (gsrb_3d.cuh:29-31),
``compute_residual_3d`` (residual_3d.cuh:30-32), ``restrict_3d``
(restrict_3d.cuh:14-15), ``prolong_3d_add`` (prolong_3d.cuh:14-15),
``MGConfig``/``MGHierarchy`` (mg_types.hpp:16-104), ``v_cycle_recursive`` and
``mg_solve`` (v_cycle.cuh:29-47). The hierarchy lives inside a
``FlowWorkspace``; the V-cycle runs in libmf3d_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from .core import BCSpec, Grid3D, PinSpec
from .runtime import HipContext, dptr


@dataclass
class MGConfig:
    """mg_types.hpp:16-24"""
    num_levels: int = 4
    pre_smooth: int = 2
    post_smooth: int = 2
    coarse_solve_iters: int = 50
    check_convergence_every: int = 1
    omega: float = 1.0
    verbose: bool = False


def hierarchy_grids(finest: Grid3D, num_levels: int) -> list[Grid3D]:
    """Level grids of MGHierarchy (mg_types.hpp:75-98): halve until a dim < 2."""
    out, cur = [], Grid3D(finest.nx, finest.ny, finest.nz, finest.dx, finest.dy, finest.dz)
    for level in range(num_levels):
        out.append(cur)
        if level < num_levels - 1:
            cur = Grid3D(cur.nx // 2, cur.ny // 2, cur.nz // 2, cur.dx * 2, cur.dy * 2, cur.dz * 2)
            if cur.nx < 2 or cur.ny < 2 or cur.nz < 2:
                break
    return out


def gsrb_smooth_3d(ctx: HipContext, grid: Grid3D, x: torch.Tensor, b: torch.Tensor,
                   K: torch.Tensor, num_iters: int, bc: BCSpec, pin: PinSpec = PinSpec()) -> None:
    """Red-black Gauss-Seidel, red ((i+j+k) even) then black, num_iters times."""
    bc.validate()
    n = grid.num_cells()
    _lib.call("mf3d_gsrb_smooth_3d", ctx.handle, C.byref(grid.c()), dptr(x, n, "x", ctx),
              dptr(b, n, "b", ctx), dptr(K, n, "K", ctx), n, int(num_iters), C.byref(bc.c()),
              1 if pin.enabled else 0)


def compute_residual_3d(ctx: HipContext, grid: Grid3D, x: torch.Tensor, b: torch.Tensor,
                        K: torch.Tensor, r: torch.Tensor, bc: BCSpec,
                        pin: PinSpec = PinSpec()) -> None:
    """r = b - A x with the smoother's operator convention."""
    bc.validate()
    n = grid.num_cells()
    _lib.call("mf3d_compute_residual_3d", ctx.handle, C.byref(grid.c()), dptr(x, n, "x", ctx),
              dptr(b, n, "b", ctx), dptr(K, n, "K", ctx), dptr(r, n, "r", ctx), n,
              C.byref(bc.c()), 1 if pin.enabled else 0)


def restrict_3d(ctx: HipContext, fine: Grid3D, coarse: Grid3D, r_fine: torch.Tensor,
                b_coarse: torch.Tensor) -> None:
    """b_c = average of the 8 children."""
    _lib.call("mf3d_restrict_3d", ctx.handle, C.byref(fine.c()), C.byref(coarse.c()),
              dptr(r_fine, fine.num_cells(), "r_fine", ctx), r_fine.numel(),
              dptr(b_coarse, coarse.num_cells(), "b_coarse", ctx), b_coarse.numel())


def prolong_3d_add(ctx: HipContext, coarse: Grid3D, fine: Grid3D, x_coarse: torch.Tensor,
                   x_fine: torch.Tensor) -> None:
    """x_f[i,j,k] += x_c[i/2, j/2, k/2] (piecewise-constant)."""
    _lib.call("mf3d_prolong_3d_add", ctx.handle, C.byref(coarse.c()), C.byref(fine.c()),
              dptr(x_coarse, coarse.num_cells(), "x_coarse", ctx), x_coarse.numel(),
              dptr(x_fine, fine.num_cells(), "x_fine", ctx), x_fine.numel())


def mg_setup(ctx: HipContext, workspace, grid: Grid3D, bc: BCSpec, cfg, K: torch.Tensor) -> int:
    """Builds the hierarchy of ``workspace`` for K (coarsened K per level and,
    with a DIRECT coarse solver, the dense coarsest inverse). Returns the level count.

    K must stay alive and unchanged while the hierarchy is used.
    """
    bc.validate()
    n = grid.num_cells()
    _lib.call("mf3d_mg_setup", workspace.handle, ctx.handle, C.byref(grid.c()), C.byref(bc.c()),
              C.byref(cfg.c()), dptr(K, n, "K", ctx), n)
    workspace._K_ref = K
    return int(_lib.lib().mf3d_mg_num_levels(workspace.handle))


def v_cycle(ctx: HipContext, workspace, r: torch.Tensor, z: torch.Tensor) -> None:
    """One V-cycle (v_cycle_recursive, v_cycle.cu:36-74) from z = 0 with b = r."""
    n = r.numel()
    _lib.call("mf3d_mg_precondition", workspace.handle, ctx.handle, dptr(r, n, "r", ctx),
              dptr(z, n, "z", ctx), n)


def smooth_level(ctx: HipContext, workspace, level: int, num_iters: int) -> None:
    """num_iters GSRB iterations on one level of the workspace's hierarchy, in
    place on that level's own buffers (timing and in-situ tests)."""
    _lib.call("mf3d_mg_smooth", workspace.handle, ctx.handle, int(level), int(num_iters))
