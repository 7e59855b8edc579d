"""Head solve and Darcy velocities: the physics layer of the path.

Mirrors MacroFlow3D's ``src/physics/flow`` and ``src/physics/common``:
``solve_head`` (solve_head.cuh:158-160), ``init_head_guess`` (:173-174),
``HeadSolverType``/``HeadSolveConfig``/``HeadSolveResult`` (:52-141),
``build_rhs_head`` (rhs_head.cuh:50-51), ``coarsen_K`` (coarsen_K.cuh:46-47),
``compute_velocity_from_head`` (velocity_from_head.cuh:54-73) with
``VelocityField``/``PaddedVelocityField`` (fields.cuh:108-319), the checksum
utilities (velocity_from_head.cuh:79-107), the velocity diagnostics
(velocity_diagnostics.cuh:31-92), ``compute_field_stats`` (stochastic.cu:369-414)
and ``FlowWorkspace`` (workspaces.cuh:111-167). Every computation runs in
libmf3d_hip.so.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass

import torch

from . import _lib
from .core import BCSpec, BCType, Grid3D, PinMode
from .runtime import HipContext, dptr


class HeadSolverType(enum.IntEnum):
    CG = 0
    MG = 1
    PCG_MG = 2


class InitGuess(enum.IntEnum):
    REFERENCE = 0  # linear in x between Dirichlet x faces, else zero (solve_head.cu:56-73)
    ZERO = 1
    PROVIDED = 2   # keep the caller's h (warm start)


class CoarseSolver(enum.IntEnum):
    GSRB = 0       # mg_coarse_iters GSRB iterations (v_cycle.cu:41-45)
    DIRECT = 1     # dense FP64 inverse of the coarsest operator


class StopNorm(enum.IntEnum):
    INITIAL_RESIDUAL = 0  # ||r|| / ||r0|| (pcg.cuh:156-164)
    RHS_NORM = 1          # ||r|| / ||b||


class PrecondPrecision(enum.IntEnum):
    FP64 = 0
    FP32 = 1


class CoefFormat(enum.IntEnum):
    FP32 = 0
    BF16 = 1
    FP16 = 2


@dataclass
class HeadSolveConfig:
    """HeadSolveConfig (solve_head.cuh:64-96) plus extensions (defaults = reference)."""
    solver_type: HeadSolverType = HeadSolverType.PCG_MG
    mg_levels: int = 4
    mg_pre_smooth: int = 2
    mg_post_smooth: int = 2
    mg_max_cycles: int = 20
    mg_coarse_iters: int = 50
    cg_max_iter: int = 1000
    cg_check_every: int = 10
    cg_rtol: float = 1e-8
    rtol: float = 1e-6
    pin: PinMode = PinMode.Auto
    # extensions
    init_guess: InitGuess = InitGuess.REFERENCE
    coarse_solver: CoarseSolver = CoarseSolver.GSRB
    stop_norm: StopNorm = StopNorm.INITIAL_RESIDUAL
    precond_precision: PrecondPrecision = PrecondPrecision.FP64
    coef_format: CoefFormat = CoefFormat.BF16
    true_residual: bool = True
    flexible: bool = False

    @staticmethod
    def from_yaml(flow: dict) -> "HeadSolveConfig":
        """HeadSolveConfig::from_yaml (solve_head.cuh:98-130) from a ``flow:`` mapping."""
        cfg = HeadSolveConfig()
        s = str(flow.get("solver", "mg"))
        cfg.solver_type = {"cg": HeadSolverType.CG, "mg": HeadSolverType.MG}.get(
            s, HeadSolverType.PCG_MG)
        for k in ("mg_levels", "mg_pre_smooth", "mg_post_smooth", "mg_max_cycles",
                  "mg_coarse_iters", "cg_max_iter", "cg_check_every"):
            if k in flow:
                setattr(cfg, k, int(flow[k]))
        for k in ("cg_rtol", "rtol"):
            if k in flow:
                setattr(cfg, k, float(flow[k]))
        pin = flow.get("pin", {})
        mode = pin.get("mode", "auto") if isinstance(pin, dict) else str(pin)
        cfg.pin = {"on": PinMode.On, "off": PinMode.Off}.get(str(mode), PinMode.Auto)
        return cfg

    @staticmethod
    def fp64_direct(rtol: float = 1e-10, mg_levels: int = 6) -> "HeadSolveConfig":
        """Direct coarsest solve, ||r||/||b|| <= rtol checked every iteration, zero guess."""
        return HeadSolveConfig(mg_levels=mg_levels, cg_rtol=rtol, cg_check_every=1,
                               init_guess=InitGuess.ZERO, coarse_solver=CoarseSolver.DIRECT,
                               stop_norm=StopNorm.RHS_NORM)

    @staticmethod
    def mixed(rtol: float = 1e-10, mg_levels: int = 6, coef: "CoefFormat" = None,
              flexible: bool = False) -> "HeadSolveConfig":
        """fp64_direct() with the FP32 V-cycle and 16-bit (default BF16) face coefficients;
        the PCG, its dot products, the operator and the coarsest solve stay FP64."""
        c = HeadSolveConfig.fp64_direct(rtol, mg_levels)
        c.precond_precision = PrecondPrecision.FP32
        c.coef_format = CoefFormat.BF16 if coef is None else coef
        c.flexible = flexible
        return c

    def c(self) -> _lib.mf3d_head_solve_config:
        o = _lib.mf3d_head_solve_config()
        _lib.lib().mf3d_default_config(C.byref(o))
        o.solver_type = int(self.solver_type)
        o.mg_levels = int(self.mg_levels)
        o.mg_pre_smooth = int(self.mg_pre_smooth)
        o.mg_post_smooth = int(self.mg_post_smooth)
        o.mg_max_cycles = int(self.mg_max_cycles)
        o.mg_coarse_iters = int(self.mg_coarse_iters)
        o.cg_max_iter = int(self.cg_max_iter)
        o.cg_check_every = int(self.cg_check_every)
        o.cg_rtol = float(self.cg_rtol)
        o.rtol = float(self.rtol)
        o.pin_mode = int(self.pin)
        o.init_guess = int(self.init_guess)
        o.coarse_solver = int(self.coarse_solver)
        o.stop_norm = int(self.stop_norm)
        o.precond_precision = int(self.precond_precision)
        o.coef_format = int(self.coef_format)
        o.true_residual = 1 if self.true_residual else 0
        o.flexible = 1 if self.flexible else 0
        return o


@dataclass
class HeadSolveResult:
    """HeadSolveResult (solve_head.cuh:136-141) plus extensions."""
    num_iterations: int = 0
    initial_residual: float = 0.0
    final_residual: float = 0.0
    converged: bool = False
    bnorm: float = 0.0
    true_residual: float = -1.0
    setup_ms: float = 0.0
    solve_ms: float = 0.0
    restarts: int = 0

    @property
    def relative_true_residual(self) -> float:
        return self.true_residual / self.bnorm if self.bnorm > 0 else math.nan


class FlowWorkspace:
    """Owns the device hierarchy/Krylov buffers (workspaces.cuh:111-167).

    Allocation is lazy and reused while the grid and level count stay the same
    (solve_head.cu:89-94). Not re-entrant: one solve at a time.
    """

    def __init__(self):
        h = C.c_void_p()
        _lib.call("mf3d_workspace_create", C.byref(h))
        self._h = h

    @property
    def handle(self):
        return self._h

    def bytes(self) -> int:
        return int(_lib.lib().mf3d_workspace_bytes(self._h))

    TUNE_SPLIT = 4
    TUNE_SWEEP_BLOCK = 5
    TUNE_FOLD_FIRST = 10
    TUNE_FOLD_R = 11
    TUNE_FOLD_RZ = 12
    TUNE_OP_TILE = 13
    TUNE_GRAPH = 14
    TUNE_OP_CHUNK = 15
    TUNE_MARCH_BLOCKS = 16

    def set_tuning(self, key: int, value: int) -> None:
        """Kernel-layout knobs (colour-split storage, sweep workgroup size,
        folded PCG steps, plane-tile operator and its chunk depth, the FP64
        sweeps' workgroup target, graph replay); applied at the next set-up
        or solve. None changes the result beyond rounding; the sweep block
        size of the mixed V-cycle and the graph replay leave it bit-identical,
        the operator tile and chunk depth regroup the (p, Ap) partials, the
        FP64 sweeps' block size and workgroup target the (r, z) partials."""
        _lib.call("mf3d_workspace_set_tuning", self._h, int(key), int(value))

    def solve_counts(self) -> tuple[int, int]:
        """(operator applications, preconditioner applications) enqueued on
        this workspace since the last single solve began, graph replays
        included."""
        op, pc = C.c_long(), C.c_long()
        _lib.call("mf3d_workspace_solve_counts", self._h, C.byref(op), C.byref(pc))
        return op.value, pc.value

    def clear(self) -> None:
        self.close()
        self.__init__()

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().mf3d_workspace_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_head(h: torch.Tensor, K: torch.Tensor, grid: Grid3D, bc: BCSpec,
               cfg: HeadSolveConfig, ctx: HipContext, workspace: FlowWorkspace) -> HeadSolveResult:
    """physics::solve_head (solve_head.cu:75-243).

    h is overwritten with the initial guess unless cfg.init_guess is PROVIDED,
    then solved in place. Synchronises the context stream (the result is host data).
    """
    bc.validate()
    n = grid.num_cells()
    res = _lib.mf3d_head_solve_result()
    _lib.call("mf3d_solve_head", workspace.handle, ctx.handle, C.byref(grid.c()), C.byref(bc.c()),
              C.byref(cfg.c()), dptr(K, n, "K", ctx), dptr(h, n, "h", ctx), n, C.byref(res))
    return HeadSolveResult(res.num_iterations, res.initial_residual, res.final_residual,
                           bool(res.converged), res.bnorm, res.true_residual, res.setup_ms,
                           res.solve_ms, res.restarts)


def solve_head_batched(hs, Ks, grids, bc: BCSpec, cfg: HeadSolveConfig, ctx: HipContext,
                       workspaces, use_graph: bool = True) -> list:
    """Many independent head solves in one batched run (extension; replaces the
    sequential loop of EnsembleRunner.cu:259-311). One FlowWorkspace per
    realisation; grids may differ; BCs and configuration are shared. Needs the
    FP32 V-cycle with the direct coarsest solve (``HeadSolveConfig.mixed``).
    Synchronises; returns one HeadSolveResult per realisation."""
    bc.validate()
    m = len(hs)
    if not (len(Ks) == len(grids) == len(workspaces) == m) or m == 0:
        raise ValueError("hs, Ks, grids and workspaces must have the same nonzero length")
    G = (_lib.mf3d_grid * m)(*[g.c() for g in grids])
    W = (C.c_void_p * m)(*[w.handle.value for w in workspaces])
    KP = (C.c_void_p * m)(*[dptr(K, g.num_cells(), "K", ctx) for K, g in zip(Ks, grids)])
    HP = (C.c_void_p * m)(*[dptr(h, g.num_cells(), "h", ctx) for h, g in zip(hs, grids)])
    N = (C.c_size_t * m)(*[g.num_cells() for g in grids])
    R = (_lib.mf3d_head_solve_result * m)()
    _lib.call("mf3d_solve_head_batched", m, W, ctx.handle, G, C.byref(bc.c()), C.byref(cfg.c()),
              KP, HP, N, 0 if use_graph else 1, R)
    return [HeadSolveResult(r.num_iterations, r.initial_residual, r.final_residual,
                            bool(r.converged), r.bnorm, r.true_residual, r.setup_ms, r.solve_ms,
                            r.restarts) for r in R]


def init_head_guess(h: torch.Tensor, grid: Grid3D, bc: BCSpec, ctx: HipContext) -> None:
    """physics::init_head_guess (solve_head.cu:56-73)."""
    _lib.call("mf3d_init_head_guess", ctx.handle, C.byref(grid.c()), C.byref(bc.c()),
              dptr(h, grid.num_cells(), "h", ctx), h.numel())


def build_rhs_head(rhs: torch.Tensor, K: torch.Tensor, grid: Grid3D, bc: BCSpec,
                   ctx: HipContext) -> None:
    """physics::build_rhs_head (rhs_head.cu:121-194)."""
    n = grid.num_cells()
    if rhs.numel() < n or K.numel() < n:
        raise RuntimeError("RHS or K buffer too small")
    _lib.call("mf3d_build_rhs_head", ctx.handle, C.byref(grid.c()), C.byref(bc.c()),
              dptr(K, n, "K", ctx), dptr(rhs, n, "rhs", ctx), n)


def coarsen_K(ctx: HipContext, coarse: Grid3D, fine: Grid3D, K_fine: torch.Tensor,
              K_coarse: torch.Tensor) -> None:
    """physics::coarsen_K (coarsen_K.cu:55-78): geometric mean of 2x2x2 children."""
    _lib.call("mf3d_coarsen_K", ctx.handle, C.byref(coarse.c()), C.byref(fine.c()),
              dptr(K_fine, fine.num_cells(), "K_fine", ctx), K_fine.numel(),
              dptr(K_coarse, coarse.num_cells(), "K_coarse", ctx), K_coarse.numel())


class VelocityField:
    """Compact staggered (MAC) velocity (fields.cuh:108-212).

    U: (nx+1)*ny*nz at i + j*(nx+1) + k*(nx+1)*ny; V: nx*(ny+1)*nz; W: nx*ny*(nz+1).
    """

    def __init__(self, grid: Grid3D, ctx: HipContext):
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        self.dx, self.dy, self.dz = grid.dx, grid.dy, grid.dz
        self.U = ctx.zeros(self.size_U())
        self.V = ctx.zeros(self.size_V())
        self.W = ctx.zeros(self.size_W())

    def size_U(self) -> int:
        return (self.nx + 1) * self.ny * self.nz

    def size_V(self) -> int:
        return self.nx * (self.ny + 1) * self.nz

    def size_W(self) -> int:
        return self.nx * self.ny * (self.nz + 1)

    def total_size(self) -> int:
        return self.size_U() + self.size_V() + self.size_W()

    def grid(self) -> Grid3D:
        return Grid3D(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz)


class PaddedVelocityField:
    """Par2-compatible padded velocity (fields.cuh:248-319): each component has
    (nx+1)*(ny+1)*(nz+1) values at k*(ny+1)*(nx+1) + j*(nx+1) + i; padding stays 0."""

    def __init__(self, grid: Grid3D, ctx: HipContext):
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        self.dx, self.dy, self.dz = grid.dx, grid.dy, grid.dz
        s = self.field_size()
        self.U, self.V, self.W = ctx.zeros(s), ctx.zeros(s), ctx.zeros(s)

    def field_size(self) -> int:
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    def total_size(self) -> int:
        return 3 * self.field_size()

    def merge_id(self, i: int, j: int, k: int) -> int:
        return k * (self.ny + 1) * (self.nx + 1) + j * (self.nx + 1) + i

    def grid(self) -> Grid3D:
        return Grid3D(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz)


def compute_velocity_from_head(vel, head: torch.Tensor, K: torch.Tensor, grid: Grid3D,
                               bc: BCSpec, ctx: HipContext) -> None:
    """physics::compute_velocity_from_head, both layouts (velocity_from_head.cu:601-777).

    Enqueued on the context stream, then the stream is synchronised (as the
    reference does, velocity_from_head.cu:690,776).
    """
    n = grid.num_cells()
    g, b = C.byref(grid.c()), C.byref(bc.c())
    if isinstance(vel, PaddedVelocityField):
        s = vel.field_size()
        _lib.call("mf3d_velocity_padded", ctx.handle, g, b, dptr(head, n, "head", ctx),
                  dptr(K, n, "K", ctx), n, dptr(vel.U, s, "U", ctx), dptr(vel.V, s, "V", ctx),
                  dptr(vel.W, s, "W", ctx), s)
    elif isinstance(vel, VelocityField):
        _lib.call("mf3d_velocity_compact", ctx.handle, g, b, dptr(head, n, "head", ctx),
                  dptr(K, n, "K", ctx), n, dptr(vel.U, vel.size_U(), "U", ctx), vel.U.numel(),
                  dptr(vel.V, vel.size_V(), "V", ctx), vel.V.numel(),
                  dptr(vel.W, vel.size_W(), "W", ctx), vel.W.numel())
    else:
        raise TypeError("vel must be a VelocityField or PaddedVelocityField")
    ctx.synchronize()


# -- checksum utilities (velocity_from_head.cu:783-962) ---------------------

class _Scratch:
    ws: FlowWorkspace | None = None


def _ws() -> FlowWorkspace:
    if _Scratch.ws is None:
        _Scratch.ws = FlowWorkspace()
    return _Scratch.ws


def compute_norm2(data: torch.Tensor, ctx: HipContext) -> float:
    """sqrt(sum x^2) (velocity_from_head.cu:783-810), on the device, synchronous."""
    n = data.numel()
    if n == 0:
        return 0.0
    out = ctx.empty(1)
    _lib.call("mf3d_nrm2", _ws().handle, ctx.handle, dptr(data, n, "data", ctx), n,
              dptr(out, 1, "out", ctx))
    ctx.synchronize()
    return float(out.item())


def compute_sum(data: torch.Tensor, ctx: HipContext) -> float:
    """sum x (velocity_from_head.cu:870-893), on the device, synchronous."""
    n = data.numel()
    if n == 0:
        return 0.0
    ones = ctx.empty(n)
    _lib.call("mf3d_fill", ctx.handle, dptr(ones, n, "ones", ctx), n, 1.0)
    out = ctx.empty(1)
    _lib.call("mf3d_dot", _ws().handle, ctx.handle, dptr(data, n, "data", ctx),
              dptr(ones, n, "ones", ctx), n, dptr(out, 1, "out", ctx))
    ctx.synchronize()
    return float(out.item())


def compute_mean(data: torch.Tensor, ctx: HipContext) -> float:
    return compute_sum(data, ctx) / data.numel() if data.numel() else 0.0


def check_no_nans(data: torch.Tensor, ctx: HipContext) -> bool:
    """True when no value is NaN (a NaN makes the sum of squares NaN)."""
    return not math.isnan(compute_norm2(data, ctx))


def compute_field_stats(data: torch.Tensor, ctx: HipContext) -> tuple:
    """(min, max, mean) of a device field (stochastic.cu:369-414), synchronous;
    zeros for an empty field."""
    n = data.numel()
    if n == 0:
        return 0.0, 0.0, 0.0
    out = ctx.empty(3)
    _lib.call("mf3d_field_stats", _ws().handle, ctx.handle, dptr(data, n, "data", ctx), n,
              dptr(out, 3, "out", ctx))
    ctx.synchronize()
    lo, hi, mean = out.tolist()
    return lo, hi, mean


class VelocityDiagnostics:
    """Cell-centred divergence, |curl| and helicity of a velocity field
    (velocity_diagnostics.cuh:31-50), each nx*ny*nz values at i + nx*(j + ny*k)."""

    def __init__(self, grid: Grid3D, ctx: HipContext):
        self.resize(grid, ctx)

    def resize(self, grid: Grid3D, ctx: HipContext) -> None:
        n = grid.num_cells()
        self.divergence = ctx.zeros(n)
        self.vorticity_mag = ctx.zeros(n)
        self.helicity = ctx.zeros(n)
        self.stats = None  # (3, 3): min, max, mean per field, after compute

    def empty(self) -> bool:
        return self.divergence.numel() == 0


def compute_velocity_diagnostics(vel: "PaddedVelocityField", diag: VelocityDiagnostics,
                                 grid: Grid3D, ctx: HipContext) -> None:
    """physics::compute_velocity_diagnostics (velocity_diagnostics.cu:147-177).

    Divergence at every cell; |curl| and helicity at interior cells, 0 at the
    border and everywhere when a dimension is below 3 (a warning is printed
    then, as the reference does). Also fills ``diag.stats`` (min, max, mean of
    each field, computed on the device in the same pass). Synchronous.
    """
    if not isinstance(vel, PaddedVelocityField):
        raise TypeError("compute_velocity_diagnostics takes a PaddedVelocityField")
    n = grid.num_cells()
    if n == 0:
        return
    if min(grid.nx, grid.ny, grid.nz) < 3:
        print(f"       [diag] WARNING: grid too small for curl/helicity "
              f"({grid.nx}x{grid.ny}x{grid.nz}, need >=3 in each dim). "
              f"Only divergence computed.")
    s = vel.field_size()
    stats = ctx.empty(9)
    _lib.call("mf3d_velocity_diagnostics", _ws().handle, ctx.handle, C.byref(grid.c()),
              dptr(vel.U, s, "U", ctx), dptr(vel.V, s, "V", ctx), dptr(vel.W, s, "W", ctx), s,
              dptr(diag.divergence, n, "divergence", ctx),
              dptr(diag.vorticity_mag, n, "vorticity_mag", ctx),
              dptr(diag.helicity, n, "helicity", ctx), n, dptr(stats, 9, "stats", ctx))
    ctx.synchronize()
    diag.stats = stats.view(3, 3).cpu()


def print_velocity_diagnostics(diag: VelocityDiagnostics, realization_id: int,
                               ctx: HipContext) -> None:
    """One line per field with min, max, mean (velocity_diagnostics.cu:179-199)."""
    rows = [("div: ", ""), ("|w|: ", "  (incl. border=0)"), ("h:   ", "  (incl. border=0)")]
    fields = (diag.divergence, diag.vorticity_mag, diag.helicity)
    for (label, tail), f in zip(rows, fields):
        lo, hi, mean = compute_field_stats(f, ctx)
        print(f"       [diag] r={realization_id}  {label} min={lo:.4e}  max={hi:.4e}  "
              f"mean={mean:.4e}{tail}")


def velocity_checksums(vel, ctx: HipContext) -> dict:
    """Per-component sum, sum of squares and NaN flag (print_velocity_checksums,
    velocity_from_head.cu:942-962), as a dict."""
    out = {}
    for name in ("U", "V", "W"):
        t = getattr(vel, name)
        nrm = compute_norm2(t, ctx)
        out[name] = {"sum": compute_sum(t, ctx), "sumsq": nrm * nrm, "norm2": nrm,
                     "no_nans": not math.isnan(nrm)}
    return out
