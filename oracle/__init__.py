"""CPU FP64 oracle for the Darcy head-solve path (TEST INFRASTRUCTURE ONLY).

ctypes binding of ``libmf3d_oracle.so`` (built from ``mf3d_oracle.cpp`` by the
Makefile next to this file). Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only to check
GPU results against; the product package ``macroflow3d_amd`` never imports it.

Each wrapper names the reference function it restates (MacroFlow3D source
paths, file:line).

Parity status of each oracle function (how it is pinned, see tests/):
  * operator / residual / GSRB / restrict / prolong / coarsen / rhs / guess:
    closed forms, discrete invariants and dense brute force
    (tests/test_oracle_pinning.py);
  * PCG, MG, CG solves: dense brute-force solve on small grids and the
    homogeneous / layered closed forms;
  * velocity: homogeneous and layered closed forms, flux balance;
  * velocity diagnostics: closed forms (linear fields with exact divergence,
    curl and helicity), telescoping flux sums, and the divergence of the
    velocity of a converged head solve (= the solve's residual);
  * field stats: numpy min / max / mean;
  * dot / nrm2: exact-rational checks; the reference's cuBLAS reduction order
    is closed source and parity unpinned (not reproducible, not needed);
  * lognormal generator: covariance / variance recovery (statistical) -- it is
    not a reference function (the reference's curand spectral generator is
    parity unpinned).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmf3d_oracle.so")

BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int * 6), ("value", C.c_double * 6)]


class Config(C.Structure):
    _fields_ = [("solver_type", C.c_int), ("mg_levels", C.c_int),
                ("mg_pre_smooth", C.c_int), ("mg_post_smooth", C.c_int),
                ("mg_max_cycles", C.c_int), ("mg_coarse_iters", C.c_int),
                ("cg_max_iter", C.c_int), ("cg_check_every", C.c_int),
                ("cg_rtol", C.c_double), ("rtol", C.c_double),
                ("pin_mode", C.c_int), ("init_guess", C.c_int),
                ("coarse_solver", C.c_int), ("stop_norm", C.c_int)]


class Result(C.Structure):
    _fields_ = [("num_iterations", C.c_int), ("converged", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("bnorm", C.c_double), ("true_residual", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        d = P(C.c_double)
        g, b, cfg = P(Grid), P(BC), P(Config)
        sig = {
            "orc_num_threads": (C.c_int, []),
            "orc_set_num_threads": (None, [C.c_int]),
            "orc_init_head_guess": (None, [g, b, d]),
            "orc_build_rhs_head": (None, [g, b, d, d]),
            "orc_coarsen_K": (None, [g, d, d]),
            "orc_gsrb_smooth_3d": (None, [g, b, C.c_int, d, d, d, C.c_int]),
            "orc_compute_residual_3d": (None, [g, b, C.c_int, d, d, d, d]),
            "orc_restrict_3d": (None, [g, d, d]),
            "orc_prolong_3d_add": (None, [g, d, d]),
            "orc_varcoeff_apply": (None, [g, b, C.c_int, d, d, d]),
            "orc_dot": (C.c_double, [C.c_size_t, d, d]),
            "orc_nrm2": (C.c_double, [C.c_size_t, d]),
            "orc_dense_inverse": (C.c_int, [g, b, C.c_int, d, d]),
            "orc_v_cycle_precond": (C.c_int, [g, b, cfg, d, d, d]),
            "orc_solve_head": (C.c_int, [g, b, cfg, d, d, P(Result)]),
            "orc_velocity_compact": (None, [g, b, d, d, d, d, d]),
            "orc_velocity_padded": (None, [g, b, d, d, d, d, d]),
            "orc_velocity_diagnostics": (None, [g, d, d, d, d, d, d]),
            "orc_field_stats": (None, [C.c_size_t, d, d]),
            "orc_lognormal_lnK": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_uint64, d, d]),
            "orc_gauss_pair": (None, [C.c_uint64, C.c_uint64, d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def grid(nx, ny, nz, dx=1.0, dy=None, dz=None) -> Grid:
    return Grid(nx, ny, nz, dx, dx if dy is None else dy, dx if dz is None else dz)


def bc(types, values=(0.0,) * 6) -> BC:
    o = BC()
    for i in range(6):
        o.type[i] = int(types[i])
        o.value[i] = float(values[i])
    return o


def darcy_bc(h_west=1.0, h_east=0.0) -> BC:
    """Dirichlet h_west at x=0, h_east at x=L, no-flow on y and z faces."""
    return bc([BC_DIRICHLET, BC_DIRICHLET] + [BC_NEUMANN] * 4,
              [h_west, h_east, 0, 0, 0, 0])


@dataclass
class SolveParams:
    solver_type: int = 2           # 0 CG, 1 MG, 2 PCG_MG
    mg_levels: int = 4
    mg_pre_smooth: int = 2
    mg_post_smooth: int = 2
    mg_max_cycles: int = 20
    mg_coarse_iters: int = 50
    cg_max_iter: int = 1000
    cg_check_every: int = 10
    cg_rtol: float = 1e-8
    rtol: float = 1e-6
    pin_mode: int = 0              # 0 Auto, 1 On, 2 Off
    init_guess: int = 0            # 0 reference guess, 1 zero, 2 caller's h
    coarse_solver: int = 0         # 0 GSRB iterations, 1 dense direct
    stop_norm: int = 0             # 0 ||r||/||r0||, 1 ||r||/||b||

    def c(self) -> Config:
        return Config(*[getattr(self, f[0]) for f in Config._fields_])


def mode_r(**kw) -> SolveParams:
    """Reference semantics (solve_head.cuh:64-96 defaults)."""
    return SolveParams(**kw)


def mode_n(**kw) -> SolveParams:
    """Extension semantics: direct coarsest, ||r||/||b||, zero guess, check every iteration."""
    base = dict(coarse_solver=1, stop_norm=1, init_guess=1, cg_check_every=1, cg_rtol=1e-10)
    base.update(kw)
    return SolveParams(**base)


def ncells(g: Grid) -> int:
    return g.nx * g.ny * g.nz


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()


# -- field operations -------------------------------------------------------

def init_head_guess(g, b):
    """init_head_guess (src/physics/flow/solve_head.cu:56-73)."""
    h = np.zeros(ncells(g))
    lib().orc_init_head_guess(C.byref(g), C.byref(b), _p(h))
    return h


def build_rhs_head(g, b, K):
    """build_rhs_head (src/physics/flow/rhs_head.cu:121-194)."""
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    r = np.zeros(ncells(g))
    lib().orc_build_rhs_head(C.byref(g), C.byref(b), _p(K), _p(r))
    return r


def coarsen_K(coarse, Kf):
    """coarsen_K (src/physics/flow/coarsen_K.cu:15-78)."""
    Kf = np.ascontiguousarray(Kf, dtype=np.float64).ravel()
    Kc = np.zeros(ncells(coarse))
    lib().orc_coarsen_K(C.byref(coarse), _p(Kf), _p(Kc))
    return Kc


def gsrb_smooth_3d(g, b_, pin, x, rhs, K, iters):
    """gsrb_smooth_3d (src/multigrid/smoothers/gsrb_3d.cu:368-576); returns new x."""
    x = np.array(x, dtype=np.float64).ravel()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).ravel()
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    lib().orc_gsrb_smooth_3d(C.byref(g), C.byref(b_), int(pin), _p(x), _p(rhs), _p(K), int(iters))
    return x


def compute_residual_3d(g, b_, pin, x, rhs, K):
    """compute_residual_3d (src/multigrid/smoothers/residual_3d.cu:378-503)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).ravel()
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    r = np.zeros(ncells(g))
    lib().orc_compute_residual_3d(C.byref(g), C.byref(b_), int(pin), _p(x), _p(rhs), _p(K), _p(r))
    return r


def restrict_3d(coarse, f):
    """restrict_3d (src/multigrid/transfer/restrict_3d.cu:12-81)."""
    f = np.ascontiguousarray(f, dtype=np.float64).ravel()
    c = np.zeros(ncells(coarse))
    lib().orc_restrict_3d(C.byref(coarse), _p(f), _p(c))
    return c


def prolong_3d_add(coarse, c, f):
    """prolong_3d_add (src/multigrid/transfer/prolong_3d.cu:520-621); returns f + P c."""
    c = np.ascontiguousarray(c, dtype=np.float64).ravel()
    f = np.array(f, dtype=np.float64).ravel()
    lib().orc_prolong_3d_add(C.byref(coarse), _p(c), _p(f))
    return f


def varcoeff_apply(g, b_, pin, K, x):
    """VarCoeffLaplacian::apply (src/numerics/operators/varcoeff_laplacian.cu:241-277)."""
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.zeros(ncells(g))
    lib().orc_varcoeff_apply(C.byref(g), C.byref(b_), int(pin), _p(K), _p(x), _p(y))
    return y


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    return lib().orc_dot(a.size, _p(a), _p(b))


def nrm2(a):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    return lib().orc_nrm2(a.size, _p(a))


def dense_inverse(g, b_, pin, K):
    """Inverse of the smoother/residual operator of one level (dense, FP64)."""
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    n = ncells(g)
    out = np.zeros(n * n)
    rc = lib().orc_dense_inverse(C.byref(g), C.byref(b_), int(pin), _p(K), _p(out))
    if rc != 0:
        raise RuntimeError("coarse operator is not definite")
    return out.reshape(n, n)


def v_cycle_precond(g, b_, params: SolveParams, K, r):
    """MultigridPreconditioner::apply (src/numerics/solvers/mg_preconditioner.cu:16-45)."""
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    z = np.zeros(ncells(g))
    cfg = params.c()
    rc = lib().orc_v_cycle_precond(C.byref(g), C.byref(b_), C.byref(cfg), _p(K), _p(r), _p(z))
    if rc:
        raise RuntimeError("oracle v-cycle failed")
    return z


def solve_head(g, b_, params: SolveParams, K, h0=None):
    """solve_head (src/physics/flow/solve_head.cu:75-243). Returns (h, Result)."""
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    h = np.zeros(ncells(g)) if h0 is None else np.array(h0, dtype=np.float64).ravel()
    cfg = params.c()
    res = Result()
    rc = lib().orc_solve_head(C.byref(g), C.byref(b_), C.byref(cfg), _p(K), _p(h), C.byref(res))
    if rc:
        raise RuntimeError(f"oracle solve failed ({rc})")
    return h, res


def velocity_compact(g, b_, h, K):
    """compute_velocity_from_head, compact MAC layout (velocity_from_head.cu:601-691)."""
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    U = np.zeros((g.nx + 1) * g.ny * g.nz)
    V = np.zeros(g.nx * (g.ny + 1) * g.nz)
    W = np.zeros(g.nx * g.ny * (g.nz + 1))
    lib().orc_velocity_compact(C.byref(g), C.byref(b_), _p(h), _p(K), _p(U), _p(V), _p(W))
    return U, V, W


def velocity_padded(g, b_, h, K):
    """compute_velocity_from_head, padded layout (velocity_from_head.cu:697-777)."""
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    K = np.ascontiguousarray(K, dtype=np.float64).ravel()
    n = (g.nx + 1) * (g.ny + 1) * (g.nz + 1)
    U, V, W = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_velocity_padded(C.byref(g), C.byref(b_), _p(h), _p(K), _p(U), _p(V), _p(W))
    return U, V, W


def velocity_diagnostics(g, U, V, W):
    """compute_velocity_diagnostics on a padded field (velocity_diagnostics.cu:46-177):
    (divergence, |curl|, helicity) at the cell centres."""
    n = (g.nx + 1) * (g.ny + 1) * (g.nz + 1)
    arrs = []
    for a in (U, V, W):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != n:
            raise ValueError(f"padded component has {a.size} values, expected {n}")
        arrs.append(a)
    nc = g.nx * g.ny * g.nz
    div, om, hel = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    lib().orc_velocity_diagnostics(C.byref(g), *(_p(a) for a in arrs), _p(div), _p(om), _p(hel))
    return div, om, hel


def field_stats(x):
    """compute_field_stats (stochastic.cu:369-414): (min, max, mean)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros(3)
    lib().orc_field_stats(x.size, _p(x), _p(out))
    return float(out[0]), float(out[1]), float(out[2])


def lognormal_lnK(nx, ny, nz, sigma2, corr_len, seed):
    """FFT circulant-embedding Gaussian field (not a reference function)."""
    out = np.zeros(nx * ny * nz)
    stats = np.zeros(5)
    lib().orc_lognormal_lnK(nx, ny, nz, float(sigma2), float(corr_len), int(seed), _p(out),
                            _p(stats))
    return out, stats


def gauss_pair(seed, k):
    out = np.zeros(2)
    lib().orc_gauss_pair(int(seed), int(k), _p(out))
    return out
