"""Shared helpers for the tests: boundary mixes, grids and random fields in
both the product's and the oracle's types."""
from __future__ import annotations

import numpy as np

import oracle as O
from macroflow3d_amd.core import BCFace, BCSpec, BCType, Grid3D

D, N, P = 0, 1, 2

# name -> (types xmin..zmax, values); pin flag for the oracle helpers
BC_MIXES = {
    "darcy_x": ([D, D, N, N, N, N], [1.0, 0.0, 0, 0, 0, 0]),
    "all_dirichlet": ([D, D, D, D, D, D], [1.0, 0.0, 0.5, -0.25, 2.0, 0.1]),
    "periodic_yz": ([D, D, P, P, P, P], [1.0, 0.0, 0, 0, 0, 0]),
    "periodic_x_dirichlet_z": ([P, P, N, N, D, D], [0, 0, 0, 0, 1.0, 0.0]),
    "all_neumann": ([N, N, N, N, N, N], [0, 0, 0, 0, 0, 0]),
    "all_periodic": ([P, P, P, P, P, P], [0, 0, 0, 0, 0, 0]),
}


def bc_pair(name: str):
    t, v = BC_MIXES[name]
    spec = BCSpec(*[BCFace(BCType(t[i]), float(v[i])) for i in range(6)])
    return spec, O.bc(t, v)


def grid_pair(nx, ny, nz, dx=1.0):
    return Grid3D(nx, ny, nz, dx), O.grid(nx, ny, nz, dx)


def random_K(n, seed=0, sigma=1.0):
    rng = np.random.default_rng(seed)
    return np.exp(sigma * rng.standard_normal(n))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
