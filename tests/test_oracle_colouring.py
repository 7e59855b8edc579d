"""Pins the oracle's 3-D red-black colouring (no GPU).

gsrb_smooth_3d (gsrb_3d.cu:368-576) updates red cells ((i + j + k) even)
then black ones; every 7-point neighbour of a cell has the other colour, so a
colour pass is the same whatever order it visits its cells in. Here one
oracle sweep is compared with a cell-by-cell Gauss-Seidel emulation built
only from the oracle's residual (x_C <- x_C - r_C / d_C, with d_C the
diagonal read off the residual of indicator vectors), visiting each colour in
natural, reversed and shuffled order, and with a Jacobi pass that updates a
whole colour from the pass's input state. A colouring that got the y or z
parity wrong would let two same-colour neighbours see each other's new
values and fail these comparisons; the planted wrong colourings below show
that they do.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import bc_pair, random_K, rel_err


def _cells(g):
    k, j, i = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def _diagonal(g, b, pin, K):
    """d_C = -(A e_C)_C of the smoother's operator, from residuals of indicator
    vectors over the 27 classes (i mod 3, j mod 3, k mod 3): cells of one class
    are never neighbours (periodic sizes are multiples of 3 here)."""
    n = O.ncells(g)
    i, j, k = _cells(g)
    cls = (i % 3) + 3 * (j % 3) + 9 * (k % 3)
    d = np.zeros(n)
    zero = np.zeros(n)
    for c in range(27):
        e = (cls == c).astype(np.float64)
        r = O.compute_residual_3d(g, b, pin, e, zero, K)
        d[cls == c] = r[cls == c]
    return d


def _emulate(g, b, pin, K, rhs, x0, colour_of, order, d):
    """One red-then-black sweep, one cell at a time, in the given visit order."""
    x = x0.copy()
    for colour in (0, 1):
        cells = np.flatnonzero(colour_of == colour)
        for c in order(cells):
            r = O.compute_residual_3d(g, b, pin, x, rhs, K)
            x[c] -= r[c] / d[c]
    return x


def _jacobi_colours(g, b, pin, K, rhs, x0, colour_of, d):
    """Each colour updated at once from the state its pass starts from."""
    x = x0.copy()
    for colour in (0, 1):
        m = colour_of == colour
        r = O.compute_residual_3d(g, b, pin, x, rhs, K)
        x[m] -= r[m] / d[m]
    return x


CASES = [
    ("darcy_x", (9, 7, 5)),
    ("periodic_yz", (8, 6, 12)),
    ("all_periodic", (6, 12, 6)),
    ("all_neumann", (7, 6, 5)),
    ("periodic_x_dirichlet_z", (12, 5, 4)),
]


@pytest.mark.parametrize("mix,shape", CASES)
def test_colour_pass_is_independent_of_visit_order(mix, shape):
    g = O.grid(*shape, 0.5)
    _, b = bc_pair(mix)
    pin = mix == "all_neumann" or mix == "all_periodic"
    n = O.ncells(g)
    rng = np.random.default_rng(sum(shape) * 7 + len(mix))
    K = random_K(n, 3, 1.0)
    rhs = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    i, j, k = _cells(g)
    red_black = (i + j + k) & 1
    d = _diagonal(g, b, pin, K)
    assert np.all(d > 0)
    ref = O.gsrb_smooth_3d(g, b, pin, x0, rhs, K, 1)
    orders = {
        "natural": lambda c: c,
        "reversed": lambda c: c[::-1],
        "shuffled": lambda c: rng.permutation(c),
    }
    for name, order in orders.items():
        x = _emulate(g, b, pin, K, rhs, x0, red_black, order, d)
        assert rel_err(x, ref) < 1e-13, name
    assert rel_err(_jacobi_colours(g, b, pin, K, rhs, x0, red_black, d), ref) < 1e-13


@pytest.mark.parametrize("wrong", ["ignore_z", "ignore_y", "x_only"])
def test_planted_wrong_colouring_is_detected(wrong):
    """Colourings that drop the y or z parity do not reproduce the oracle's
    sweep, and their passes depend on the visit order."""
    g = O.grid(8, 6, 6)
    _, b = bc_pair("darcy_x")
    n = O.ncells(g)
    rng = np.random.default_rng(11)
    K = random_K(n, 4, 1.0)
    rhs = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    i, j, k = _cells(g)
    colour = {"ignore_z": (i + j) & 1, "ignore_y": (i + k) & 1, "x_only": i & 1}[wrong]
    d = _diagonal(g, b, False, K)
    ref = O.gsrb_smooth_3d(g, b, False, x0, rhs, K, 1)
    fwd = _emulate(g, b, False, K, rhs, x0, colour, lambda c: c, d)
    rev = _emulate(g, b, False, K, rhs, x0, colour, lambda c: c[::-1], d)
    assert rel_err(fwd, ref) > 1e-6
    assert rel_err(fwd, rev) > 1e-6
