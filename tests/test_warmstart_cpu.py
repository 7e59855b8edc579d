"""Host-side pieces of the warm start: the symmetry augmentation maps a head
solution to the head solution of the transformed field (checked with the CPU
oracle on a small grid), and the parameter layout matches the oracle's."""
from __future__ import annotations

import numpy as np
import torch

import oracle as O
from macroflow3d_amd import warmstart as W


def test_augmentations_map_solutions_to_solutions():
    n = 8
    lnK = np.random.default_rng(3).standard_normal(n ** 3)
    K = np.exp(lnK)
    g, bc = O.grid(n, n, n), O.darcy_bc()
    p = O.mode_n(mg_levels=2, cg_rtol=1e-13)
    h, _ = O.solve_head(g, bc, p, K)
    for code in range(8):
        Kt = W.augment(torch.as_tensor(K), n, code).numpy()
        ht = W.augment(torch.as_tensor(h), n, code).numpy()
        h2, _ = O.solve_head(g, bc, p, Kt)
        assert np.abs(h2 - ht).max() < 1e-10, code


def test_init_params_shape():
    p = W.init_params(0)
    assert p.dtype == np.float32 and p.size == sum(27 * a * b + b for a, b in W.LAYERS)
