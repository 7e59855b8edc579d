"""Log-normal K generator (FFT circulant embedding): device-independent
draws, agreement with the oracle's C++ generator, and the covariance it is
meant to have."""
from __future__ import annotations

import numpy as np
import torch

import oracle as O
from macroflow3d_amd import kfield


def test_draws_match_oracle_and_uint64_reference():
    k = torch.arange(0, 64, dtype=torch.int64)
    n1, n2 = kfield.gauss_pairs(12345, k)
    for i in (0, 1, 17, 63):
        o = O.gauss_pair(12345, i)
        assert abs(n1[i].item() - o[0]) < 1e-14 and abs(n2[i].item() - o[1]) < 1e-14
    # splitmix finaliser against exact Python integers
    z = 0x0123456789ABCDEF
    t = torch.tensor([kfield._s64(z)], dtype=torch.int64)
    assert (kfield._mix64(t).item() & ((1 << 64) - 1)) == kfield._mix64_int(z)


def test_field_matches_oracle_generator():
    Y, st = kfield.lognormal_lnK(24, 16, 20, 2.0, 3.0, 77)
    Yo, sto = O.lognormal_lnK(24, 16, 20, 2.0, 3.0, 77)
    assert tuple(st["embedding"]) == tuple(int(v) for v in sto[:3])
    assert np.abs(Y.numpy() - Yo).max() < 1e-12 * np.abs(Yo).max()


def test_exponential_covariance_recovered():
    sig2, lam = 1.0, 3.0
    acc = {L: [] for L in (0, 1, 2, 4, 6)}
    for seed in range(6):
        Y, st = kfield.lognormal_lnK(48, 48, 48, sig2, lam, seed)
        assert st["clipped_mass"] == 0.0
        cov = kfield.empirical_covariance(Y, (48, 48, 48), lags=tuple(acc))
        for L in acc:
            acc[L].append(cov[L])
    for L, vals in acc.items():
        target = sig2 * np.exp(-L / lam)
        assert abs(np.mean(vals) - target) < 0.12 * sig2, (L, np.mean(vals), target)


def test_seed_rule_gives_distinct_realisations():
    a, _ = kfield.lognormal_lnK(16, 16, 16, 1.0, 2.0, 100)
    b, _ = kfield.lognormal_lnK(16, 16, 16, 1.0, 2.0, 101)
    assert float(torch.corrcoef(torch.stack([a, b]))[0, 1]) < 0.3
