"""Pins the CPU oracle of the warm-start network (oracle/cnn_oracle.py):
brute-force convolution, closed forms (identity and shift kernels),
adjoint identities of every backward routine, BF16 rounding against torch's,
and the full network gradient against central finite differences."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cnn_oracle as CO


def _brute_conv(x, w, b):
    nz, ny, nx, ci = x.shape
    co = w.shape[0]
    y = np.zeros((nz, ny, nx, co))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for t, (dx, dy, dz) in enumerate(CO.TAPS):
                    ii, jj, kk = i + dx, j + dy, k + dz
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        y[k, j, i] += w[:, t, :] @ x[kk, jj, ii]
                y[k, j, i] += b
    return y


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 5, 2))
    w = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    assert np.allclose(CO.conv(x, w, b), _brute_conv(x, w, b), rtol=1e-13, atol=1e-13)


def test_conv_identity_and_shift_kernels():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 4, 4, 3))
    w = np.zeros((3, 27, 3))
    for c in range(3):
        w[c, 13, c] = 1.0  # centre tap
    assert np.array_equal(CO.conv(x, w), x)
    w = np.zeros((3, 27, 3))
    for c in range(3):
        w[c, 14, c] = 1.0  # tap 14 = (dx=+1, 0, 0): y[i] = x[i+1]
    y = CO.conv(x, w)
    assert np.array_equal(y[:, :, :-1], x[:, :, 1:]) and np.all(y[:, :, -1] == 0)


def test_backward_routines_are_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 4, 6, 5))
    w = rng.standard_normal((7, 27, 5))
    dy = rng.standard_normal((4, 4, 6, 7))
    lhs = float((CO.conv(x, w) * dy).sum())
    assert np.isclose(lhs, float((x * CO.conv_bwd_data(dy, w)).sum()), rtol=1e-12)
    gw, gb = CO.conv_bwd_weights(x, dy)
    assert np.isclose(lhs, float((w * gw).sum()), rtol=1e-12)
    assert np.allclose(gb, dy.reshape(-1, 7).sum(axis=0))
    x = rng.standard_normal((4, 2, 6, 5))
    p = rng.standard_normal((2, 1, 3, 5))
    assert np.isclose(float((CO.pool(x) * p).sum()), float((x * CO.pool_bwd(p)).sum()))
    assert np.isclose(float((CO.up(p) * x).sum()), float((p * CO.up_bwd(x)).sum()))


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    v = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.0078125, 1.01171875, 3.0e38]])
    ref = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(CO.bf16(v), ref)


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    n = 32
    lnK = rng.standard_normal((n, n, n))
    h = rng.standard_normal((n, n, n)) * 0.01 + np.linspace(1, 0, n)
    p = CO.init_params(7)
    p[-1] = 0.05  # nonzero output bias
    _, g = CO.loss_and_grads(p, lnK, h, 1.0, 0.0)
    offs = np.cumsum([0] + [27 * ci * co + co for ci, co in CO.LAYERS])
    idx = []
    for li in range(len(CO.LAYERS)):
        idx += list(rng.integers(offs[li], offs[li + 1], 4))
        idx.append(offs[li + 1] - 1)  # a bias
    eps = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (CO.loss_and_grads(pp, lnK, h, 1.0, 0.0)[0] -
              CO.loss_and_grads(pm, lnK, h, 1.0, 0.0)[0]) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-5 * max(1e-3, abs(fd)) + 1e-9, (q, fd, g[q])


def test_target_and_output_closed_forms():
    """A head linear in x plus a constant offset c gives the target c / J on
    every block, and the trilinear output of a constant y is h_lin + J y; a y
    linear in the coarse x index is reproduced exactly between the outer centres."""
    nx = 16
    hl = 2.0 + (-3.0) * (np.arange(nx) + 0.5) / nx
    h = np.broadcast_to(hl, (8, 4, nx)) + 0.3
    J = 3.0 / nx
    assert np.allclose(CO.target(h, 2.0, -1.0), 0.3 / J)
    y = np.full((2, 1, 4), 1.7)
    assert np.allclose(CO.output(y, 2.0, -1.0, (8, 4, nx)), hl + J * 1.7)
    y = np.broadcast_to(np.arange(4.0), (2, 1, 4))
    out = CO.output(y, 2.0, -1.0, (8, 4, nx))
    u = (np.arange(nx) + 0.5) / 4 - 0.5
    inner = (u >= 0) & (u <= 3)
    assert np.allclose(((out - hl) / J)[..., inner], u[inner])


def test_inputs_block_statistics():
    rng = np.random.default_rng(8)
    lnK = rng.standard_normal((8, 8, 8))
    x0 = CO.inputs(lnK)
    blk = lnK[:4, :4, 4:8]
    K = np.exp(blk)
    assert np.isclose(x0[0, 0, 1, 0], 0.5 * blk.mean())
    assert np.isclose(x0[0, 0, 1, 1], 0.5 * np.log(K.mean()))
    assert np.isclose(x0[0, 0, 1, 2], -0.5 * np.log((1 / K).mean()))
    assert x0[0, 0, 1, 1] >= x0[0, 0, 1, 0] >= x0[0, 0, 1, 2]  # arithmetic >= geometric >= harmonic
    assert np.isclose(x0[0, 0, 1, 3], 2 * 1.5 / 2 - 1)
