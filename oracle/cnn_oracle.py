"""CPU FP64 oracle of the learned warm-start network (TEST INFRASTRUCTURE ONLY).

numpy restatement of the encoder-decoder that libmf3d_hip runs on BF16 MFMA
kernels (macroflow3d_amd/csrc/cnn.hip); it has no counterpart in MacroFlow3D.
Only tests/ may import it, to check the GPU forward and backward passes
against. With ``bf16=True`` every tensor the GPU stores in BF16 (inputs,
weights, activations, activation gradients) is rounded to BF16 at the same
place, so the remaining difference is FP32-vs-FP64 accumulation; with
``bf16=False`` it is the exact FP64 function, whose gradient the tests pin
by central finite differences.

Tensors are channels-last arrays of shape (nz, ny, nx, C); convolution taps
are ordered tap = (dx+1) + 3 (dy+1) + 9 (dz+1), weights w[cout][27][cin],
zero padding, cross-correlation: y[v][co] = sum w[co][tap][ci] x[v + d(tap)][ci].
"""
from __future__ import annotations

import numpy as np

LAYERS = [(4, 16), (16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 1)]
S = 4  # fine cells per network cell, per axis
TAPS = [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]


def param_count() -> int:
    return sum(27 * ci * co + co for ci, co in LAYERS)


def unflatten(p: np.ndarray):
    """[(w (cout, 27, cin), b (cout,)), ...] from the flat parameter vector."""
    out, o = [], 0
    for ci, co in LAYERS:
        w = p[o:o + 27 * ci * co].reshape(co, 27, ci)
        o += 27 * ci * co
        b = p[o:o + co]
        o += co
        out.append((w, b))
    return out


def flatten(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(x, dx, dy, dz):
    """s[k, j, i] = x[k + dz, j + dy, i + dx] with zeros outside."""
    nz, ny, nx = x.shape[:3]
    s = np.zeros_like(x)
    zs, ze = max(0, -dz), min(nz, nz - dz)
    ys, ye = max(0, -dy), min(ny, ny - dy)
    xs, xe = max(0, -dx), min(nx, nx - dx)
    s[zs:ze, ys:ye, xs:xe] = x[zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return s


def conv(x, w, b=None):
    """y = conv(x, w) + b; x (nz, ny, nx, cin), w (cout, 27, cin)."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dx, dy, dz) in enumerate(TAPS):
        y += _shift(x, dx, dy, dz) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv_bwd_data(dy, w):
    """dx = sum_tap shift(dy, -d) w[:, tap, :]."""
    dx_ = np.zeros(dy.shape[:3] + (w.shape[2],))
    for t, (dx, dy_, dz) in enumerate(TAPS):
        dx_ += _shift(dy, -dx, -dy_, -dz) @ w[:, t, :]
    return dx_


def conv_bwd_weights(x, dy):
    """gw (cout, 27, cin), gb (cout,)."""
    co, ci = dy.shape[3], x.shape[3]
    gw = np.zeros((co, 27, ci))
    d2 = dy.reshape(-1, co)
    for t, (dx, dy_, dz) in enumerate(TAPS):
        gw[:, t, :] = d2.T @ _shift(x, dx, dy_, dz).reshape(-1, ci)
    return gw, d2.sum(axis=0)


F = 2  # pooling factor per axis between network levels


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // F, F, ny // F, F, nx // F, F, c).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(F, axis=0).repeat(F, axis=1).repeat(F, axis=2)


def pool_bwd(dy):
    return up(dy) / float(F ** 3)


def up_bwd(dy):
    nz, ny, nx, c = dy.shape
    return dy.reshape(nz // F, F, ny // F, F, nx // F, F, c).sum(axis=(1, 3, 5))


def block_mean(a):
    """Mean over the S^3 fine blocks of a fine (nz, ny, nx) array."""
    nz, ny, nx = a.shape
    return a.reshape(nz // S, S, ny // S, S, nx // S, S).mean(axis=(1, 3, 5))


def inputs(lnK):
    """x0 on the coarse grid: [<lnK>/2, log<K>/2, -log<1/K>/2, 2 (I + 1/2)/NX - 1]."""
    K = np.exp(lnK)
    a = 0.5 * block_mean(lnK)
    b = 0.5 * np.log(block_mean(K))
    c = -0.5 * np.log(block_mean(1.0 / K))
    NX = a.shape[2]
    xc = 2.0 * (np.arange(NX) + 0.5) / NX - 1.0
    return np.stack([a, b, c, np.broadcast_to(xc, a.shape)], axis=-1)


def target(h, hw, he):
    """Block mean of (h - h_lin) / J, J = (hw - he) / nx (coarse grid)."""
    nx = h.shape[2]
    hl = hw + (he - hw) * (np.arange(nx) + 0.5) / nx
    return block_mean((h - hl) / ((hw - he) / nx))


def _lin_axis(n_fine, n):
    u = (np.arange(n_fine) + 0.5) / S - 0.5
    i0 = np.clip(np.floor(u), 0, n - 1).astype(int)
    i1 = np.minimum(i0 + 1, n - 1)
    w = np.where((u <= 0) | (u >= n - 1), 0.0, u - np.floor(u))
    return i0, i1, w


def output(y, hw, he, fine_shape):
    """h0 = h_lin + J trilinear(y) on the fine grid."""
    nz, ny, nx = fine_shape
    Z, Y, X = y.shape
    x0, x1, wx = _lin_axis(nx, X)
    y0, y1, wy = _lin_axis(ny, Y)
    z0, z1, wz = _lin_axis(nz, Z)
    a = y[:, :, x0] * (1 - wx) + y[:, :, x1] * wx
    b = a[:, y0, :] * (1 - wy)[:, None] + a[:, y1, :] * wy[:, None]
    c = b[z0] * (1 - wz)[:, None, None] + b[z1] * wz[:, None, None]
    J = (hw - he) / nx
    return hw + (he - hw) * (np.arange(nx) + 0.5) / nx + J * c


def forward(p, lnK, rb=False):
    """Network output y on the coarse grid and the activations, for flat params
    p and the fine (nz, ny, nx) ln K."""
    r = bf16 if rb else (lambda a: a)
    L = [(r(w), b) for w, b in unflatten(np.asarray(p, dtype=np.float64))]
    A = {"x0": r(inputs(lnK))}
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    A["e0"] = r(relu(conv(A["x0"], *L[0])))
    A["p1"] = r(pool(A["e0"]))
    A["e1"] = r(relu(conv(A["p1"], *L[1])))
    A["p2"] = r(pool(A["e1"]))
    A["e2"] = r(relu(conv(A["p2"], *L[2])))
    A["e2b"] = r(relu(conv(A["e2"], *L[3])))
    A["u1"] = up(A["e2b"])
    A["d1"] = r(relu(conv(A["u1"], *L[4]) + A["e1"]))
    A["u0"] = up(A["d1"])
    A["d0"] = r(relu(conv(A["u0"], *L[5]) + A["e0"]))
    A["y"] = conv(A["d0"], *L[6])[..., 0]
    A["L"] = L
    return A["y"], A


def loss_and_grads(p, lnK, h, hw, he, rb=False):
    """mean over the coarse grid of (y - t)^2 and the flat gradient, t the
    block mean of (h - h_lin)/J."""
    r = bf16 if rb else (lambda a: a)
    y, A = forward(p, lnK, rb)
    L = A["L"]
    t = target(h, hw, he)
    n = y.size
    loss = float(((y - t) ** 2).mean())
    g = [None] * 7
    mask = lambda a: (a > 0).astype(np.float64)  # noqa: E731
    dOut = r((2.0 * (y - t) / n)[..., None])
    g[6] = conv_bwd_weights(A["d0"], dOut)
    dpre_f = r(conv_bwd_data(dOut, L[6][0]) * mask(A["d0"]))
    g[5] = conv_bwd_weights(A["u0"], dpre_f)
    du0 = r(conv_bwd_data(dpre_f, L[5][0]))
    dpre_e = r(up_bwd(du0) * mask(A["d1"]))
    g[4] = conv_bwd_weights(A["u1"], dpre_e)
    du1 = r(conv_bwd_data(dpre_e, L[4][0]))
    dpre_d = r(up_bwd(du1) * mask(A["e2b"]))
    g[3] = conv_bwd_weights(A["e2"], dpre_d)
    dpre_c = r(conv_bwd_data(dpre_d, L[3][0]) * mask(A["e2"]))
    g[2] = conv_bwd_weights(A["p2"], dpre_c)
    dp2 = r(conv_bwd_data(dpre_c, L[2][0]))
    dpre_b = r((pool_bwd(dp2) + dpre_e) * mask(A["e1"]))
    g[1] = conv_bwd_weights(A["p1"], dpre_b)
    dp1 = r(conv_bwd_data(dpre_b, L[1][0]))
    dpre_a = r((pool_bwd(dp1) + dpre_f) * mask(A["e0"]))
    g[0] = conv_bwd_weights(A["x0"], dpre_a)
    return loss, flatten(g)


def init_params(seed: int) -> np.ndarray:
    """He-normal weights (std sqrt(2 / (27 cin))), zero biases; the last layer
    scaled by 0.1 so the initial prediction stays near the linear head."""
    rng = np.random.default_rng(seed)
    layers = []
    for li, (ci, co) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (27 * ci)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        layers.append((rng.standard_normal((co, 27, ci)) * std, np.zeros(co)))
    return flatten(layers)
