"""Learned warm start: the BF16 MFMA convolution kernels and the full
network's gradients against the CPU FP64 oracle (oracle/cnn_oracle.py, with
BF16 rounding at the same places), training, and the gain of the trained
network over the zero and linear initial heads (true residual <= 1e-10)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest
import torch

from macroflow3d_amd import _lib
from macroflow3d_amd import warmstart as W
from macroflow3d_amd.core import BCSpec, Grid3D
from oracle import cnn_oracle as CO

gpu = pytest.mark.gpu


def test_param_layout_matches_oracle():
    assert W.param_count() == CO.param_count() == sum(27 * a * b + b for a, b in W.LAYERS)
    assert W.LAYERS == CO.LAYERS


def _bf(a, ctx):
    """numpy (nz, ny, nx, C) -> device BF16 channels-last tensor (rounded)."""
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32).to(torch.bfloat16).to(ctx.device)


def _host(t):
    torch.cuda.synchronize()
    return t.float().cpu().numpy().astype(np.float64)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


SHAPES = [((8, 8, 8), 16, 32), ((12, 8, 4), 8, 16), ((16, 16, 8), 32, 64), ((8, 4, 20), 64, 16)]


@gpu
@pytest.mark.parametrize("shape,cin,cout", SHAPES)
@pytest.mark.parametrize("relu,res", [(False, False), (True, True)])
def test_conv_forward_matches_oracle(ctx, shape, cin, cout, relu, res):
    nx, ny, nz = shape
    rng = np.random.default_rng(cin * 7 + cout)
    x = CO.bf16(rng.standard_normal((nz, ny, nx, cin)))
    w = rng.standard_normal((cout, 27, cin)) * 0.2
    b = rng.standard_normal(cout) * 0.1
    R = CO.bf16(rng.standard_normal((nz, ny, nx, cout))) if res else None
    ref = CO.conv(x, CO.bf16(w), b) + (R if res else 0.0)
    if relu:
        ref = np.maximum(ref, 0.0)
    X = _bf(x, ctx)
    Y = torch.empty((nz, ny, nx, cout), dtype=torch.bfloat16, device=ctx.device)
    wd = torch.as_tensor(w, dtype=torch.float32).to(ctx.device)
    bd = torch.as_tensor(b, dtype=torch.float32).to(ctx.device)
    Rd = _bf(R, ctx) if res else None
    g = Grid3D(nx, ny, nz).c()
    print(f"conv forward {shape} {cin}->{cout} relu={relu} res={res}", flush=True)
    _lib.call("mf3d_conv3d_forward", ctx.handle, C.byref(g), cin, cout, _ptr(X), _ptr(wd),
              _ptr(bd), _ptr(Rd), 1 if relu else 0, _ptr(Y))
    got = _host(Y)
    assert _rel(got, ref) < 8e-3
    assert np.abs(got - ref).max() <= 0.02 * np.abs(ref).max()


@gpu
@pytest.mark.parametrize("shape,cin,cout", [((8, 8, 8), 16, 32), ((16, 8, 12), 64, 64),
                                            ((8, 12, 4), 32, 16)])
def test_conv_backward_matches_oracle(ctx, shape, cin, cout):
    nx, ny, nz = shape
    rng = np.random.default_rng(cin + 3 * cout)
    x = CO.bf16(np.maximum(rng.standard_normal((nz, ny, nx, cin)), 0.0))
    dy = CO.bf16(rng.standard_normal((nz, ny, nx, cout)))
    w = rng.standard_normal((cout, 27, cin)) * 0.2
    g = Grid3D(nx, ny, nz).c()
    wd = torch.as_tensor(w, dtype=torch.float32).to(ctx.device)
    X, dY = _bf(x, ctx), _bf(dy, ctx)
    dX = torch.empty((nz, ny, nx, cin), dtype=torch.bfloat16, device=ctx.device)
    print(f"conv backward {shape} {cin}->{cout}", flush=True)
    _lib.call("mf3d_conv3d_backward_data", ctx.handle, C.byref(g), cin, cout, _ptr(dY),
              _ptr(wd), _ptr(X), _ptr(dX))
    ref = CO.conv_bwd_data(dy, CO.bf16(w)) * (x > 0)
    assert _rel(_host(dX), ref) < 8e-3
    gw = torch.empty((cout, 27, cin), dtype=torch.float32, device=ctx.device)
    gb = torch.empty(cout, dtype=torch.float32, device=ctx.device)
    _lib.call("mf3d_conv3d_backward_weights", ctx.handle, C.byref(g), cin, cout, _ptr(X),
              _ptr(dY), _ptr(gw), _ptr(gb))
    rw, rb = CO.conv_bwd_weights(x, dy)
    assert _rel(_host(gw), rw) < 1e-4
    assert _rel(_host(gb), rb) < 1e-4


def _problem(n3, seed):
    nx, ny, nz = n3
    rng = np.random.default_rng(seed)
    lnK = rng.standard_normal((nz, ny, nx)) * 1.2
    h = 1.0 - (np.arange(nx) + 0.5) / nx + 0.05 * rng.standard_normal((nz, ny, nx))
    return lnK, h


@gpu
@pytest.mark.parametrize("n3", [(32, 32, 32), (64, 32, 48)])
def test_network_gradients_match_oracle(ctx, n3):
    nx, ny, nz = n3
    lnK, h = _problem(n3, 11)
    p = CO.init_params(3)
    p[-1] = 0.1
    loss_o, g_o = CO.loss_and_grads(p, lnK, h, 1.0, 0.0, rb=True)
    net = W.WarmStartNet(p.astype(np.float32))
    g = Grid3D(nx, ny, nz)
    bc = BCSpec.darcy_x()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a).ravel()).to(ctx.device)  # noqa: E731
    print(f"network gradients {n3}", flush=True)
    loss = net.gradients(ctx, dev(lnK), dev(h), g, bc, k_is_log=True)
    grads = net.get_grads().astype(np.float64)
    assert abs(loss - loss_o) <= 1e-3 * loss_o
    # prediction: h_lin + J trilinear(y)
    h0 = ctx.zeros(g.num_cells())
    net.predict(ctx, dev(lnK), g, bc, h0, k_is_log=True)
    y_o, _ = CO.forward(p, lnK, rb=True)
    ref = CO.output(y_o, 1.0, 0.0, (nz, ny, nx))
    assert np.abs(_host(h0).reshape(nz, ny, nx) - ref).max() < 2e-4 / nx
    off = 0
    for li, (ci, co) in enumerate(CO.LAYERS):
        nw = 27 * ci * co
        for name, sl in (("w", slice(off, off + nw)), ("b", slice(off + nw, off + nw + co))):
            assert _rel(grads[sl], g_o[sl]) < 2e-2, (li, name, _rel(grads[sl], g_o[sl]))
        off += nw + co
    net.close()


@gpu
def test_training_reduces_loss(ctx):
    """Adam on one field whose head correction is a smooth function of ln K:
    the loss falls well below its starting value."""
    n = 32
    g = Grid3D(n, n, n)
    bc = BCSpec.darcy_x()
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    lnK = np.sin(0.4 * x + 0.3 * y) * np.cos(0.35 * z) * 1.5
    hl = 1.0 - (np.arange(n) + 0.5) / n
    h = hl + (1.0 / n) * 0.5 * lnK
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a).ravel()).to(ctx.device)  # noqa: E731
    K, H = dev(lnK), dev(h)
    net = W.WarmStartNet(seed=5)
    first = net.train_step(ctx, K, H, g, bc, 3e-3, k_is_log=True)
    for _ in range(100):
        last = net.train_step(ctx, K, H, g, bc, 3e-3, k_is_log=True)
    assert np.isfinite(last) and last < 0.3 * first, (first, last)
    net.close()


@gpu
@pytest.mark.skipif(not os.path.exists(W.DEFAULT_WEIGHTS), reason="no trained weights")
def test_trained_network_beats_zero_and_linear_guess(ctx):
    """24 held-out 128^3 fields (sigma2 2-4, correlation 4-12 cells, seeds not
    used in training): PCG from the predicted head reaches ||b - A h|| <= 1e-10
    ||b|| on every field, in fewer iterations on average than from the zero or
    the linear head, and never in more than the linear head needs plus one."""
    net = W.WarmStartNet.load()
    params = W.field_params(77_000_000, 24)
    print("warm start held-out 128^3", flush=True)
    samples = W.make_samples(ctx, 128, params)
    rows = W.compare_guesses(ctx, net, samples, 128)
    for r in rows:
        for k in ("zero", "linear", "predicted"):
            assert r[k]["converged"] and r[k]["rel_true_residual"] <= 1e-10
        assert r["predicted"]["iterations"] <= r["linear"]["iterations"] + 1
    it = {k: np.mean([r[k]["iterations"] for r in rows]) for k in ("zero", "linear", "predicted")}
    print(f"mean PCG iterations: {it}", flush=True)
    assert it["predicted"] < it["linear"] < it["zero"], it
    net.close()


@gpu
@pytest.mark.skipif(not os.path.exists(W.DEFAULT_WEIGHTS), reason="no trained weights")
def test_warm_start_pays_in_wall_time_at_256cubed(ctx):
    """The headline step at 256^3 (mixed-precision solve to 1e-10): inference
    plus the solve from the predicted head takes less wall time, on average
    over held-out fields (sigma2 2-4, correlation 4-12 cells, seeds not used in
    training), than the solve from the linear head; every solve converges."""
    net = W.WarmStartNet.load()
    print("warm start wall time 256^3", flush=True)
    rows = W.wall_time_gain(ctx, net, 256, W.field_params(88_000_256, 4), levels=6, reps=5)
    for r in rows:
        assert r["linear"]["converged"] and r["predicted"]["converged"]
    it = {k: np.mean([r[k]["iterations"] for r in rows]) for k in ("linear", "predicted")}
    ms = {k: np.mean([r[k]["ms"] for r in rows]) for k in ("linear", "predicted")}
    print(f"256^3 iterations {it}, ms {ms}", flush=True)
    assert it["predicted"] < it["linear"], it
    assert ms["predicted"] < ms["linear"], ms
    net.close()
