"""mf3d_dot_noalloc, the dot product that allocates nothing (its partial sums
sit in a fixed device array inside the library), against mf3d_dot with a
workspace, both called through the C ABI: bitwise equal results on calls one
after another with shrinking and growing sizes (the array keeps the previous
call's partials; every partial a call reads is rewritten by that call), at the
maximum size, and a clear error beyond it."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from macroflow3d_amd import _lib
from macroflow3d_amd.numerics import ReductionWorkspace

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 27


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _noalloc(ctx, x, y, n, out):
    return _lib.lib().mf3d_dot_noalloc(ctx.handle, _ptr(x), _ptr(y), n, _ptr(out))


def _with_workspace(ctx, red, x, y, n, out):
    _lib.call("mf3d_dot", red.ws.handle, ctx.handle, _ptr(x), _ptr(y), n, _ptr(out))


def test_noalloc_dot_matches_workspace_dot(ctx):
    red = ReductionWorkspace()
    g = torch.Generator(device="cpu").manual_seed(5)
    for n in (5_000_003, 100_003, 1, 2048, 4_194_305):  # second call smaller, then mixed
        print(f"noalloc dot: n = {n}", flush=True)
        xh = torch.randn(n, dtype=torch.float64, generator=g)
        yh = torch.randn(n, dtype=torch.float64, generator=g)
        x, y = xh.to(ctx.device), yh.to(ctx.device)
        a = torch.full((1,), np.nan, dtype=torch.float64, device=ctx.device)
        b = torch.full((1,), np.nan, dtype=torch.float64, device=ctx.device)
        assert _noalloc(ctx, x, y, n, a) == 0, _lib.lib().mf3d_last_error()
        _with_workspace(ctx, red, x, y, n, b)
        ctx.synchronize()
        assert a.item() == b.item(), (n, a.item(), b.item())
        exact = float(np.dot(xh.numpy().astype(np.longdouble), yh.numpy().astype(np.longdouble)))
        assert abs(a.item() - exact) <= 1e-13 * float(np.abs(xh.numpy()) @ np.abs(yh.numpy()))
    # n = 0: the reference's dot of an empty vector is 0
    z = torch.full((1,), np.nan, dtype=torch.float64, device=ctx.device)
    assert _noalloc(ctx, x, y, 0, z) == 0
    ctx.synchronize()
    assert z.item() == 0.0


def test_noalloc_dot_at_and_beyond_its_maximum(ctx):
    print("noalloc dot: n = 2^27 (maximum)", flush=True)
    x = torch.empty(MAX_CELLS + 1, dtype=torch.float64, device=ctx.device)
    x.uniform_(-1.0, 1.0)
    y = torch.empty_like(x)
    y.uniform_(-1.0, 1.0)
    a = torch.zeros(1, dtype=torch.float64, device=ctx.device)
    b = torch.zeros(1, dtype=torch.float64, device=ctx.device)
    assert _noalloc(ctx, x, y, MAX_CELLS, a) == 0, _lib.lib().mf3d_last_error()
    _with_workspace(ctx, ReductionWorkspace(), x, y, MAX_CELLS, b)
    ctx.synchronize()
    assert a.item() == b.item()
    # one element more is refused before any work is queued
    a.fill_(3.0)
    st = _noalloc(ctx, x, y, MAX_CELLS + 1, a)
    assert st == 1  # MF3D_ERR_INVALID_ARGUMENT
    assert b"exceeds the maximum" in _lib.lib().mf3d_last_error()
    ctx.synchronize()
    assert a.item() == 3.0
    del x, y
    torch.cuda.empty_cache()
