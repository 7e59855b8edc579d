"""Learned warm start for the head solve (no counterpart in MacroFlow3D).

A 3-D convolutional encoder-decoder (3 levels, 16/32/64 channels, 3^3
kernels) maps ln K to an initial head for the Darcy set-up with Dirichlet x
faces: h0 = h_lin + J trilinear(y(ln K)), h_lin linear between the x faces and
J = (h_W - h_E) / nx, so one network serves any grid size. The network sees
block statistics of K on the grid coarsened 4x per axis and predicts the
smooth part of the head there: MG-PCG removes high-frequency error in one
V-cycle, so only the smooth part of a guess shortens the solve. Forward, backward
and the Adam step run in libmf3d_hip.so (BF16 MFMA convolutions with FP32
accumulation, FP32 master weights); this module holds the parameters, makes
training data with the plain FP64 MG-PCG solve and drives the training.

The solve keeps its own guarantees: the prediction is only the initial
iterate (``InitGuess.PROVIDED``), PCG still runs to ``||b - A h|| <= rtol ||b||``
decided on the true residual, and a zero guess remains the fallback.

Command line (the GPU runs; writes FP32 parameters as .npy):
    python -m macroflow3d_amd.warmstart train --out weights.npy [--train 2000 --held 200]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np
import torch

from . import _lib
from . import physics as ph
from .core import BCSpec, Grid3D
from .kfield import lognormal_lnK
from .runtime import HipContext, dptr

LAYERS = [(4, 16), (16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 1)]
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "warmstart_128.npy")


def param_count() -> int:
    return int(_lib.lib().mf3d_wsnet_param_count())


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / (27 cin))), zero biases; the output layer
    scaled by 0.1 so an untrained network predicts close to the linear head."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (ci, co) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (27 * ci)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        parts.append((rng.standard_normal(co * 27 * ci) * std).astype(np.float32))
        parts.append(np.zeros(co, dtype=np.float32))
    return np.concatenate(parts)


class WarmStartNet:
    """Device copy of the network (parameters, Adam state, activations)."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        h = C.c_void_p()
        _lib.call("mf3d_wsnet_create", C.byref(h))
        self._h = h
        self.n = param_count()
        self.set_params(init_params(seed) if params is None else params)

    @classmethod
    def load(cls, path: str | None = None) -> "WarmStartNet":
        """Network with the parameters of path (default: the shipped 128^3 training)."""
        return cls(np.load(path or DEFAULT_WEIGHTS, allow_pickle=False))

    def save(self, path: str) -> None:
        np.save(path, self.get_params())

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n:
            raise ValueError(f"expected {self.n} parameters, got {p.size}")
        _lib.call("mf3d_wsnet_set_params", self._h, p.ctypes.data, self.n)

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n, dtype=np.float32)
        _lib.call("mf3d_wsnet_get_params", self._h, p.ctypes.data, self.n)
        return p

    def get_grads(self) -> np.ndarray:
        g = np.empty(self.n, dtype=np.float32)
        _lib.call("mf3d_wsnet_get_grads", self._h, g.ctypes.data, self.n)
        return g

    def predict(self, ctx: HipContext, K: torch.Tensor, grid: Grid3D, bc: BCSpec,
                h0: torch.Tensor, k_is_log: bool = False) -> None:
        """h0 <- network prediction (enqueued on ctx's stream)."""
        n = grid.num_cells()
        _lib.call("mf3d_wsnet_predict", self._h, ctx.handle, C.byref(grid.c()),
                  C.byref(bc.c()), dptr(K, n, "K", ctx), 1 if k_is_log else 0,
                  dptr(h0, n, "h0", ctx), n)

    def gradients(self, ctx: HipContext, K: torch.Tensor, h: torch.Tensor, grid: Grid3D,
                  bc: BCSpec, k_is_log: bool = False) -> float:
        n = grid.num_cells()
        loss = C.c_double()
        _lib.call("mf3d_wsnet_gradients", self._h, ctx.handle, C.byref(grid.c()),
                  C.byref(bc.c()), dptr(K, n, "K", ctx), 1 if k_is_log else 0,
                  dptr(h, n, "h", ctx), n, C.byref(loss))
        return loss.value

    def train_step(self, ctx: HipContext, K: torch.Tensor, h: torch.Tensor, grid: Grid3D,
                   bc: BCSpec, lr: float, k_is_log: bool = False, want_loss: bool = True) -> float:
        n = grid.num_cells()
        loss = C.c_double()
        _lib.call("mf3d_wsnet_train_step", self._h, ctx.handle, C.byref(grid.c()),
                  C.byref(bc.c()), dptr(K, n, "K", ctx), 1 if k_is_log else 0,
                  dptr(h, n, "h", ctx), n, float(lr), C.byref(loss) if want_loss else None)
        return loss.value if want_loss else float("nan")

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().mf3d_wsnet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- training data ------------------------------------------------------------
def field_params(seed: int, count: int, sigma2=(2.0, 4.0), corr=(4.0, 12.0)):
    """(sigma2, correlation length, field seed) per realisation, from a seeded
    host generator; field seed = seed + r."""
    rng = np.random.default_rng(seed)
    s2 = rng.uniform(*sigma2, count)
    lam = rng.uniform(*corr, count)
    return [(float(s2[r]), float(lam[r]), seed + r) for r in range(count)]


def make_samples(ctx: HipContext, n: int, params, levels: int = 5, rtol: float = 1e-10,
                 log=None):
    """lnK and the plain FP64 MG-PCG head for each (sigma2, corr, seed); device tensors."""
    g = Grid3D(n, n, n)
    bc = BCSpec.darcy_x()
    cfg = ph.HeadSolveConfig.fp64_direct(rtol, levels)
    ws = ph.FlowWorkspace()
    out = []
    t0 = time.perf_counter()
    for r, (s2, lam, seed) in enumerate(params):
        lnK, _ = lognormal_lnK(n, n, n, s2, lam, seed, device=ctx.device)
        h = ctx.zeros(g.num_cells())
        res = ph.solve_head(h, torch.exp(lnK), g, bc, cfg, ctx, ws)
        if not res.converged:
            raise RuntimeError(f"training solve {r} did not converge")
        out.append((lnK, h))
        if log and (r + 1) % 100 == 0:
            log(f"samples {r + 1}/{len(params)} ({time.perf_counter() - t0:.1f} s)")
    ws.close()
    return out


def augment(t: torch.Tensor, n: int, code: int) -> torch.Tensor:
    """One of the 8 symmetries of the Darcy set-up on an n^3 field (x fastest):
    bit 0 mirrors y, bit 1 mirrors z, bit 2 swaps y and z (the x faces carry
    the head drop, the y and z faces are alike)."""
    if code == 0:
        return t
    v = t.view(n, n, n)  # (z, y, x)
    if code & 1:
        v = torch.flip(v, dims=[1])
    if code & 2:
        v = torch.flip(v, dims=[0])
    if code & 4:
        v = v.transpose(0, 1)
    return v.contiguous().view(-1)


def train(ctx: HipContext, net: WarmStartNet, samples, n: int, epochs: int, lr: float,
          lr_final: float, log=None, seed: int = 0, time_budget: float = float("inf"),
          augment_data: bool = True):
    """Adam over the samples (one field per step, shuffled each epoch, each
    step under a random symmetry of the set-up when augment_data, learning
    rate decayed geometrically from lr to lr_final). Stops after the epoch in
    which time_budget seconds have passed. Returns mean sampled loss per epoch."""
    t_start = time.perf_counter()
    g = Grid3D(n, n, n)
    bc = BCSpec.darcy_x()
    rng = np.random.default_rng(seed)
    total = epochs * len(samples)
    hist = []
    step = 0
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        acc = 0.0
        t0 = time.perf_counter()
        for q, r in enumerate(order):
            rate = lr * (lr_final / lr) ** (step / max(1, total - 1))
            want = (q % 50 == 0)
            code = int(rng.integers(8)) if augment_data else 0
            lnK = augment(samples[r][0], n, code)
            h = augment(samples[r][1], n, code)
            loss = net.train_step(ctx, lnK, h, g, bc, rate, k_is_log=True, want_loss=want)
            if want:
                acc += loss
            step += 1
            if log and want and q % 500 == 0:
                log(f"  step {q}/{len(samples)} loss {loss:.4e} lr {rate:.2e}")
        hist.append(acc / max(1, (len(samples) + 49) // 50))
        if log:
            log(f"epoch {ep + 1}/{epochs}: sampled loss {hist[-1]:.4e} "
                f"({time.perf_counter() - t0:.1f} s)")
        if time.perf_counter() - t_start > time_budget:
            break
    return hist


def compare_guesses(ctx: HipContext, net: WarmStartNet, samples, n: int, levels: int = 5,
                    rtol: float = 1e-10, precond=ph.PrecondPrecision.FP64):
    """PCG iterations and times to ||b - A h|| <= rtol ||b|| from the zero, the
    linear and the predicted initial head, per sample."""
    g = Grid3D(n, n, n)
    bc = BCSpec.darcy_x()
    ws = ph.FlowWorkspace()
    rows = []
    for lnK, _ in samples:
        K = torch.exp(lnK)
        row = {}
        for name in ("zero", "linear", "predicted"):
            cfg = ph.HeadSolveConfig.fp64_direct(rtol, levels)
            cfg.precond_precision = precond
            h = ctx.zeros(g.num_cells())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "linear":
                cfg.init_guess = ph.InitGuess.REFERENCE
            elif name == "predicted":
                net.predict(ctx, lnK, g, bc, h, k_is_log=True)
                cfg.init_guess = ph.InitGuess.PROVIDED
            res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
            ms = (time.perf_counter() - t0) * 1e3
            row[name] = {"iterations": res.num_iterations, "converged": res.converged,
                         "rel_true_residual": res.true_residual / res.bnorm, "ms": ms,
                         "initial_rel_residual": res.initial_residual / res.bnorm}
        rows.append(row)
    ws.close()
    return rows


def wall_time_gain(ctx: HipContext, net: WarmStartNet, n: int, params, levels: int,
                   rtol: float = 1e-10, reps: int = 3, precond=ph.PrecondPrecision.FP32):
    """Per field: iterations and median wall time (reps timed runs after one
    untimed) of the solve from the linear head, and of prediction + solve from
    the predicted head (mixed-precision solve by default, as in the headline
    step). Fields are generated on the device from params."""
    g = Grid3D(n, n, n)
    bc = BCSpec.darcy_x()
    ws = ph.FlowWorkspace()
    h = ctx.zeros(g.num_cells())
    rows = []
    for s2, lam, seed in params:
        lnK, _ = lognormal_lnK(n, n, n, s2, lam, seed, device=ctx.device)
        K = torch.exp(lnK)
        row = {}
        for name in ("linear", "predicted"):
            cfg = ph.HeadSolveConfig.fp64_direct(rtol, levels)
            cfg.precond_precision = precond
            if precond == ph.PrecondPrecision.FP32:
                cfg = ph.HeadSolveConfig.mixed(rtol, levels)
            cfg.init_guess = (ph.InitGuess.REFERENCE if name == "linear"
                              else ph.InitGuess.PROVIDED)

            def run():
                if name == "predicted":
                    net.predict(ctx, K, g, bc, h)
                return ph.solve_head(h, K, g, bc, cfg, ctx, ws)
            run()
            times = []
            for _ in range(max(1, reps)):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                res = run()
                torch.cuda.synchronize()
                times.append((time.perf_counter() - t0) * 1e3)
            row[name] = {"iterations": res.num_iterations, "ms": float(np.median(times)),
                         "converged": bool(res.converged and res.true_residual <= rtol * res.bnorm)}
        rows.append(row)
        del lnK, K
    ws.close()
    return rows


def _main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("compare", help="wall time from the predicted vs the linear head")
    c.add_argument("--sizes", type=int, nargs="+", default=[128, 256, 512])
    c.add_argument("--fields", type=int, default=4)
    c.add_argument("--seed", type=int, default=88_000_000)
    c.add_argument("--weights", default=None)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--n", type=int, default=128)
    t.add_argument("--train", type=int, default=2000)
    t.add_argument("--held", type=int, default=200)
    t.add_argument("--epochs", type=int, default=4)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--lr-final", type=float, default=1e-4)
    t.add_argument("--seed", type=int, default=1000)
    t.add_argument("--report", default=None, help="JSON summary path")
    t.add_argument("--time-budget", type=float, default=float("inf"),
                   help="stop after the epoch that passes this many seconds of training")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    ctx = HipContext(0)
    if a.cmd == "compare":
        net = WarmStartNet.load(a.weights)
        out = {}
        for n in a.sizes:
            levels = {128: 5, 256: 6, 512: 7}.get(n, 5)
            rows = wall_time_gain(ctx, net, n, field_params(a.seed + n, a.fields), levels)
            out[n] = {k: {"mean_iterations": float(np.mean([r[k]["iterations"] for r in rows])),
                          "mean_ms": float(np.mean([r[k]["ms"] for r in rows])),
                          "all_converged": all(r[k]["converged"] for r in rows)}
                      for k in ("linear", "predicted")}
            log(f"{n}^3: {json.dumps(out[n])}")
        log(json.dumps(out))
        net.close()
        return 0
    ptrain = field_params(a.seed, a.train)
    pheld = field_params(a.seed + 10 ** 6, a.held)
    log(f"making {a.train} training + {a.held} held-out fields at {a.n}^3")
    tr = make_samples(ctx, a.n, ptrain, log=log)
    net = WarmStartNet(seed=a.seed)
    hist = train(ctx, net, tr, a.n, a.epochs, a.lr, a.lr_final, log=log, seed=a.seed,
                 time_budget=a.time_budget)
    net.save(a.out)
    del tr
    held = make_samples(ctx, a.n, pheld, log=log)
    g = Grid3D(a.n, a.n, a.n)
    bc = BCSpec.darcy_x()
    held_loss = float(np.mean([net.gradients(ctx, lnK, h, g, bc, k_is_log=True)
                               for lnK, h in held]))
    rows = compare_guesses(ctx, net, held, a.n)
    summ = {k: {"mean_iterations": float(np.mean([r[k]["iterations"] for r in rows])),
                "mean_ms": float(np.mean([r[k]["ms"] for r in rows])),
                "all_converged": all(r[k]["converged"] for r in rows),
                "max_rel_true_residual": max(r[k]["rel_true_residual"] for r in rows)}
            for k in ("zero", "linear", "predicted")}
    rep = {"n": a.n, "train": a.train, "held": a.held, "epochs": a.epochs, "lr": a.lr,
           "lr_final": a.lr_final, "seed": a.seed, "loss_per_epoch": hist,
           "held_loss": held_loss, "held_out": summ}
    log(json.dumps(rep))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(_main())
