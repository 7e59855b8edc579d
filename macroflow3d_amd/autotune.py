"""Autotuner of the solve's launch shapes (no counterpart in MacroFlow3D).

Every choice it makes leaves the results bitwise unchanged (each cell's
arithmetic is the same for every workgroup size, and the reductions keep their
fixed 256-thread first stage), so tuning only moves time:

* ``sweep_block``: threads per workgroup of the colour-split red-black sweeps
  (64, 128, 256, 512), timed on the finest level of the V-cycle in place;
* ``graph``: for batched solves, replaying the captured PCG iteration or
  launching its kernels one by one, timed on the whole batch.

Results are cached per device and problem shape in a JSON file (default
``~/.cache/macroflow3d_amd/autotune.json``, or ``$MF3D_AUTOTUNE_CACHE``);
reads and writes hold a process lock and an exclusive file lock.
"""
from __future__ import annotations

import fcntl
import json
import os
import threading
import time

import torch

from . import multigrid as mg
from . import physics as ph

BLOCKS = (64, 128, 256, 512)
_LOCK = threading.Lock()


def cache_path() -> str:
    return os.environ.get("MF3D_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "macroflow3d_amd",
                                       "autotune.json"))


def _device_name(ctx) -> str:
    return torch.cuda.get_device_name(ctx.device)


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            fcntl.flock(f, fcntl.LOCK_SH)
            try:
                return json.load(f)
            finally:
                fcntl.flock(f, fcntl.LOCK_UN)
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, value: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "a+") as f:
        fcntl.flock(f, fcntl.LOCK_EX)
        try:
            f.seek(0)
            try:
                data = json.load(f)
            except ValueError:
                data = {}
            data[key] = value
            f.seek(0)
            f.truncate()
            json.dump(data, f, indent=1, sort_keys=True)
        finally:
            fcntl.flock(f, fcntl.LOCK_UN)


def _key(kind: str, ctx, shapes, cfg) -> str:
    sh = ";".join("x".join(str(v) for v in s) for s in sorted(shapes))
    return (f"{kind}|{_device_name(ctx)}|{sh}|levels={cfg.mg_levels}|"
            f"prec={int(cfg.precond_precision)}|coef={int(cfg.coef_format)}")


def _time_events(ctx, fn, reps: int) -> float:
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()  # warm
    e0.record(ctx.stream)
    for _ in range(reps):
        fn()
    e1.record(ctx.stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_sweep_block(ctx, grid, bc, cfg, K, reps: int = 10, use_cache: bool = True) -> dict:
    """Fastest sweep workgroup size for this grid (finest-level smoothing, in place)."""
    key = _key("sweep", ctx, [(grid.nx, grid.ny, grid.nz)], cfg)
    path = cache_path()
    with _LOCK:
        if use_cache:
            hit = _load(path).get(key)
            if hit:
                return hit
        times = {}
        for b in BLOCKS:
            ws = ph.FlowWorkspace()
            ws.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, b)
            mg.mg_setup(ctx, ws, grid, bc, cfg, K)
            times[b] = _time_events(ctx, lambda: mg.smooth_level(ctx, ws, 0, 1), reps)
            ws.close()
        best = min(times, key=times.get)
        out = {"sweep_block": best, "ms_per_iteration": {str(b): t for b, t in times.items()}}
        _store(path, key, out)
        return out


def tune_batch(ctx, hs, Ks, grids, bc, cfg, workspaces, use_cache: bool = True) -> dict:
    """Fastest (sweep_block, graph) for a batch: one timed batched solve per candidate."""
    key = _key("batch", ctx, [(g.nx, g.ny, g.nz) for g in grids], cfg)
    path = cache_path()
    with _LOCK:
        if use_cache:
            hit = _load(path).get(key)
            if hit:
                return hit
        times = {}
        for b in (128, 256, 512):
            for graph in (True, False):
                for w in workspaces:
                    w.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, b)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, workspaces, use_graph=graph)
                times[(b, graph)] = (time.perf_counter() - t0) * 1e3
        (b, graph) = min(times, key=times.get)
        out = {"sweep_block": b, "graph": graph,
               "ms": {f"{k[0]}/{'graph' if k[1] else 'launches'}": v for k, v in times.items()}}
        _store(path, key, out)
        return out


def apply(workspaces, params: dict) -> None:
    """Sets the tuned workgroup size on the workspaces (takes effect at set-up)."""
    for w in workspaces:
        w.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, int(params["sweep_block"]))
