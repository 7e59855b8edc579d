"""Autotuner of the solve's launch shapes (no counterpart in MacroFlow3D).

Its choices change where and when the work runs, not what is computed:

* ``sweep_block``: threads per workgroup of the colour-split red-black sweeps
  of the mixed V-cycle (64, 128, 256, 512): a runtime launch parameter; each
  cell's arithmetic and every partial sum are the same for each size, so the
  head is bitwise unchanged;
* ``graph``: replaying one captured PCG iteration or launching its kernels
  one by one (single mixed-precision solves and batches): bitwise unchanged;
* ``op_tile``: the finest-level FP64 operator on plane tiles staged in LDS or
  on register cell pairs: the same faces and per-cell sums, but the (p, Ap)
  partials are grouped by a different tile geometry, so the head agrees to
  the rounding of that reduction (iteration counts may differ by one);
* ``op_chunk``: the operator's z-planes per workgroup (at most; fewer where
  the launch would have too few workgroups): the same per-cell sums, the
  (p, Ap) partials regrouped;
* ``march_blocks``: the workgroup count the FP64 V-cycle's colour-split
  sweeps aim for by shortening their chunks of planes: the same per-cell
  values, the (r, z) partials regrouped.

The FP64 V-cycle's sweeps also take ``sweep_block``; there the (r, z)
partials are one per workgroup, so a different size regroups that sum.

``tune_sweep_block`` times the finest-level smoothing in place;
``tune_solve`` times whole single solves over (sweep_block, graph, op_tile),
then, from the fastest of those, over ``op_chunk`` and, when the V-cycle is
FP64, ``march_blocks`` (one knob at a time); ``tune_batch`` times whole
batched solves over (sweep_block, graph). Each candidate gets one untimed
run, then the median of ``reps`` timed runs.

Results are cached per device, problem shape, solver configuration and
boundary types in a JSON file (default
``~/.cache/macroflow3d_amd/autotune.json``, or ``$MF3D_AUTOTUNE_CACHE``);
reads and writes hold a process lock and an exclusive file lock.
"""
from __future__ import annotations

import fcntl
import json
import os
import statistics
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
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
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
            f.flush()
        finally:
            fcntl.flock(f, fcntl.LOCK_UN)


def key_of(kind: str, device: str, shapes, cfg, bc) -> str:
    """Cache key: device, shapes, the configuration fields that change the
    launch sequence, and the boundary types (they select kernel variants)."""
    sh = ";".join("x".join(str(v) for v in s) for s in sorted(shapes))
    types = "".join(str(int(f.type)) for f in bc.faces())
    return (f"{kind}|{device}|{sh}|levels={cfg.mg_levels}|pre={cfg.mg_pre_smooth}|"
            f"post={cfg.mg_post_smooth}|prec={int(cfg.precond_precision)}|"
            f"coef={int(cfg.coef_format)}|bc={types}")


def _key(kind: str, ctx, shapes, cfg, bc) -> str:
    return key_of(kind, _device_name(ctx), shapes, cfg, bc)


def cached_or_tune(key: str, tune, use_cache: bool = True) -> dict:
    """The cached entry for key, or tune() stored under it (process and file
    locks held while tuning, so concurrent tuners do not repeat the work)."""
    path = cache_path()
    with _LOCK:
        if use_cache:
            hit = _load(path).get(key)
            if hit:
                return hit
        out = tune()
        _store(path, key, out)
        return out


def _median_ms(ctx, fn, reps: int, events: bool) -> float:
    """fn() once untimed, then the median of reps timed runs (HIP events on
    the context stream, or the host clock around a synchronised call)."""
    fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(max(1, reps)):
        if events:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(ctx.stream)
            fn()
            e1.record(ctx.stream)
            e1.synchronize()
            times.append(e0.elapsed_time(e1))
        else:
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(times)


def tune_sweep_block(ctx, grid, bc, cfg, K, reps: int = 5, use_cache: bool = True) -> dict:
    """Fastest sweep workgroup size for this grid (finest-level smoothing, in place)."""
    def tune():
        times = {}
        for b in BLOCKS:
            ws = ph.FlowWorkspace()
            ws.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, b)
            mg.mg_setup(ctx, ws, grid, bc, cfg, K)
            times[b] = _median_ms(ctx, lambda: mg.smooth_level(ctx, ws, 0, 1), reps, True)
            ws.close()
        best = min(times, key=times.get)
        return {"sweep_block": best, "ms_per_iteration": {str(b): t for b, t in times.items()}}
    return cached_or_tune(_key("sweep", ctx, [(grid.nx, grid.ny, grid.nz)], cfg, bc), tune,
                          use_cache)


OP_CHUNKS = (4, 8, 16, 32)
MARCH_BLOCKS = (1024, 2048, 4096)


def tune_solve(ctx, grid, bc, cfg, K, reps: int = 3, use_cache: bool = True,
               blocks=(128, 256, 512), op_chunks=OP_CHUNKS, march_blocks=MARCH_BLOCKS) -> dict:
    """Fastest (sweep_block, graph, op_tile), then op_chunk and (FP64
    V-cycle) march_blocks, for single solves of this grid: whole solves into
    a scratch head, one untimed then the median of reps."""
    h = ctx.zeros(grid.num_cells())

    def time_params(params):
        ws = ph.FlowWorkspace()
        apply([ws], params)
        t = _median_ms(ctx, lambda: ph.solve_head(h, K, grid, bc, cfg, ctx, ws), reps, False)
        ws.close()
        return t

    def tune():
        times = {}
        for b in blocks:
            for graph in (True, False):
                for tile in (True, False):
                    times[(b, graph, tile)] = time_params(
                        {"sweep_block": b, "graph": graph, "op_tile": tile})
        b, graph, tile = min(times, key=times.get)
        best = {"sweep_block": b, "graph": graph, "op_tile": tile}
        ms = {f"{k[0]}/{'graph' if k[1] else 'launches'}/{'tile' if k[2] else 'pairs'}": v
              for k, v in times.items()}
        stages = [("op_chunk", op_chunks)]
        if cfg.precond_precision == ph.PrecondPrecision.FP64:
            stages.append(("march_blocks", march_blocks))
        for knob, values in stages:
            cand = {v: time_params({**best, knob: v}) for v in values}
            best[knob] = min(cand, key=cand.get)
            ms.update({f"{knob}={v}": t for v, t in cand.items()})
        return {**best, "ms": ms}
    return cached_or_tune(_key("solve", ctx, [(grid.nx, grid.ny, grid.nz)], cfg, bc), tune,
                          use_cache)


def tune_batch(ctx, hs, Ks, grids, bc, cfg, workspaces, reps: int = 3,
               use_cache: bool = True) -> dict:
    """Fastest (sweep_block, graph) for a batch: whole batched solves per candidate."""
    def tune():
        times = {}
        for b in (128, 256, 512):
            for graph in (True, False):
                apply(workspaces, {"sweep_block": b})
                times[(b, graph)] = _median_ms(
                    ctx, lambda: ph.solve_head_batched(hs, Ks, grids, bc, cfg, ctx, workspaces,
                                                       use_graph=graph), reps, False)
        (b, graph) = min(times, key=times.get)
        return {"sweep_block": b, "graph": graph,
                "ms": {f"{k[0]}/{'graph' if k[1] else 'launches'}": v for k, v in times.items()}}
    return cached_or_tune(_key("batch", ctx, [(g.nx, g.ny, g.nz) for g in grids], cfg, bc),
                          tune, use_cache)


def apply(workspaces, params: dict) -> None:
    """Sets tuned choices on the workspaces (they take effect at the next set-up
    or solve); keys absent from params are left as they are."""
    for w in workspaces:
        if "sweep_block" in params:
            w.set_tuning(ph.FlowWorkspace.TUNE_SWEEP_BLOCK, int(params["sweep_block"]))
        if "graph" in params:
            w.set_tuning(ph.FlowWorkspace.TUNE_GRAPH, int(bool(params["graph"])))
        if "op_tile" in params:
            w.set_tuning(ph.FlowWorkspace.TUNE_OP_TILE, int(bool(params["op_tile"])))
        if "op_chunk" in params:
            w.set_tuning(ph.FlowWorkspace.TUNE_OP_CHUNK, int(params["op_chunk"]))
        if "march_blocks" in params:
            w.set_tuning(ph.FlowWorkspace.TUNE_MARCH_BLOCKS, int(params["march_blocks"]))
