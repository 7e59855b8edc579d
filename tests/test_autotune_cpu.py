"""The autotuner's cache without a GPU: keys separate what changes the launch
sequence (shapes, levels, smoothing counts, precision, coefficient format,
boundary types), entries survive a round trip through the JSON file, a cached
entry is not tuned again, and concurrent writers (threads and processes)
leave one valid file holding every entry."""
from __future__ import annotations

import json
import multiprocessing as mp
import threading

import pytest

from macroflow3d_amd import autotune
from macroflow3d_amd import physics as ph
from macroflow3d_amd.core import BCFace, BCSpec, BCType


def _periodic_y():
    d = BCSpec.darcy_x()
    return BCSpec(d.xmin, d.xmax, BCFace(BCType.Periodic, 0.0), BCFace(BCType.Periodic, 0.0),
                  d.zmin, d.zmax)


def test_key_separates_launch_relevant_fields():
    cfg = ph.HeadSolveConfig.mixed(1e-10, 6)
    base = autotune.key_of("solve", "gpu", [(256, 256, 256)], cfg, BCSpec.darcy_x())
    assert base == autotune.key_of("solve", "gpu", [(256, 256, 256)], cfg, BCSpec.darcy_x())
    others = [autotune.key_of("solve", "gpu", [(256, 256, 255)], cfg, BCSpec.darcy_x()),
              autotune.key_of("solve", "gpu", [(256, 256, 256)], cfg, _periodic_y()),
              autotune.key_of("batch", "gpu", [(256, 256, 256)], cfg, BCSpec.darcy_x()),
              autotune.key_of("solve", "gpu2", [(256, 256, 256)], cfg, BCSpec.darcy_x()),
              autotune.key_of("solve", "gpu", [(256, 256, 256)], ph.HeadSolveConfig.mixed(1e-10, 5),
                              BCSpec.darcy_x()),
              autotune.key_of("solve", "gpu", [(256, 256, 256)],
                              ph.HeadSolveConfig.fp64_direct(1e-10, 6), BCSpec.darcy_x())]
    pre = ph.HeadSolveConfig.mixed(1e-10, 6)
    pre.mg_pre_smooth = 3
    others.append(autotune.key_of("solve", "gpu", [(256, 256, 256)], pre, BCSpec.darcy_x()))
    assert base not in others and len(set(others)) == len(others)
    # the shapes of a batch are a set: order does not matter
    a = autotune.key_of("batch", "gpu", [(48,) * 3, (96,) * 3], cfg, BCSpec.darcy_x())
    b = autotune.key_of("batch", "gpu", [(96,) * 3, (48,) * 3], cfg, BCSpec.darcy_x())
    assert a == b


def test_cache_round_trip_and_hit(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("MF3D_AUTOTUNE_CACHE", str(path))
    calls = []

    def tune():
        calls.append(1)
        return {"sweep_block": 128, "graph": True}
    assert autotune.cached_or_tune("k1", tune) == {"sweep_block": 128, "graph": True}
    assert autotune.cached_or_tune("k1", tune) == {"sweep_block": 128, "graph": True}
    assert len(calls) == 1  # the second call was a cache hit
    assert json.loads(path.read_text())["k1"]["sweep_block"] == 128
    autotune.cached_or_tune("k1", tune, use_cache=False)
    assert len(calls) == 2
    path.write_text("not json")  # a damaged file reads as empty and is rewritten
    autotune.cached_or_tune("k2", tune)
    assert set(json.loads(path.read_text())) == {"k2"}


def _writer(path, i):
    for j in range(20):
        autotune._store(path, f"p{i}-{j}", {"v": j})


def test_concurrent_writers_keep_every_entry(tmp_path):
    path = str(tmp_path / "tune.json")
    threads = [threading.Thread(target=_writer, args=(path, i)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    ctx = mp.get_context("spawn")
    procs = [ctx.Process(target=_writer, args=(path, 10 + i)) for i in range(3)]
    for p in procs:
        p.start()
    for p in procs:
        p.join(timeout=120)
        assert p.exitcode == 0
    data = autotune._load(path)
    assert len(data) == 7 * 20
    assert all(data[f"p{i}-{j}"] == {"v": j} for i in (0, 1, 2, 3, 10, 11, 12) for j in range(20))
