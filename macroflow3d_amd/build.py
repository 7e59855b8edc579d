"""Builds libmf3d_hip.so (gfx950) in-tree with hipcc.

Usage: python -m macroflow3d_amd.build [--jobs N]
Objects go to macroflow3d_amd/_build/, the shared library next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmf3d_hip.so")
ROOT = os.path.dirname(HERE)

SOURCES = ["stencil.hip", "faces.hip", "split64.hip", "blas.hip", "velocity.hip", "solver.hip",
           "cnn.hip"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # never emit packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


# C++ drivers over include/macroflow3d (host code only, linked to the library)
DRIVERS = ["dropin_solve", "ensemble_dropin"]
DROPIN_BIN = os.path.join(BUILD, "dropin_solve")
ENSEMBLE_BIN = os.path.join(BUILD, "ensemble_dropin")


def build_dropin() -> list[str]:
    """The C++ drop-in drivers (tests/cpp/*.cpp) over include/macroflow3d."""
    out = []
    for name in DRIVERS:
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        if not os.path.exists(src):
            continue
        dst = os.path.join(BUILD, name)
        cmd = [hipcc(), "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
               src, "-L" + HERE, "-lmf3d_hip", "-Wl,-rpath,$ORIGIN/..", "-o", dst]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {name}.cpp:\n{r.stdout}\n{r.stderr}")
        out.append(dst)
    return out


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_dropin()
    if verbose:
        print(f"built {LIB}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
