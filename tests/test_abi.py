"""C ABI checks that need no GPU: the library loads, exports exactly what
include/mf3d.h declares, and rejects bad arguments before touching HIP."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from macroflow3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mf3d.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|void|size_t|const char\*)\s+(mf3d_\w+)\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 28
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (mf3d_\w+)", out))
    assert set(names) == exported


def test_binding_table_covers_header():
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_library_metadata_and_defaults():
    L = _lib.lib()
    assert L.mf3d_abi_version() == 1
    assert L.mf3d_status_string(0) == b"ok"
    assert L.mf3d_status_string(2) == b"size mismatch"
    cfg = _lib.mf3d_head_solve_config()
    L.mf3d_default_config(C.byref(cfg))
    # HeadSolveConfig defaults (src/physics/flow/solve_head.cuh:64-96)
    assert (cfg.solver_type, cfg.mg_levels, cfg.mg_pre_smooth, cfg.mg_post_smooth,
            cfg.mg_max_cycles, cfg.mg_coarse_iters, cfg.cg_max_iter, cfg.cg_check_every) == \
        (2, 4, 2, 2, 20, 50, 1000, 10)
    assert cfg.cg_rtol == 1e-8 and cfg.rtol == 1e-6 and cfg.pin_mode == 0
    assert (cfg.init_guess, cfg.coarse_solver, cfg.stop_norm, cfg.precond_precision) == (0, 0, 0, 0)


def test_struct_layouts_match_header():
    assert C.sizeof(_lib.mf3d_grid) == 40
    assert C.sizeof(_lib.mf3d_bc) == 56
    assert _lib.mf3d_bc.value.offset == 8


def test_argument_validation_before_any_device_work():
    L = _lib.lib()
    ws = C.c_void_p()
    assert L.mf3d_workspace_create(C.byref(ws)) == 0
    g = _lib.mf3d_grid(8, 8, 8, 1.0, 1.0, 1.0)
    bc = _lib.mf3d_bc()
    cfg = _lib.mf3d_head_solve_config()
    L.mf3d_default_config(C.byref(cfg))
    res = _lib.mf3d_head_solve_result()
    fake = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    # null grid
    assert L.mf3d_solve_head(ws, None, None, C.byref(bc), C.byref(cfg), fake, fake, 512,
                             C.byref(res)) == 1
    # periodic on one side only (BCSpec::validate, BCSpec.hpp:66-80)
    bad = _lib.mf3d_bc()
    bad.type[2] = 2
    assert L.mf3d_solve_head(ws, None, C.byref(g), C.byref(bad), C.byref(cfg), fake, fake, 512,
                             C.byref(res)) == 1
    assert b"Periodic BC must be specified on both ymin and ymax" in L.mf3d_last_error()
    # short buffer (solve_head.cu:85-87)
    assert L.mf3d_solve_head(ws, None, C.byref(g), C.byref(bc), C.byref(cfg), fake, fake, 100,
                             C.byref(res)) == 2
    assert b"too small" in L.mf3d_last_error()
    # non-finite BC value
    nan_bc = _lib.mf3d_bc()
    nan_bc.value[0] = float("nan")
    assert L.mf3d_init_head_guess(None, C.byref(g), C.byref(nan_bc), fake, 512) == 1
    # non-2:1 transfer sizes
    c = _lib.mf3d_grid(3, 4, 4, 2.0, 2.0, 2.0)
    assert L.mf3d_restrict_3d(None, C.byref(g), C.byref(c), fake, 512, fake, 48) == 2
    # bad config values
    cfg.mg_levels = 0
    assert L.mf3d_solve_head(ws, None, C.byref(g), C.byref(bc), C.byref(cfg), fake, fake, 512,
                             C.byref(res)) == 1
    assert L.mf3d_workspace_bytes(ws) == 0
    assert L.mf3d_workspace_destroy(ws) == 0


def test_python_layer_validates_like_reference():
    from macroflow3d_amd.core import BCFace, BCSpec, BCType, PinMode, needs_pin, pin_enabled
    from macroflow3d_amd.physics import HeadSolveConfig, HeadSolverType
    bc = BCSpec.darcy_x()
    bc.validate()
    assert not needs_pin(bc) and not pin_enabled(PinMode.Auto, bc)
    allN = BCSpec(*[BCFace(BCType.Neumann, 0.0) for _ in range(6)])
    assert needs_pin(allN) and allN.is_all_homog_neumann()
    bad = BCSpec.darcy_x()
    bad.zmax = BCFace(BCType.Periodic, 0.0)
    with pytest.raises(RuntimeError, match="zmin and zmax"):
        bad.validate()
    cfg = HeadSolveConfig.from_yaml({"solver": "mg_cg", "mg_levels": 6, "cg_rtol": 1e-6,
                                     "pin": {"mode": "on"}})
    assert cfg.solver_type == HeadSolverType.PCG_MG and cfg.mg_levels == 6
    assert cfg.pin == PinMode.On and cfg.cg_rtol == 1e-6
    assert HeadSolveConfig.from_yaml({"solver": "cg"}).solver_type == HeadSolverType.CG
    assert HeadSolveConfig.from_yaml({"solver": "mg"}).solver_type == HeadSolverType.MG


def test_tuning_keys_validate_their_ranges():
    """Workspace tuning keys accept their documented ranges and reject the
    rest without device work; a fresh workspace counts no applications."""
    L = _lib.lib()
    ws = C.c_void_p()
    assert L.mf3d_workspace_create(C.byref(ws)) == 0
    cases = {14: ([-1, 0, 1], [-2, 2]),           # graph replay
             15: ([1, 16, 64], [0, 65]),          # operator chunk depth
             16: ([256, 2048, 8192], [255, 8193]),  # FP64 sweep workgroup target
             5: ([64, 256, 512], [32, 100, 576])}  # sweep workgroup size
    for key, (good, bad) in cases.items():
        for v in good:
            assert L.mf3d_workspace_set_tuning(ws, key, v) == 0, (key, v)
        for v in bad:
            assert L.mf3d_workspace_set_tuning(ws, key, v) != 0, (key, v)
    assert L.mf3d_workspace_set_tuning(ws, 99, 1) != 0
    op, pc = C.c_long(-1), C.c_long(-1)
    assert L.mf3d_workspace_solve_counts(ws, C.byref(op), C.byref(pc)) == 0
    assert (op.value, pc.value) == (0, 0)
    assert L.mf3d_workspace_destroy(ws) == 0
