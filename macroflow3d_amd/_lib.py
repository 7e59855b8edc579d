"""ctypes binding of libmf3d_hip.so (the C ABI declared in include/mf3d.h).

The library is required: importing a module that computes raises at the first
call when the shared object is missing or cannot be loaded. There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmf3d_hip.so")

MF3D_OK = 0
STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARGUMENT", 2: "SIZE_MISMATCH", 3: "BREAKDOWN", 4: "HIP",
    5: "ALLOC", 6: "SINGULAR", 7: "UNSUPPORTED",
}


class MF3DError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"mf3d error {status} ({STATUS_NAMES.get(status, '?')}): {message}")
        self.status = status


class mf3d_grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class mf3d_bc(C.Structure):
    _fields_ = [("type", C.c_uint8 * 6), ("value", C.c_double * 6)]


class mf3d_head_solve_config(C.Structure):
    _fields_ = [("solver_type", C.c_int), ("mg_levels", C.c_int),
                ("mg_pre_smooth", C.c_int), ("mg_post_smooth", C.c_int),
                ("mg_max_cycles", C.c_int), ("mg_coarse_iters", C.c_int),
                ("cg_max_iter", C.c_int), ("cg_check_every", C.c_int),
                ("cg_rtol", C.c_double), ("rtol", C.c_double), ("pin_mode", C.c_int),
                ("init_guess", C.c_int), ("coarse_solver", C.c_int), ("stop_norm", C.c_int),
                ("precond_precision", C.c_int), ("coef_format", C.c_int),
                ("true_residual", C.c_int), ("flexible", C.c_int), ("reserved", C.c_int * 4)]


class mf3d_head_solve_result(C.Structure):
    _fields_ = [("num_iterations", C.c_int), ("converged", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("bnorm", C.c_double), ("true_residual", C.c_double),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double),
                ("restarts", C.c_int), ("reserved", C.c_int * 3)]


_P = C.POINTER
_vp = C.c_void_p
_d = C.c_double
_sz = C.c_size_t
_i = C.c_int
_g = _P(mf3d_grid)
_b = _P(mf3d_bc)
_cfg = _P(mf3d_head_solve_config)

# name -> (restype, argtypes); the full C ABI of include/mf3d.h
SIGNATURES = {
    "mf3d_abi_version": (_i, []),
    "mf3d_status_string": (C.c_char_p, [_i]),
    "mf3d_last_error": (C.c_char_p, []),
    "mf3d_default_config": (None, [_cfg]),
    "mf3d_workspace_create": (_i, [_P(_vp)]),
    "mf3d_workspace_destroy": (_i, [_vp]),
    "mf3d_workspace_bytes": (_sz, [_vp]),
    "mf3d_solve_head": (_i, [_vp, _vp, _g, _b, _cfg, _vp, _vp, _sz, _P(mf3d_head_solve_result)]),
    "mf3d_init_head_guess": (_i, [_vp, _g, _b, _vp, _sz]),
    "mf3d_build_rhs_head": (_i, [_vp, _g, _b, _vp, _vp, _sz]),
    "mf3d_coarsen_K": (_i, [_vp, _g, _g, _vp, _sz, _vp, _sz]),
    "mf3d_velocity_compact": (_i, [_vp, _g, _b, _vp, _vp, _sz, _vp, _sz, _vp, _sz, _vp, _sz]),
    "mf3d_velocity_padded": (_i, [_vp, _g, _b, _vp, _vp, _sz, _vp, _vp, _vp, _sz]),
    "mf3d_velocity_diagnostics": (_i, [_vp, _vp, _g, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _sz,
                                       _vp]),
    "mf3d_field_stats": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "mf3d_gsrb_smooth_3d": (_i, [_vp, _g, _vp, _vp, _vp, _sz, _i, _b, _i]),
    "mf3d_compute_residual_3d": (_i, [_vp, _g, _vp, _vp, _vp, _vp, _sz, _b, _i]),
    "mf3d_restrict_3d": (_i, [_vp, _g, _g, _vp, _sz, _vp, _sz]),
    "mf3d_prolong_3d_add": (_i, [_vp, _g, _g, _vp, _sz, _vp, _sz]),
    "mf3d_mg_setup": (_i, [_vp, _vp, _g, _b, _cfg, _vp, _sz]),
    "mf3d_mg_num_levels": (_i, [_vp]),
    "mf3d_workspace_set_tuning": (_i, [_vp, _i, C.c_long]),
    "mf3d_workspace_solve_counts": (_i, [_vp, _vp, _vp]),
    "mf3d_workspace_reserve": (_i, [_vp, _g, _i]),
    "mf3d_mg_smooth": (_i, [_vp, _vp, _i, _i]),
    "mf3d_mg_precondition": (_i, [_vp, _vp, _vp, _vp, _sz]),
    "mf3d_varcoeff_apply": (_i, [_vp, _g, _b, _i, _vp, _vp, _vp, _sz]),
    "mf3d_dot": (_i, [_vp, _vp, _vp, _vp, _sz, _vp]),
    "mf3d_nrm2": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "mf3d_dot_noalloc": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "mf3d_axpy": (_i, [_vp, _d, _vp, _vp, _sz]),
    "mf3d_axpby": (_i, [_vp, _d, _vp, _d, _vp, _sz]),
    "mf3d_copy": (_i, [_vp, _vp, _vp, _sz]),
    "mf3d_fill": (_i, [_vp, _vp, _sz, _d]),
    "mf3d_scal": (_i, [_vp, _vp, _sz, _d]),
    "mf3d_solve_head_batched": (_i, [_i, _P(_vp), _vp, _P(mf3d_grid), _P(mf3d_bc), _cfg,
                                     _P(_vp), _P(_vp), _P(_sz), _i,
                                     _P(mf3d_head_solve_result)]),
    # learned warm start
    "mf3d_wsnet_param_count": (C.c_long, []),
    "mf3d_wsnet_create": (_i, [_P(_vp)]),
    "mf3d_wsnet_destroy": (_i, [_vp]),
    "mf3d_wsnet_set_params": (_i, [_vp, _vp, C.c_long]),
    "mf3d_wsnet_get_params": (_i, [_vp, _vp, C.c_long]),
    "mf3d_wsnet_get_grads": (_i, [_vp, _vp, C.c_long]),
    "mf3d_wsnet_predict": (_i, [_vp, _vp, _P(mf3d_grid), _P(mf3d_bc), _vp, _i, _vp, _sz]),
    "mf3d_wsnet_gradients": (_i, [_vp, _vp, _P(mf3d_grid), _P(mf3d_bc), _vp, _i, _vp, _sz,
                                  _P(_d)]),
    "mf3d_wsnet_train_step": (_i, [_vp, _vp, _P(mf3d_grid), _P(mf3d_bc), _vp, _i, _vp, _sz, _d,
                                   _P(_d)]),
    "mf3d_conv3d_forward": (_i, [_vp, _P(mf3d_grid), _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "mf3d_conv3d_backward_data": (_i, [_vp, _P(mf3d_grid), _i, _i, _vp, _vp, _vp, _vp]),
    "mf3d_conv3d_backward_weights": (_i, [_vp, _P(mf3d_grid), _i, _i, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """Loads libmf3d_hip.so once; raises if it is absent (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m macroflow3d_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != MF3D_OK:
        msg = lib().mf3d_last_error()
        raise MF3DError(status, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
