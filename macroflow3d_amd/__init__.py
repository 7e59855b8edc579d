"""MI355X-native Darcy head solver (MacroFlow3D head-solve path, gfx950).

Host-side mirror of MacroFlow3D's solver interfaces over the C ABI of
libmf3d_hip.so (include/mf3d.h). Layers:
  core       Grid3D, BCSpec, PinSpec ...             (src/core, pin_spec.hpp)
  runtime    HipContext (= CudaContext)               (src/runtime/CudaContext.*)
  multigrid  gsrb_smooth_3d, residual, transfers      (src/multigrid)
  numerics   VarCoeffLaplacian, blas, pcg_solve       (src/numerics)
  physics    solve_head, velocities, FlowWorkspace    (src/physics)
  kfield     log-normal K by FFT circulant embedding  (synthetic input)
"""
from .core import BCFace, BCSpec, BCType, Grid3D, PinMode, PinSpec, needs_pin, pin_enabled  # noqa: F401

__all__ = ["BCFace", "BCSpec", "BCType", "Grid3D", "PinMode", "PinSpec", "needs_pin",
           "pin_enabled"]
