"""Device context: device, HIP stream and argument checks.

Counterpart of ``CudaContext`` (src/runtime/CudaContext.cuh:8-35,
CudaContext.cu:6-42): one context per device and host thread; every
operation of this package is enqueued on ``ctx.stream`` and
``ctx.synchronize()`` waits for that stream only. Device arrays are torch
tensors (float64, contiguous, on the context's device); torch provides the
memory and the stream object, the kernels are this package's own.
"""
from __future__ import annotations

import torch

from . import _lib


class HipContext:
    def __init__(self, device: int = 0, stream: "torch.cuda.Stream | None" = None):
        if not torch.cuda.is_available():
            raise RuntimeError("HipContext needs a GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", device)
        # By default the context works on torch's current stream of the device,
        # so tensors made by torch need no cross-stream wait; pass a stream
        # (default priority) to use a dedicated one.
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        _lib.lib()  # fail loudly now if the library is missing

    @property
    def handle(self) -> int:
        return int(self.stream.cuda_stream)

    def synchronize(self) -> None:
        self.stream.synchronize()

    def empty(self, n: int) -> torch.Tensor:
        return torch.empty(int(n), dtype=torch.float64, device=self.device)

    def zeros(self, n: int) -> torch.Tensor:
        return torch.zeros(int(n), dtype=torch.float64, device=self.device)

    def __enter__(self):
        self._ctx = torch.cuda.stream(self.stream)
        self._ctx.__enter__()
        return self

    def __exit__(self, *exc):
        self._ctx.__exit__(*exc)
        return False


# The reference spells its context ``CudaContext``; same object here.
CudaContext = HipContext


def dptr(t: torch.Tensor, need: int, name: str, ctx: HipContext | None = None,
         dtype: torch.dtype = torch.float64) -> int:
    """Checks a device array argument (dtype, contiguity, device, length) and
    returns its address."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a device tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name}: must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if ctx is not None and t.device != ctx.device:
        raise ValueError(f"{name}: on {t.device}, context is on {ctx.device}")
    if t.numel() < need:
        raise RuntimeError(f"{name}: buffer too small ({t.numel()} < {need})")
    return int(t.data_ptr())


def record(ctx: HipContext, *tensors: torch.Tensor) -> None:
    """Marks tensors as used on the context stream (torch caching allocator)."""
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            t.record_stream(ctx.stream)
