"""Tensor handoff between callers and the native library.

PyTorch-ROCm only carries device memory and the stream here; every step of
the solver runs in the HIP kernels of ``libcardiax_amd.so``.
"""
from __future__ import annotations

import numpy as np
import torch


def gpu_available() -> bool:
    return torch.cuda.is_available()


def default_device() -> torch.device:
    """Where host-built fields are placed: the current GPU when one is
    present, else host memory (the solver itself always needs a GPU)."""
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def as_field(a) -> torch.Tensor:
    """A float32 tensor on the default device from a NumPy array or tensor."""
    if isinstance(a, torch.Tensor):
        return a.to(device=default_device(), dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(default_device())


def require_gpu(what: str) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "%s needs a ROCm GPU: cardiax_amd runs only through its HIP kernels and has no "
            "CPU fallback" % what)
    return torch.device("cuda", torch.cuda.current_device())


def on_device(x, device: torch.device, name: str = "array") -> torch.Tensor:
    """float32, contiguous, on ``device``; copies only when needed."""
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x), dtype=np.float32))
    if t.dtype != torch.float32:
        t = t.to(torch.float32)
    if t.device != device:
        t = t.to(device)
    if not t.is_contiguous():
        t = t.contiguous()
    return t


def stream_handle() -> int:
    """hipStream_t of torch's current stream on the current device."""
    return torch.cuda.current_stream().cuda_stream
