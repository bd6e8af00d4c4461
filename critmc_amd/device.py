"""Device plumbing: PyTorch-ROCm owns the buffers, the HIP library computes.

Only allocation, copies and the current stream come from torch; every update
and reduction runs in the library's hand-written kernels.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch


class NoGPU(RuntimeError):
    pass


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise NoGPU(
            "critmc_amd runs on an AMD GPU through its HIP kernels; torch.cuda.is_available() is "
            "False here"
        )
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> ctypes.c_void_p:
    """The torch current stream as a ``hipStream_t``."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def host_ptr(a: np.ndarray, typ):
    return a.ctypes.data_as(ctypes.POINTER(typ))


def from_numpy_u32(a: np.ndarray, dev: torch.device) -> torch.Tensor:
    """uint32 data in an int32 tensor (bit pattern preserved)."""
    a = np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)
    return torch.from_numpy(a.copy()).to(dev)


def from_numpy_u64(a: np.ndarray, dev: torch.device) -> torch.Tensor:
    a = np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)
    return torch.from_numpy(a.copy()).to(dev)


def to_numpy_u64(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64)
