"""Device plumbing: current stream, raw pointers, reusable scratch buffers.

torch-ROCm is used only to own device memory and the stream; every kernel on
the hot path is a call into libparsec_hip.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise _lib.ParsecHipError("parsec_python_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.lib()


def stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    return C.c_void_p(t.data_ptr())


F64 = torch.float64
F32 = torch.float32


def empty(shape, dtype=F64) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device="cuda")


def zeros(shape, dtype=F64) -> torch.Tensor:
    return torch.zeros(shape, dtype=dtype, device="cuda")


class Scratch:
    """Grow-only named device buffers (allocated outside the hot loop once)."""

    def __init__(self):
        self._bufs: dict[str, torch.Tensor] = {}

    def get(self, name: str, numel: int, dtype=F64) -> torch.Tensor:
        b = self._bufs.get(name)
        if b is None or b.numel() < numel or b.dtype != dtype:
            b = torch.empty(max(numel, 1), dtype=dtype, device="cuda")
            self._bufs[name] = b
        return b[:numel]

    def bytes(self, name: str, nbytes: int) -> torch.Tensor:
        return self.get(name, (nbytes + 7) // 8, torch.float64)
