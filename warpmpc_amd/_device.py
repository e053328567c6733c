"""Device-memory plumbing: PyTorch-ROCm tensors as the buffers handed to the C ABI."""

from __future__ import annotations

import numpy as np


def torch():
    import torch as _torch

    return _torch


def require_gpu() -> None:
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("the HIP backend needs a ROCm GPU (torch.cuda.is_available() is False)")


def stream_ptr() -> int:
    return int(torch().cuda.current_stream().cuda_stream)


def as_device_f32(x, name: str):
    """Contiguous float32 tensor on the current GPU (copies numpy input / other dtypes)."""
    t = torch()
    if isinstance(x, np.ndarray) or not isinstance(x, t.Tensor):
        x = t.as_tensor(np.asarray(x, dtype=np.float32))
    if x.dtype != t.float32:
        x = x.to(t.float32)
    if not x.is_cuda:
        x = x.cuda()
    if not x.is_contiguous():
        x = x.contiguous()
    return x


def ptr(x) -> int | None:
    return None if x is None else int(x.data_ptr())


__all__ = ["as_device_f32", "ptr", "require_gpu", "stream_ptr", "torch"]
