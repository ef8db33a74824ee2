"""Device memory and stream plumbing (torch-ROCm tensors carry the buffers).

The kernels never see torch types: only raw device pointers and the HIP stream
handle of torch's current stream cross the C ABI.
"""
from __future__ import annotations

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t

        _torch = _t
    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("mdtraj_amd computes on the GPU only and no GPU is visible")
    return t


def is_device_array(x) -> bool:
    if x is None or isinstance(x, np.ndarray):
        return False
    t = torch()
    return isinstance(x, t.Tensor) and x.is_cuda


def device():
    return torch().device("cuda", torch().cuda.current_device())


def to_device_f32(x):
    """float32 C-contiguous device tensor for x (numpy or tensor); no copy if
    x already is one."""
    t = require_gpu()
    if isinstance(x, t.Tensor):
        y = x
        if not y.is_cuda:
            y = y.to(device())
        if y.dtype != t.float32:
            y = y.to(t.float32)
        return y.contiguous()
    a = np.ascontiguousarray(x, dtype=np.float32)
    return t.from_numpy(a).to(device())


def to_device_i32(idx):
    if idx is None:
        return None
    t = require_gpu()
    if isinstance(idx, t.Tensor):
        return idx.to(device=device(), dtype=t.int32).contiguous()
    a = np.ascontiguousarray(idx, dtype=np.int32)
    return t.from_numpy(a).to(device())


def empty_f32(n):
    t = require_gpu()
    return t.empty(int(n), dtype=t.float32, device=device())


def workspace(nbytes: int):
    t = require_gpu()
    return t.empty(max(int(nbytes), 256), dtype=t.uint8, device=device())


def ptr(x):
    return None if x is None else x.data_ptr()


def stream():
    """HIP stream handle of torch's current stream (kernels launch on it)."""
    return torch().cuda.current_stream().cuda_stream


def to_host(x):
    return x.detach().cpu().numpy()
