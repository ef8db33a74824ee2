"""Device memory and streams for the Python layer.

Two kinds of device buffer cross the C ABI, always as raw pointers:

* ``DeviceArray``, this module's own buffer from ``hr_alloc``.  Host (numpy)
  inputs are uploaded into these and every kernel then runs on the default
  stream, so a process that never imports torch runs every call;
* torch CUDA tensors, when the caller passes them: results come back as
  tensors and the kernels run on torch's current stream.

Each public function selects the kind from its main input with
``with using(x):``; allocations inside follow it.  No torch type crosses the
ABI, and torch is imported only for torch inputs (or by modules that need its
streams and events: the serving loop, batch plans and the autotuner).
"""
from __future__ import annotations

import ctypes
import functools
import sys
import threading

import numpy as np

from ._lib import call, load

_tls = threading.local()


# ------------------------------------------------------------------- torch
def torch():
    """The torch module (imported on first use)."""
    import torch as _t

    return _t


def _torch_loaded():
    return sys.modules.get("torch")


def is_torch_cuda(x) -> bool:
    t = _torch_loaded()
    return t is not None and isinstance(x, t.Tensor) and x.is_cuda


def is_torch_tensor(x) -> bool:
    t = _torch_loaded()
    return t is not None and isinstance(x, t.Tensor)


# -------------------------------------------------------- library buffers
class _BlockCache:
    """Freed hr_alloc blocks kept for reuse, by rounded size.  Every
    DeviceArray kernel runs on the default stream, so a block handed out again
    is only touched by work queued after the work that used it before."""

    LIMIT = 8 << 30

    def __init__(self):
        self.lock = threading.Lock()
        self.free = {}
        self.held = 0

    @staticmethod
    def rounded(nbytes):
        nbytes = max(int(nbytes), 256)
        if nbytes <= (1 << 20):
            return 1 << (nbytes - 1).bit_length()
        return (nbytes + (2 << 20) - 1) // (2 << 20) * (2 << 20)

    def get(self, nbytes):
        size = self.rounded(nbytes)
        with self.lock:
            lst = self.free.get(size)
            if lst:
                self.held -= size
                return lst.pop(), size
        p = ctypes.c_void_p()
        try:
            call("hr_alloc", ctypes.byref(p), size)
        except RuntimeError:
            self.release()
            call("hr_alloc", ctypes.byref(p), size)
        return p.value, size

    def put(self, ptr, size):
        with self.lock:
            if self.held + size <= self.LIMIT:
                self.free.setdefault(size, []).append(ptr)
                self.held += size
                return
        load().hr_free(ctypes.c_void_p(ptr))

    def release(self):
        """Return every cached block to the device."""
        with self.lock:
            blocks = [p for lst in self.free.values() for p in lst]
            self.free.clear()
            self.held = 0
        lib = load()
        for p in blocks:
            lib.hr_free(ctypes.c_void_p(p))


_cache = _BlockCache()


class _Block:
    __slots__ = ("ptr", "size", "__weakref__")

    def __init__(self, nbytes):
        self.ptr, self.size = _cache.get(nbytes)

    def __del__(self):
        try:
            if self.ptr:
                _cache.put(self.ptr, self.size)
        except Exception:   # interpreter shutdown
            pass


class DeviceArray:
    """A C-ordered device array in library-allocated memory (float32, float64,
    int32, int64 or uint8).  The subset of the tensor interface the package uses:
    ``data_ptr``, ``shape``, ``dim``, ``numel``, ``reshape``, integer and
    unit-step slice indexing on the first axis (views), ``clone``, ``copy_``,
    ``zero_``, ``numpy``."""

    __slots__ = ("_block", "_offset", "shape", "dtype")
    is_cuda = True

    def __init__(self, block, offset, shape, dtype):
        self._block = block
        self._offset = int(offset)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)

    @classmethod
    def empty(cls, shape, dtype=np.float32):
        shape = (int(shape),) if np.isscalar(shape) else tuple(int(s) for s in shape)
        n = int(np.prod(shape, dtype=np.int64)) * np.dtype(dtype).itemsize
        return cls(_Block(n), 0, shape, dtype)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls.empty(a.shape, a.dtype)
        if a.nbytes:
            call("hr_memcpy_h2d", ctypes.c_void_p(d.data_ptr()), a.ctypes.data_as(ctypes.c_void_p),
                 a.nbytes, None)
        return d

    # -- tensor-like surface
    def data_ptr(self):
        return self._block.ptr + self._offset

    def dim(self):
        return len(self.shape)

    @property
    def ndim(self):
        return len(self.shape)

    def numel(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def __len__(self):
        return self.shape[0]

    def element_size(self):
        return self.dtype.itemsize

    @property
    def nbytes(self):
        return self.numel() * self.dtype.itemsize

    def is_contiguous(self):
        return True

    def contiguous(self):
        return self

    def stride(self, axis):
        return int(np.prod(self.shape[axis + 1:], dtype=np.int64))

    def _row_bytes(self):
        return int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize

    def __getitem__(self, key):
        if isinstance(key, (int, np.integer)):
            k = int(key)
            if k < 0:
                k += self.shape[0]
            if not 0 <= k < self.shape[0]:
                raise IndexError(f"index {key} out of range for {self.shape[0]} rows")
            return DeviceArray(self._block, self._offset + k * self._row_bytes(), self.shape[1:],
                               self.dtype)
        if isinstance(key, slice):
            a, b, step = key.indices(self.shape[0])
            if step != 1:
                raise IndexError("only unit-step slices of a DeviceArray are views")
            b = max(a, b)
            return DeviceArray(self._block, self._offset + a * self._row_bytes(),
                               (b - a,) + self.shape[1:], self.dtype)
        raise IndexError("DeviceArray supports integer and slice indexing on its first axis")

    def reshape(self, *shape):
        if len(shape) == 1 and not np.isscalar(shape[0]):
            shape = tuple(shape[0])
        shape = list(int(s) for s in shape)
        n = self.numel()
        if shape.count(-1) == 1:
            known = int(np.prod([s for s in shape if s != -1], dtype=np.int64))
            shape[shape.index(-1)] = n // known if known else 0
        if int(np.prod(shape, dtype=np.int64)) != n:
            raise ValueError(f"cannot reshape {self.shape} to {tuple(shape)}")
        return DeviceArray(self._block, self._offset, shape, self.dtype)

    view = reshape

    def clone(self):
        d = DeviceArray.empty(self.shape, self.dtype)
        if self.nbytes:
            call("hr_memcpy_d2d", ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(self.data_ptr()),
                 self.nbytes, None)
        return d

    def copy_(self, src):
        if isinstance(src, np.ndarray):
            a = np.ascontiguousarray(src, dtype=self.dtype)
            if a.nbytes != self.nbytes:
                raise ValueError("copy_ between arrays of different sizes")
            call("hr_memcpy_h2d", ctypes.c_void_p(self.data_ptr()),
                 a.ctypes.data_as(ctypes.c_void_p), a.nbytes, None)
            return self
        n = src.numel() * src.element_size()
        if n != self.nbytes:
            raise ValueError("copy_ between arrays of different sizes")
        if is_torch_cuda(src):
            torch().cuda.current_stream().synchronize()
        call("hr_memcpy_d2d", ctypes.c_void_p(self.data_ptr()), ctypes.c_void_p(src.data_ptr()),
             n, None)
        return self

    def zero_(self):
        call("hr_memset", ctypes.c_void_p(self.data_ptr()), 0, self.nbytes, None)
        return self

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            call("hr_memcpy_d2h", out.ctypes.data_as(ctypes.c_void_p),
                 ctypes.c_void_p(self.data_ptr()), out.nbytes, None)
            call("hr_stream_synchronize", None)
        return out

    def cpu(self):
        """Host copy (a numpy array that also answers ``.numpy()``)."""
        return self.numpy().view(_HostArray)

    def fill_(self, value):
        return self.copy_(np.full(self.shape, value, self.dtype))

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


class _HostArray(np.ndarray):
    """numpy array returned by DeviceArray.cpu(), with the tensor-style
    ``.numpy()``."""

    def numpy(self):
        return self.view(np.ndarray)


# ------------------------------------------------------------- selection
class using:
    """Context: buffers allocated inside follow the kind of the given inputs
    (torch tensors when any is a torch CUDA tensor, else DeviceArray)."""

    def __init__(self, *xs):
        self.use_torch = any(is_torch_cuda(x) for x in xs)

    def __enter__(self):
        self.prev = getattr(_tls, "torch", False)
        _tls.torch = self.use_torch
        return self

    def __exit__(self, *exc):
        _tls.torch = self.prev


def _torch_mode():
    return getattr(_tls, "torch", False)


def is_torch_mode():
    """Whether buffers allocated now are torch tensors."""
    return _torch_mode()


def follows(pos, name):
    """Decorator: inside the call, buffers follow the kind of the argument at
    position `pos` (keyword `name`), or of its ``.xyz``."""
    def deco(fn):
        @functools.wraps(fn)
        def wrapper(*args, **kwargs):
            x = args[pos] if len(args) > pos else kwargs.get(name)
            x = getattr(x, "xyz", x)
            prev = getattr(_tls, "torch", False)
            _tls.torch = is_torch_cuda(x)
            try:
                return fn(*args, **kwargs)
            finally:
                _tls.torch = prev
        return wrapper
    return deco


def any_nonzero(x):
    """Whether device or host array x holds a non-zero element."""
    if is_torch_tensor(x):
        return bool(x.any())
    return bool(np.any(to_host(x)))


def indices_on_host(idx):
    """int64 numpy copy of an index array held anywhere."""
    return np.asarray(to_host(idx)).astype(np.int64)


def gpu_count():
    n = ctypes.c_int(0)
    call("hr_device_count", ctypes.byref(n))
    return n.value


_gpu_seen = False


def require_gpu():
    """RuntimeError unless a GPU is visible (no CPU fallback)."""
    global _gpu_seen
    if _gpu_seen:
        return
    if gpu_count() < 1:
        raise RuntimeError("mdtraj_amd computes on the GPU only and no GPU is visible")
    _gpu_seen = True


def require_torch_gpu():
    """torch, for modules that run on torch streams and events (serving);
    RuntimeError unless it sees a GPU."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("mdtraj_amd computes on the GPU only and no GPU is visible")
    return t


class torch_buffers:
    """Context: buffers allocated inside are torch tensors (serving, whose
    batches run on torch streams)."""

    def __enter__(self):
        self.prev = getattr(_tls, "torch", False)
        _tls.torch = True
        return self

    def __exit__(self, *exc):
        _tls.torch = self.prev


def is_device_array(x) -> bool:
    if x is None or isinstance(x, np.ndarray):
        return False
    return isinstance(x, DeviceArray) or is_torch_cuda(x)


def device():
    t = torch()
    return t.device("cuda", t.cuda.current_device())


def _to_device(x, np_dtype, torch_dtype_name):
    if isinstance(x, DeviceArray):
        if x.dtype == np_dtype:
            return x
        return DeviceArray.from_host(x.numpy().astype(np_dtype))
    if is_torch_tensor(x):
        t = torch()
        y = x if x.is_cuda else x.to(device())
        dt = getattr(t, torch_dtype_name)
        if y.dtype != dt:
            y = y.to(dt)
        return y.contiguous()
    a = np.ascontiguousarray(x, dtype=np_dtype)
    if _torch_mode():
        return torch().from_numpy(a).to(device())
    require_gpu()
    return DeviceArray.from_host(a)


class using_loaded:
    """Context for calls with no array input (file readers): torch tensors
    when torch is already imported, else DeviceArray."""

    def __enter__(self):
        self.prev = getattr(_tls, "torch", False)
        _tls.torch = _torch_loaded() is not None
        return self

    def __exit__(self, *exc):
        _tls.torch = self.prev


def to_device(x, dtype):
    """C-contiguous device copy of host array x with numpy dtype `dtype`."""
    dt = np.dtype(dtype)
    name = {np.dtype(np.float32): "float32", np.dtype(np.int32): "int32",
            np.dtype(np.int64): "int64", np.dtype(np.uint8): "uint8",
            np.dtype(np.float64): "float64"}[dt]
    return _to_device(x, dt, name)


def to_device_f32(x):
    """float32 C-contiguous device array for x (numpy, tensor or DeviceArray);
    no copy if x already is one."""
    return _to_device(x, np.float32, "float32")


def to_device_i32(idx):
    if idx is None:
        return None
    return _to_device(idx, np.int32, "int32")


def empty(shape, dtype=np.float32):
    if _torch_mode():
        t = torch()
        tdt = {np.dtype(np.float32): t.float32, np.dtype(np.int32): t.int32,
               np.dtype(np.int64): t.int64, np.dtype(np.uint8): t.uint8,
               np.dtype(np.float64): t.float64}[np.dtype(dtype)]
        return t.empty(shape if not np.isscalar(shape) else int(shape), dtype=tdt, device=device())
    require_gpu()
    return DeviceArray.empty(shape, dtype)


def empty_f32(n):
    return empty(int(n), np.float32)


def workspace(nbytes: int):
    return empty(max(int(nbytes), 256), np.uint8)


def ptr(x):
    return None if x is None else x.data_ptr()


def stream():
    """Stream the kernels launch on: torch's current stream for torch
    buffers, else the default stream (NULL)."""
    if _torch_mode():
        return torch().cuda.current_stream().cuda_stream
    return None


def synchronize():
    if _torch_mode():
        torch().cuda.synchronize()
    else:
        call("hr_stream_synchronize", None)


def to_host(x):
    if isinstance(x, DeviceArray):
        return x.numpy()
    if is_torch_tensor(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def release_cached_blocks():
    """Free the library buffers kept for reuse."""
    _cache.release()
