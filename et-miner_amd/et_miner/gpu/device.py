"""Device memory and streams for the counting path.

PyTorch (ROCm) provides the allocations and the stream handles: every device
buffer is a torch tensor on ``cuda:N`` whose ``data_ptr()`` crosses the C ABI.
This module has no kernels of its own.
"""

from __future__ import annotations

import numpy as np

_TORCH = None


def torch():
    global _TORCH
    if _TORCH is None:
        import torch as t

        _TORCH = t
    return _TORCH


def require_gpu() -> None:
    """Raises unless a HIP device and libetm.so are both present (the path has no CPU fallback)."""
    from .._lib import lib

    lib()
    if not torch().cuda.is_available():
        raise RuntimeError("the GPU counting path needs a ROCm device; none is visible")


def device():
    return torch().device("cuda", torch().cuda.current_device())


def stream_ptr():
    """The current HIP stream as an integer handle for the C ABI."""
    return torch().cuda.current_stream().cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


_DT = {
    np.dtype(np.int32): "int32",
    np.dtype(np.int64): "int64",
    np.dtype(np.uint8): "uint8",
    np.dtype(np.uint64): "uint64",
}


def empty(n, dtype: str):
    t = torch()
    shape = (n,) if isinstance(n, int) else tuple(n)
    return t.empty(shape, dtype=getattr(t, dtype), device=device())


def zeros(n, dtype: str):
    t = torch()
    shape = (n,) if isinstance(n, int) else tuple(n)
    return t.zeros(shape, dtype=getattr(t, dtype), device=device())


#: Arrays of at least this many bytes go through ``etm_upload`` (chunked pinned staging, parallel host copies).
UPLOAD_CHUNKED_MIN = 4 << 20


def to_device(a: np.ndarray, pinned: bool = True):
    """A host array copied to the device (through pinned memory, asynchronously on the current stream).

    Large arrays are staged chunk by chunk through libetm's pinned ring (``etm_upload``): no pinned copy of the
    whole array is made, and the host copy of each chunk overlaps the DMA of the previous one. uint64 arrays
    (bitvector words) travel as int64 tensors: the kernels read the same bits."""
    t = torch()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    if pinned and a.nbytes >= UPLOAD_CHUNKED_MIN:
        from .. import _lib

        out = t.empty(a.shape, dtype=t.from_numpy(a[:0].reshape(-1)).dtype, device=device())
        _lib.call("etm_upload", a.ctypes.data, a.nbytes, out.data_ptr(), stream_ptr())
        return out
    host = t.from_numpy(a)
    if pinned and host.numel() > 0:
        host = host.pin_memory()
    return host.to(device(), non_blocking=pinned)


class _RawDevice:
    """A device range described by ``__cuda_array_interface__`` (for torch.as_tensor)."""

    def __init__(self, p: int, n: int, typestr: str):
        self.__cuda_array_interface__ = {"shape": (int(n),), "typestr": typestr, "data": (int(p), False),
                                         "version": 2, "strides": None}


def copy_from_pointer(p: int, n: int, dtype: str = "int32"):
    """A new device tensor holding the n elements at device address p, copied on the current stream."""
    t = torch()
    typestr = {"int32": "<i4", "int64": "<i8", "uint8": "|u1"}[dtype]
    if n == 0:
        return t.empty(0, dtype=getattr(t, dtype), device=device())
    return t.as_tensor(_RawDevice(p, n, typestr), device=device()).clone()


def to_host(x) -> np.ndarray:
    """A device tensor copied to a numpy array (synchronizes)."""
    return x.cpu().numpy()


def download(x) -> np.ndarray:
    """A contiguous device tensor copied into a new (pageable) numpy array through the library's pinned ring
    (``etm_download``: DMAs in flight while host threads copy out the chunks that landed), after the work queued
    on the current stream; returns with the data in place."""
    from .. import _lib

    t = torch()
    x = x.contiguous()
    out = np.empty(tuple(x.shape), dtype=t.empty(0, dtype=x.dtype).numpy().dtype)
    if out.nbytes:
        _lib.call("etm_download", x.data_ptr(), out.nbytes, out.ctypes.data, stream_ptr())
    return out


def synchronize() -> None:
    torch().cuda.synchronize()


class CudaArray:
    """A borrowed device array described by ``__cuda_array_interface__`` (pointer, shape, dtype, strides)."""

    def __init__(self, obj):
        cai = obj.__cuda_array_interface__
        self.obj = obj
        self.ptr = int(cai["data"][0])
        self.shape = tuple(int(x) for x in cai["shape"])
        self.typestr = cai["typestr"]
        self.strides = cai.get("strides")
        self.dtype = np.dtype(self.typestr)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def c_contiguous(self) -> bool:
        if self.strides is None:
            return True
        expect, step = [], self.dtype.itemsize
        for dim in reversed(self.shape):
            expect.append(step)
            step *= dim
        return tuple(self.strides) == tuple(reversed(expect))


def describe(obj) -> CudaArray:
    """A device array from a torch tensor or any object exposing ``__cuda_array_interface__``."""
    t = torch()
    if isinstance(obj, t.Tensor):
        if obj.device.type != "cuda":
            raise TypeError("bitvecs_gpu must live on the GPU, got a tensor on " + str(obj.device))
        arr = CudaArray.__new__(CudaArray)
        arr.obj = obj
        arr.ptr = obj.data_ptr()
        arr.shape = tuple(obj.shape)
        np_dtype = {t.uint64: np.uint64, t.int64: np.int64, t.int32: np.int32, t.uint8: np.uint8}.get(obj.dtype)
        arr.dtype = np.dtype(np_dtype) if np_dtype is not None else None
        arr.typestr = arr.dtype.str if arr.dtype is not None else str(obj.dtype)
        arr.strides = None if obj.is_contiguous() else tuple(s * obj.element_size() for s in obj.stride())
        return arr
    if hasattr(obj, "__cuda_array_interface__"):
        return CudaArray(obj)
    raise TypeError(
        f"bitvecs_gpu must be a device array (has __cuda_array_interface__), got {type(obj).__name__}"
    )
