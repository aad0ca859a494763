"""Minimal device-memory plumbing over the HIP runtime that libsogrand.so links.

The library links ROCm's libamdhip64.so.7; loading a second HIP runtime in the
same process (for example the one bundled with a PyTorch wheel) makes the
device unavailable to one of them. Tests and benchmarks therefore allocate
device buffers through this module, which binds the very runtime the library
already loaded.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import load_library

_hip = None

H2D = 1
D2H = 2
D2D = 3


def runtime() -> C.CDLL:
    global _hip
    if _hip is None:
        load_library()  # pulls in libamdhip64.so.7; the next load returns the same handle
        lib = C.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        lib.hipSetDevice.argtypes = [C.c_int]
        lib.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        lib.hipStreamDestroy.argtypes = [C.c_void_p]
        lib.hipStreamSynchronize.argtypes = [C.c_void_p]
        lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        lib.hipEventDestroy.argtypes = [C.c_void_p]
        lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        lib.hipEventSynchronize.argtypes = [C.c_void_p]
        lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip = lib
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def device_count() -> int:
    n = C.c_int(0)
    rc = runtime().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize() -> None:
    _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


def copy(dst, src, nbytes: int) -> None:
    """Device-to-device copy of nbytes (the runtime's copy kernel)."""
    _ok(runtime().hipMemcpy(dst.data_ptr(), src.data_ptr(), nbytes, D2D), "hipMemcpy D2D")


class DeviceArray:
    """A contiguous device buffer with a numpy shape/dtype."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self._p = p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def data_ptr(self) -> int:
        return self._p.value

    def is_contiguous(self) -> bool:
        return True

    def numel(self) -> int:
        return int(np.prod(self.shape))

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _ok(runtime().hipMemcpy(self._p, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _ok(runtime().hipMemcpy(out.ctypes.data, self._p, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        _ok(runtime().hipMemset(self._p, 0, self.nbytes), "hipMemset")

    def free(self) -> None:
        if self._p is not None and self._p.value:
            runtime().hipFree(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(runtime().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s

    def synchronize(self) -> None:
        _ok(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            runtime().hipStreamDestroy(self.handle)
        except Exception:
            pass
