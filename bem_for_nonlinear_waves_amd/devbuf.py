"""Device vectors for the benchmark and the tests, through the HIP runtime that libbemfmm.so is linked with.

libbemfmm.so is built against the ROCm HIP runtime of this image (/opt/rocm). Another HIP runtime with the
same soname loaded into the process first (PyTorch bundles one) would be the one the library binds to, so
device memory for the C ABI's device-pointer entry points is taken from the library's own runtime here.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def hip():
    """The HIP runtime loaded by libbemfmm.so (resolved through the library's own dependency)."""
    global _hip
    if _hip is None:
        _lib.load()
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return _hip


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """n doubles of device memory; freed by free() or when garbage-collected."""

    def __init__(self, n: int):
        self.n = int(n)
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(1, self.n) * 8)), "hipMalloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a, np.float64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        _check(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, _H2D), "hipMemcpy H2D")

    def download(self) -> np.ndarray:
        a = np.empty(self.n)
        _check(hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes, _D2H), "hipMemcpy D2H")
        return a

    def copy_from(self, other: "DeviceArray"):
        assert other.n == self.n
        _check(hip().hipMemcpy(self.ptr, other.ptr, self.n * 8, _D2D), "hipMemcpy D2D")

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def memory_in_use() -> float:
    """Bytes of device memory in use on the current device."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(hip().hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)), "hipMemGetInfo")
    return float(t.value - f.value)
