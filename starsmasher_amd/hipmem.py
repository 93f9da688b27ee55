"""Device buffers through the HIP runtime that libsphgrav_hip.so itself uses.

torch ships its own HIP runtime (torch/lib/libamdhip64.so); the library links
ROCm's (/opt/rocm/lib/libamdhip64.so.7).  Two runtimes in one process must not
exchange device pointers, so device-resident inputs for the sg_* device API
are allocated here, with the library's own runtime, found by reading which
libamdhip64 the loaded library mapped.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_hip = None

H2D = 1
D2H = 2
D2D = 3


def runtime() -> C.CDLL:
    global _hip
    if _hip is None:
        _lib.load()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "/torch/" not in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("libamdhip64 used by libsphgrav_hip.so not found in the process maps")
        h = C.CDLL(path)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [C.c_int]
        for fn in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize, h.hipSetDevice):
            fn.restype = C.c_int
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """A float64 (or other dtype) array in device memory."""

    def __init__(self, size: int, dtype=np.float64, device: int = 0):
        self.size = int(size)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(1, self.size * self.dtype.itemsize)
        h = runtime()
        _ok(h.hipSetDevice(int(device)), "hipSetDevice")
        p = C.c_void_p()
        _ok(h.hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = int(p.value)

    @classmethod
    def from_host(cls, a, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, device)
        d.upload(a)
        return d

    def upload(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError("size mismatch")
        _ok(runtime().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D),
            "hipMemcpy H2D")

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=self.dtype)
        _ok(runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, D2H),
            "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        """Device-to-device copy (the runtime's plain copy kernel)."""
        n = min(self.nbytes, other.nbytes)
        _ok(runtime().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), n, D2D),
            "hipMemcpy D2D")

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            runtime().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
