"""Device buffers from the HIP runtime that libidb_hip.so itself links.

The PyTorch wheel ships its own copy of the HIP runtime; device pointers are
therefore allocated here, through the runtime the library uses, so that
HBM-resident inputs and outputs never depend on a second runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib  # noqa: F401  (loads libidb_hip.so and with it libamdhip64.so.7)

_hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipMalloc.restype = C.c_int
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipFree.restype = C.c_int
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipMemcpy.restype = C.c_int
_hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
_hip.hipMemset.restype = C.c_int
_hip.hipDeviceSynchronize.argtypes = []
_hip.hipDeviceSynchronize.restype = C.c_int
_hip.hipSetDevice.argtypes = [C.c_int]
_hip.hipSetDevice.restype = C.c_int

H2D, D2H = 1, 2


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def copy_device(dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int) -> None:
    """Device-to-device copy of nbytes (runs the runtime's copy kernel)."""
    _check(_hip.hipMemcpy(dst._p, src._p, int(nbytes), 3), "hipMemcpy D2D")


def synchronize() -> None:
    _check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """An owning device allocation; data_ptr() gives the device address."""

    def __init__(self, nbytes: int, device: int = 0):
        _check(_hip.hipSetDevice(device), "hipSetDevice")
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        _check(_hip.hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self._p = p

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, device)
        if a.nbytes:
            _check(_hip.hipMemcpy(b._p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return b

    @classmethod
    def zeros(cls, nbytes: int, device: int = 0) -> "DeviceBuffer":
        b = cls(nbytes, device)
        _check(_hip.hipMemset(b._p, 0, b.nbytes), "hipMemset")
        return b

    def data_ptr(self) -> int:
        return int(self._p.value)

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype)
        if out.nbytes:
            _check(_hip.hipMemcpy(out.ctypes.data, self._p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if getattr(self, "_p", None) is not None and self._p.value:
            _hip.hipFree(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
