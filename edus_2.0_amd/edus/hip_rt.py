"""Minimal ctypes access to the HIP runtime the engine library links against:
device buffers and events for the device-pointer entry points (tests and the
benchmark)."""
import ctypes as C

import numpy as np

_RT = None


def rt():
    global _RT
    if _RT is None:
        from .engine import lib

        lib()  # loads libamdhip64 through the engine library first
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipEventDestroy.argtypes = [C.c_void_p]
        _RT = L
    return _RT


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {rc}")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _ok(rt().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _ok(rt().hipMemcpy(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _ok(rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


def synchronize():
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Timer:
    """HIP events on the null stream."""

    def __init__(self):
        self.a, self.b = C.c_void_p(), C.c_void_p()
        _ok(rt().hipEventCreate(C.byref(self.a)), "hipEventCreate")
        _ok(rt().hipEventCreate(C.byref(self.b)), "hipEventCreate")

    def start(self):
        _ok(rt().hipEventRecord(self.a, None), "hipEventRecord")

    def stop_ms(self) -> float:
        _ok(rt().hipEventRecord(self.b, None), "hipEventRecord")
        _ok(rt().hipEventSynchronize(self.b), "hipEventSynchronize")
        ms = C.c_float()
        _ok(rt().hipEventElapsedTime(C.byref(ms), self.a, self.b), "hipEventElapsedTime")
        return ms.value
