"""Minimal HIP runtime access (ctypes) for device buffers in the bench and
tests: allocation, copies, synchronisation."""
import ctypes

import numpy as np

_hip = None


def lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().hipGetErrorString(rc).decode()}")


class DeviceArray:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        check(lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 8)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        check(lib().hipMemcpy(d.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        check(lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def synchronize():
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def copy_async(dst, src, nbytes, kind, stream):
    """hipMemcpyAsync between raw addresses on `stream`."""
    check(lib().hipMemcpyAsync(dst, src, int(nbytes), kind, stream), "hipMemcpyAsync")


def stream_synchronize(stream):
    check(lib().hipStreamSynchronize(stream), "hipStreamSynchronize")
