"""Device buffers and synchronisation for callers of the device-resident
entry points (`*_device`), through the HIP runtime the library itself links
(libamdhip64.so.7 from ROCm).  Using that one runtime keeps a process to a
single HIP/HSA instance; a second runtime bundled with another package (for
example a framework wheel) must not be initialised in the same process.
"""
import ctypes as C

import numpy as np

from . import _lib

_HIP = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    """The HIP runtime the library links (loaded with it)."""
    global _HIP
    if _HIP is None:
        _lib.lib()  # loads libamdhip64.so.7 as a dependency
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [C.c_int]
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        _HIP = h
    return _HIP


def check(err):
    if err != 0:
        raise RuntimeError(f"HIP error {err}: {hip().hipGetErrorString(err).decode()}")


def synchronize():
    check(hip().hipDeviceSynchronize())


class DeviceArray:
    """A 1-D float64 / int array in HBM."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n, 1) * self.dtype.itemsize
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.n
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D))

    def download(self, count=None):
        out = np.empty(self.n if count is None else count, dtype=self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H))
        return out

    def free(self):
        if self.ptr:
            check(hip().hipFree(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """Elapsed device time between two events on the null stream."""

    def __init__(self):
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        check(hip().hipEventCreate(C.byref(self.e0)))
        check(hip().hipEventCreate(C.byref(self.e1)))

    def start(self):
        check(hip().hipEventRecord(self.e0, None))

    def stop_ms(self):
        check(hip().hipEventRecord(self.e1, None))
        check(hip().hipEventSynchronize(self.e1))
        ms = C.c_float()
        check(hip().hipEventElapsedTime(C.byref(ms), self.e0, self.e1))
        return ms.value
