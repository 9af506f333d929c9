"""Device buffers for the tests of the *_device entry points, through the HIP
runtime libgwn.so itself links (libamdhip64.so.7 from /opt/rocm): loading it
by its soname after libgwn.so returns the copy already in the process, so the
buffers and libgwn share one runtime.  (torch bundles its own HIP runtime; a
process that initialises both can lose the device in the second.)"""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from libigl_amd import _native as nat
        nat.lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


H2D, D2H = 1, 2


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {rc}")


class DeviceArray:
    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    @property
    def data_ptr(self):
        return self.ptr.value

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
