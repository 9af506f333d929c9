"""Device buffers and streams for callers of the *_device entry points, through
the HIP runtime that libgwn.so itself links (libamdhip64.so.7 from
/opt/rocm): loading it by its soname after libgwn.so returns the copy already
in the process, so buffers, streams and the library share one runtime.  (torch
bundles a HIP runtime of its own under the same soname; a process that loads
torch first binds the library to that other copy.)"""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    """The HIP runtime shared with libgwn.so."""
    global _hip
    if _hip is None:
        from . import _native as nat
        nat.lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: HIP error {rc}")


def synchronize():
    """Waits for all work on the device."""
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream at the default priority, non-blocking with respect to the
    null stream (hipStreamNonBlocking)."""

    def __init__(self):
        self.handle = C.c_void_p()
        _ok(hip().hipStreamCreateWithFlags(C.byref(self.handle), 1), "hipStreamCreateWithFlags")

    @property
    def ptr(self):
        return self.handle.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            hip().hipStreamSynchronize(self.handle)
            hip().hipStreamDestroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """hipMalloc'ed array of a fixed shape and dtype."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(int(s) for s in np.atleast_1d(shape)), np.dtype(dtype)
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
        """Copies to a new host array after all device work has finished."""
        out = np.empty(self.shape, self.dtype)
        synchronize()
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other):
        """Device-to-device copy of another array of the same size."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _ok(hip().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D), "hipMemcpy D2D")

    @property
    def data_ptr(self):
        return self.ptr.value

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
