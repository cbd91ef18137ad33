"""Device memory for callers that keep particle arrays resident in HBM.

The arrays are allocated with the same HIP runtime that libphsph.so links
(/opt/rocm/lib/libamdhip64.so.7, loaded once per process), so their
pointers are recognised by the library as device memory and used in place.
(PyTorch wheels bundle a separate HIP runtime; tensors from it are not
passed to the library.)
"""
import ctypes as C

import numpy as np

from . import sph

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


def hip():
    global _hip
    if _hip is None:
        sph.load_library()  # brings libamdhip64.so.7 into the process
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        L.hipGetErrorString.restype = C.c_char_p
        L.hipGetErrorString.argtypes = [C.c_int]
        _hip = L
    return _hip


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A contiguous device array with a numpy-like shape and dtype."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        _chk(hip().hipMemset(C.c_void_p(self.ptr), 0, max(self.nbytes, 1)), "hipMemset")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _chk(hip().hipMemcpy(C.c_void_p(d.ptr), C.c_void_p(a.ctypes.data), d.nbytes, hipMemcpyHostToDevice),
             "hipMemcpy H2D")
        return d

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _chk(hip().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), self.nbytes, hipMemcpyDeviceToDevice),
             "hipMemcpy D2D")
        synchronize()  # a device-to-device hipMemcpy may return before it completes

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _chk(hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), self.nbytes, hipMemcpyDeviceToHost),
             "hipMemcpy D2H")
        return out

    @property
    def ndim(self):
        return len(self.shape)

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
