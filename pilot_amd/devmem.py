"""Device buffers through the HIP runtime the renderer links against (ROCm's
libamdhip64), for callers that hand device pointers to the C ABI.

torch bundles its own HIP runtime; mixing it with the system one in a process is
avoided, so device memory for BEAMS_MEM_DEVICE calls is allocated here.
"""

import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = None
        for cand in ("/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            if os.path.exists(cand):
                path = cand
                break
        if path is None:
            raise ImportError("ROCm HIP runtime (libamdhip64) not found")
        h = C.CDLL(path, mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.restype = C.c_char_p
        _hip = h
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _ok(e, what):
    if e != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(e).decode()}")


class DeviceArray:
    """A flat device allocation with numpy-style shape/dtype bookkeeping."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 256)), "hipMalloc")
        self.ptr = p.value
        self.is_cuda = True

    def data_ptr(self):
        return self.ptr

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def zero(self):
        _ok(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        return self

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
