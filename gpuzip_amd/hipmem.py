"""Device buffers through the system HIP runtime (/opt/rocm/lib/libamdhip64.so), the runtime
libgpuinflate.so is linked against. Plumbing for callers of gi_inflate_device (tests, the
benchmark): it keeps one HIP runtime per process (torch-ROCm wheels bundle a runtime of their
own, and two runtimes in one process do not share a device)."""
import ctypes
import os

import numpy as np

H2D, D2H, D2D = 1, 2, 3
_hip = None


def hip():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        h = ctypes.CDLL(path)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        _hip = h
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def mem_info():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(hip().hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)), "hipMemGetInfo")
    return f.value, t.value


class DeviceBuffer:
    """nbytes of device memory (plus `pad` zero bytes after them)"""

    def __init__(self, nbytes, pad=0):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes + pad, 1)), "hipMalloc")
        if pad:
            check(hip().hipMemset(ctypes.c_void_p(self.ptr.value + self.nbytes), 0, pad), "hipMemset")

    @property
    def addr(self):
        return self.ptr.value

    @classmethod
    def from_host(cls, a, pad=64):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, pad)
        check(hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return b

    def to_host(self, n=None, dtype=np.uint8, offset=0):
        n = self.nbytes - offset if n is None else n
        a = np.empty(n, np.uint8)
        if n:
            check(hip().hipMemcpy(a.ctypes.data, ctypes.c_void_p(self.ptr.value + offset), n, D2H), "hipMemcpy D2H")
        return a.view(dtype)

    def free(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
