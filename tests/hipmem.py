"""Device buffers for tests that pass device pointers through the C ABI
(kb_sunset, kb_sunset_noalloc), allocated with the HIP runtime that
libqftkb.so itself links: one runtime per process.  (torch bundles a HIP
runtime of its own; a process holding both fails to initialise the second.)
Test plumbing only."""
import ctypes

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        from qftkb import load_library
        load_library()  # libqftkb.so first: the runtime below is the one it linked
        L = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.hipSetDevice.argtypes = [ctypes.c_int]
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """float64 (or raw byte) device buffer."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = cls(a.size)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, H2D), "hipMemcpy H2D")
        return d

    def fill(self, value):
        self.upload(np.full(self.n, value, dtype=self.dtype))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        assert a.size == self.n
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def host(self):
        out = np.empty(self.n, dtype=self.dtype)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
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
