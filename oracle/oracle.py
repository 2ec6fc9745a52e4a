"""ctypes front-end for the oracle: the CPU restatement of the reference
2PI Kadanoff-Baym step (src/TwoPIScalar, CPUfull + LambdaNLO quantum).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
as the checker, and by bench.py's cpu_baseline leg.  The product path
(libqftkb.so and the qftkb host package) never imports this module.

Times are 1-based and the window is [tmin, tmax], exactly as the reference
(src/DefSimDataElements.jl:141-165).  All lattices are float64 arrays of
length V = Nx**sdim in column-major site order (x fastest), matching a Julia
Array{Float64,sdim}.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libkboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_last_error.restype = ctypes.c_char_p
        L.or_last_error.argtypes = [ctypes.c_void_p]
        for name in ("or_get_kL2", "or_set_kL2"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _dp]
        L.or_init_particles.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.or_init_rows.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.or_evolve.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_expand.argtypes = [ctypes.c_void_p]
        L.or_window.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int)]
        L.or_selfenergies.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_selfenergies_direct.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_ninv_resum.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_ninv_selfenergies.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_get_ninv.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.or_ninv_transforms.restype = ctypes.c_longlong
        L.or_ninv_transforms.argtypes = [ctypes.c_void_p]
        L.or_hartree.restype = ctypes.c_double
        L.or_hartree.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_set_hartree.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_get_hartree.restype = ctypes.c_double
        L.or_get_hartree.argtypes = [ctypes.c_void_p]
        L.or_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.or_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.or_get_sigma.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.or_set_sigma.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.or_measure.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                 ctypes.POINTER(ctypes.c_int)]
        L.or_dft.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.or_fill_free_window.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Oracle:
    """One reference-equivalent CPUfull simulation state.

    pexp: 0 = LambdaLO, 1 = LambdaNLO (quantum sunset), 2 = NinverseLO,
    3 = NinverseNLO, 4 = LambdaNLO classical sunset.
    """

    def __init__(self, Nx, sdim, Nmem, dt, lam, ONgroup, Mass, Mass2, pexp=1, nthreads=None):
        if nthreads is None:
            nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        self.Nx, self.sdim, self.N = int(Nx), int(sdim), int(Nmem)
        self.V = self.Nx ** self.sdim
        self.dt, self.lam, self.ON, self.Mass, self.Mass2 = dt, lam, ONgroup, Mass, Mass2
        self._L = lib()
        self._c = self._L.or_create(self.Nx, self.sdim, self.N, float(dt), float(lam), int(ONgroup),
                                    float(Mass), float(Mass2), int(pexp), int(nthreads))
        if not self._c:
            raise MemoryError("or_create failed (bad parameters or out of memory)")

    def close(self):
        if getattr(self, "_c", None):
            self._L.or_destroy(self._c)
            self._c = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.or_last_error(self._c).decode())

    def kL2(self):
        out = np.empty(self.V)
        self._chk(self._L.or_get_kL2(self._c, _ptr(out)))
        return out

    def set_kL2(self, k2):
        k2 = np.ascontiguousarray(k2, dtype=np.float64)
        self._chk(self._L.or_set_kL2(self._c, _ptr(k2)))

    def init_particles(self, n, initmass):
        n = np.ascontiguousarray(n, dtype=np.float64)
        assert n.size == self.V
        self._chk(self._L.or_init_particles(self._c, _ptr(n), float(initmass)))

    def init_rows(self, F11, F21, F22):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (F11, F21, F22)]
        self._chk(self._L.or_init_rows(self._c, *[_ptr(x) for x in a]))

    def fill_free_window(self, tmax, n=None, mass=1.0):
        nn = None if n is None else np.ascontiguousarray(n, dtype=np.float64)
        self._chk(self._L.or_fill_free_window(self._c, int(tmax), None if nn is None else _ptr(nn),
                                              float(mass)))

    def evolve(self, t):
        self._chk(self._L.or_evolve(self._c, int(t)))

    def expand(self):
        self._chk(self._L.or_expand(self._c))

    def window(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._L.or_window(self._c, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def selfenergies(self, tmone, direct=False):
        f = self._L.or_selfenergies_direct if direct else self._L.or_selfenergies
        self._chk(f(self._c, int(tmone)))

    def ninv_resum(self, tmone, simple=False):
        """NinverseNLO: Irx and IFx for every tp at tmone (calcIrs!/calcIFs!,
        or the simple variants)."""
        self._chk(self._L.or_ninv_resum(self._c, int(tmone), int(bool(simple))))

    def ninv_selfenergies(self, tmone, simple=False):
        self._chk(self._L.or_ninv_selfenergies(self._c, int(tmone), int(bool(simple))))

    def ninv(self, which, tp):
        """IFx (which='F') or Irx (which='r') at tp: complex, full grid."""
        out = np.empty(2 * self.V)
        self._chk(self._L.or_get_ninv(self._c, 1 if which == "r" else 0, int(tp), _ptr(out)))
        return out.view(np.complex128).copy()

    def ninv_transforms(self):
        return int(self._L.or_ninv_transforms(self._c))

    def hartree(self, i, j):
        return self._L.or_hartree(self._c, int(i), int(j))

    @property
    def hartreemass2(self):
        return self._L.or_get_hartree(self._c)

    @hartreemass2.setter
    def hartreemass2(self, v):
        self._L.or_set_hartree(self._c, float(v))

    def get(self, which, i, j):
        """Physical F(i,j) (which='F') or rho(i,j) (which='r')."""
        out = np.empty(self.V)
        self._chk(self._L.or_get(self._c, 1 if which == "r" else 0, int(i), int(j), _ptr(out)))
        return out

    def set(self, which, i, j, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self._L.or_set(self._c, 1 if which == "r" else 0, int(i), int(j), _ptr(arr)))

    def sigma(self, which, tp):
        out = np.empty(self.V)
        self._chk(self._L.or_get_sigma(self._c, 1 if which == "r" else 0, int(tp), _ptr(out)))
        return out

    def set_sigma(self, which, tp, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self._L.or_set_sigma(self._c, 1 if which == "r" else 0, int(tp), _ptr(arr)))

    def measure(self, t):
        """createMeasurement at time t; every site is its own level."""
        scal = np.empty(6)
        arrs = {k: np.empty(self.V) for k in ("F", "ddF", "r", "n", "omega", "c")}
        nst = ctypes.c_int()
        self._chk(self._L.or_measure(self._c, int(t), _ptr(scal), _ptr(arrs["F"]), _ptr(arrs["ddF"]),
                                     _ptr(arrs["r"]), _ptr(arrs["n"]), _ptr(arrs["omega"]),
                                     _ptr(arrs["c"]), ctypes.byref(nst)))
        out = dict(time=scal[0], E=scal[1], E_pot=scal[2], E_kin=scal[3], E_mem=scal[4], M2=scal[5],
                   nstab=nst.value)
        out.update(arrs)
        return out

    def dft(self, z: np.ndarray, sign: int = -1) -> np.ndarray:
        """Unnormalised complex DFT over all sdim axes (x fastest)."""
        buf = np.ascontiguousarray(z, dtype=np.complex128).reshape(-1).copy()
        assert buf.size == self.V
        view = buf.view(np.float64)
        self._chk(self._L.or_dft(self._c, _ptr(view), int(sign)))
        return buf
