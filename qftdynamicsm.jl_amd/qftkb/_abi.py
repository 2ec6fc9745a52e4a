"""ctypes binding of libqftkb.so (include/qftkb.h).

The shared library is built in-tree next to this file.  There is no CPU
fallback: if the library is missing, or no HIP device is present, the calls
fail loudly.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

LIB_NAME = "libqftkb.so"
# QFTKB_LIBRARY may point at another build of the same library (A/B timing)
LIB_PATH = os.environ.get("QFTKB_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

KB_OK = 0
KB_ERR_ARG = 1
KB_ERR_WINDOW = 2
KB_ERR_OOM = 3
KB_ERR_HIP = 4
KB_ERR_STATE = 5
KB_ERR_NONFINITE = 6

KB_PEXP_LAMBDA_LO = 0
KB_PEXP_LAMBDA_NLO = 1
KB_PEXP_NINVERSE_LO = 2
KB_PEXP_NINVERSE_NLO = 3
# self-energy form of orbit-storage LambdaNLO contexts: orbit cosine products
FFT_ORBIT_PRODUCT = -1
# TwoPIScalarLambdaNLOclassical: the sunset without its rho^2 terms
KB_PEXP_LAMBDA_NLO_CLASSICAL = 4
PEXP_CODES = {"LambdaLO": KB_PEXP_LAMBDA_LO, "LambdaNLO": KB_PEXP_LAMBDA_NLO, "NinverseLO": KB_PEXP_NINVERSE_LO,
              "NinverseNLO": KB_PEXP_NINVERSE_NLO, "LambdaNLOclassical": KB_PEXP_LAMBDA_NLO_CLASSICAL}

# every symbol include/qftkb.h declares
EXPORTED = (
    "kb_create", "kb_destroy", "kb_last_error", "kb_info", "kb_get_kL2", "kb_set_kL2",
    "kb_init_particles", "kb_init_rows", "kb_step", "kb_window", "kb_hartree", "kb_measure",
    "kb_measure_history", "kb_get_entry", "kb_set_entry", "kb_get_sigma", "kb_ninv_get", "kb_read_bandwidth",
    "kb_check_finite",
    "kb_fill_free_window", "kb_window_bytes", "kb_save_window", "kb_load_window", "kb_synchronize", "kb_set_stream", "kb_enable_timing", "kb_timing",
    "kb_reset_timing", "kb_use_graph", "kb_ens_create", "kb_ens_create_sym", "kb_ens_set_tiles", "kb_ens_get_tiles", "kb_ens_autotune", "kb_ens_destroy", "kb_ens_last_error", "kb_ens_member",
    "kb_ens_step", "kb_ens_use_graph", "kb_ens_synchronize", "kb_ens_enable_timing", "kb_ens_timing",
    "kb_adj_record", "kb_adj_record_host", "kb_adj_gradient", "kb_adj_info", "kb_adj_stop", "kb_adj_debug_stop", "kb_adj_get_entry",
    "kb_tile_shape", "kb_set_tiles", "kb_get_tiles", "kb_autotune",
    "kb_fft_shape", "kb_set_fft_tiles", "kb_get_fft_tiles", "kb_autotune_fft",
    "kb_create_sym", "kb_storage_info", "kb_sunset_workspace_bytes", "kb_sunset", "kb_sunset_noalloc",
)


class KBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libqftkb error {code}: {msg}")
        self.code = code


class kb_params(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("sdim", ctypes.c_int), ("NstepsinMemory", ctypes.c_int),
        ("ONgroup", ctypes.c_int), ("dt", ctypes.c_double), ("Lambda", ctypes.c_double),
        ("Mass", ctypes.c_double), ("Mass2", ctypes.c_double), ("pexp", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class kb_meas(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("E", ctypes.c_double), ("E_pot", ctypes.c_double),
        ("E_kin", ctypes.c_double), ("E_mem", ctypes.c_double), ("M2", ctypes.c_double),
        ("n_unstable", ctypes.c_int),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def load_library(path: str = LIB_PATH):
    """Load libqftkb.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.kb_create.argtypes = [ctypes.POINTER(kb_params), ctypes.POINTER(vp)]
    L.kb_create_sym.argtypes = [ctypes.POINTER(kb_params), ctypes.POINTER(vp)]
    L.kb_storage_info.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.kb_destroy.argtypes = [vp]
    L.kb_last_error.restype = ctypes.c_char_p
    L.kb_last_error.argtypes = [vp]
    L.kb_info.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.kb_get_kL2.argtypes = [vp, _dp]
    L.kb_set_kL2.argtypes = [vp, _dp]
    L.kb_init_particles.argtypes = [vp, _dp, ctypes.c_double]
    L.kb_init_rows.argtypes = [vp, _dp, _dp, _dp]
    L.kb_step.argtypes = [vp, ctypes.c_int]
    L.kb_window.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.kb_hartree.argtypes = [vp, _dp]
    L.kb_measure.argtypes = [vp, ctypes.c_int, ctypes.POINTER(kb_meas), _dp, _dp, _dp, _dp, _dp, _dp]
    L.kb_measure_history.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.kb_get_entry.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.kb_set_entry.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.kb_get_sigma.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp]
    L.kb_ninv_get.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.kb_read_bandwidth.argtypes = [ctypes.c_int, ctypes.c_longlong, _dp]
    L.kb_check_finite.argtypes = [vp]
    L.kb_fill_free_window.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_double]
    L.kb_window_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong)]
    L.kb_save_window.argtypes = [vp, vp, ctypes.c_ulonglong]
    L.kb_load_window.argtypes = [vp, vp, ctypes.c_ulonglong]
    L.kb_synchronize.argtypes = [vp]
    L.kb_set_stream.argtypes = [vp, vp]
    L.kb_enable_timing.argtypes = [vp, ctypes.c_int]
    L.kb_timing.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_longlong)]
    L.kb_reset_timing.argtypes = [vp]
    L.kb_use_graph.argtypes = [vp, ctypes.c_int]
    L.kb_tile_shape.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int)]
    L.kb_set_tiles.argtypes = [vp, ctypes.c_int]
    L.kb_get_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.kb_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]
    L.kb_fft_shape.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.kb_set_fft_tiles.argtypes = [vp, ctypes.c_int]
    L.kb_get_fft_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.kb_autotune_fft.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]
    L.kb_adj_record.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.kb_adj_record_host.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.kb_adj_gradient.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.kb_adj_info.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.kb_adj_stop.argtypes = [vp]
    L.kb_adj_debug_stop.argtypes = [vp, ctypes.c_int]
    L.kb_adj_get_entry.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.kb_ens_create.argtypes = [ctypes.POINTER(kb_params), ctypes.c_int, ctypes.POINTER(vp)]
    L.kb_ens_create_sym.argtypes = [ctypes.POINTER(kb_params), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                    ctypes.POINTER(vp)]
    L.kb_ens_set_tiles.argtypes = [vp, ctypes.c_int]
    L.kb_ens_get_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.kb_ens_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    L.kb_ens_destroy.argtypes = [vp]
    L.kb_ens_last_error.restype = ctypes.c_char_p
    L.kb_ens_last_error.argtypes = [vp]
    L.kb_ens_member.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.kb_ens_step.argtypes = [vp, ctypes.c_int]
    L.kb_ens_use_graph.argtypes = [vp, ctypes.c_int]
    L.kb_ens_synchronize.argtypes = [vp]
    L.kb_ens_enable_timing.argtypes = [vp, ctypes.c_int]
    L.kb_ens_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]
    L.kb_sunset_workspace_bytes.argtypes = [ctypes.POINTER(kb_params), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_ulonglong)]
    L.kb_sunset.argtypes = [ctypes.POINTER(kb_params), ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_ulonglong, vp]
    L.kb_sunset_noalloc.argtypes = [ctypes.POINTER(kb_params), ctypes.c_int, vp, vp, vp, vp, vp]
    _lib = L
    return L


TILES_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles.json")


def tiles_key(Nx, sdim, NstepsinMemory, symmetric=False):
    return f"{int(Nx)}^{int(sdim)} N_mem={int(NstepsinMemory)}" + (" orbits" if symmetric else "")


def read_bandwidth(device=0, nbytes=16 << 30):
    """Streaming-read rate of the device in GB/s (kb_read_bandwidth)."""
    L = load_library()
    out = ctypes.c_double()
    rc = L.kb_read_bandwidth(int(device), int(nbytes), ctypes.byref(out))
    if rc != KB_OK:
        raise KBError(rc, "kb_read_bandwidth failed")
    return out.value


def committed_tiles(Nx, sdim, NstepsinMemory, symmetric=False):
    """(memory-integral variant, self-energy variant) chosen by the autotuner
    for this lattice and window (tiles.json beside the library, written by
    `bench.py --autotune`); shapes are matched by value, None if absent."""
    try:
        with open(TILES_PATH) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None, None
    choices = table.get("choices", {})
    e = choices.get(tiles_key(Nx, sdim, NstepsinMemory, symmetric))
    only_fft = False
    if e is None and symmetric:
        # orbit storage transforms the same full lattice: reuse the full-storage
        # self-energy choice (its memory integral runs in the split form)
        e = choices.get(tiles_key(Nx, sdim, NstepsinMemory, False))
        only_fft = True
    if e is None:
        return None, None
    L = load_library()
    v = f = None
    r, b, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    i = 0
    while L.kb_tile_shape(i, ctypes.byref(r), ctypes.byref(b), ctypes.byref(n)) == 0:
        if (r.value, b.value, n.value) == (e.get("rows"), e.get("batch"), e.get("threads", 256)):
            v = i
        i += 1
    ff = e.get("fft")
    if ff is not None:
        i = 0
        while L.kb_fft_shape(i, ctypes.byref(n)) == 0:
            if n.value == ff.get("threads"):
                f = i
            i += 1
    return (None if only_fft else v), f


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a, V):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if a.size != V:
        raise ValueError(f"expected {V} lattice values, got {a.size}")
    return a


def make_params(Nx, sdim, NstepsinMemory, dt, Lambda, ONgroup, Mass, Mass2, pexp=KB_PEXP_LAMBDA_NLO, device=0):
    return kb_params(int(Nx), int(sdim), int(NstepsinMemory), int(ONgroup), float(dt), float(Lambda),
                     float(Mass), float(Mass2), int(pexp), int(device))


def _free_chk(rc):
    """Raise for a failed context-free call (message: kb_last_error(NULL))."""
    if rc != KB_OK:
        raise KBError(rc, load_library().kb_last_error(None).decode())


def sunset_workspace_bytes(Nx, sdim, nrows):
    """Device workspace bytes kb_sunset needs for nrows lattices."""
    L = load_library()
    n = ctypes.c_ulonglong()
    p = make_params(Nx, sdim, 3, 0.1, 1.0, 1, 1.0, 1.0)
    _free_chk(L.kb_sunset_workspace_bytes(ctypes.byref(p), int(nrows), ctypes.byref(n)))
    return n.value


def sunset(Nx, sdim, Lambda, ONgroup, nrows, F, rho, SigF, Sigr, workspace, workspace_bytes, stream=None,
           device=0, classical=False):
    """calcSelfEnergyattp! (TwoPIScalar_Evolution.jl:25-31) on nrows device
    lattices: F, rho, SigF, Sigr and workspace are device addresses (ints),
    rows of Nx^sdim doubles.  Enqueued on `stream` (a hipStream_t address,
    None = the null stream); does not wait."""
    L = load_library()
    p = make_params(Nx, sdim, 3, 0.1, Lambda, ONgroup, 1.0, 1.0, device=device,
                    pexp=KB_PEXP_LAMBDA_NLO_CLASSICAL if classical else KB_PEXP_LAMBDA_NLO)
    _free_chk(L.kb_sunset(ctypes.byref(p), int(nrows), F, rho, SigF, Sigr, workspace, int(workspace_bytes),
                          stream))


def sunset_noalloc(Nx, sdim, Lambda, ONgroup, nrows, F, rho, SigF, Sigr, stream=None, device=0):
    """sunset() with the library's 64 MiB per-device scratch as workspace: no
    allocation; calls on one device must not overlap."""
    L = load_library()
    p = make_params(Nx, sdim, 3, 0.1, Lambda, ONgroup, 1.0, 1.0, device=device)
    _free_chk(L.kb_sunset_noalloc(ctypes.byref(p), int(nrows), F, rho, SigF, Sigr, stream))


class KBContext:
    """One device-resident KB simulation (a kb_ctx)."""

    def __init__(self, Nx, sdim, NstepsinMemory, dt, Lambda, ONgroup, Mass, Mass2,
                 pexp=KB_PEXP_LAMBDA_NLO, device=0, tiles="auto", symmetric=False, _handle=None, _owner=None):
        self._L = load_library()
        p = make_params(Nx, sdim, NstepsinMemory, dt, Lambda, ONgroup, Mass, Mass2, pexp, device)
        if _handle is None:
            h = ctypes.c_void_p()
            create = self._L.kb_create_sym if symmetric else self._L.kb_create
            rc = create(ctypes.byref(p), ctypes.byref(h))
            if rc != KB_OK:
                raise KBError(rc, "kb_create failed (invalid parameters, no HIP device, or out of memory)")
        else:
            h = _handle
        self._h = h
        self._owner = _owner  # ensemble members are destroyed with their ensemble
        self.params = p
        self.V = int(Nx) ** int(sdim)
        self.symmetric = bool(symmetric)
        # orbit storage, LambdaNLO: created on the orbit product
        self.orbit_product = self.symmetric and self.fft_tiles() == FFT_ORBIT_PRODUCT
        if _owner is None and tiles is not None:
            if tiles == "auto" and symmetric:
                v, f = committed_tiles(Nx, sdim, NstepsinMemory, symmetric=True)
            elif tiles == "auto":
                v, f = committed_tiles(Nx, sdim, NstepsinMemory)
            else:
                v, f = int(tiles), None
            if v is not None:
                self.set_tiles(v)
            # an orbit-storage context on the orbit product keeps it: the stored
            # choice is for the plane kernels of full storage
            if f is not None and not self.orbit_product:
                self.set_fft_tiles(f)

    def close(self):
        if getattr(self, "_h", None):
            if self._owner is None:
                self._L.kb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != KB_OK:
            raise KBError(rc, self._L.kb_last_error(self._h).decode())

    # ---- set-up -------------------------------------------------------
    def storage_info(self):
        """(stored sites, lattice sites): orbits in cubic-orbit storage."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self._L.kb_storage_info(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def device_bytes(self):
        V, b = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self._L.kb_info(self._h, ctypes.byref(V), ctypes.byref(b)))
        return b.value

    def kL2(self):
        out = np.empty(self.V)
        self._chk(self._L.kb_get_kL2(self._h, _ptr(out)))
        return out

    def set_kL2(self, k2):
        self._chk(self._L.kb_set_kL2(self._h, _ptr(_f64(k2, self.V))))

    def init_particles(self, n, initmass):
        self._chk(self._L.kb_init_particles(self._h, _ptr(_f64(n, self.V)), float(initmass)))

    def init_rows(self, F11, F21, F22):
        a = [_f64(x, self.V) for x in (F11, F21, F22)]
        self._chk(self._L.kb_init_rows(self._h, *[_ptr(x) for x in a]))

    def save_window(self):
        """Snapshot of the run (window, self-energy ring, kL2, Hartree mass,
        window indices) as a uint8 array, for load_window of a context with
        the same parameters: checkpoint and restart."""
        n = ctypes.c_ulonglong()
        self._chk(self._L.kb_window_bytes(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        self._chk(self._L.kb_save_window(self._h, buf.ctypes.data_as(ctypes.c_void_p), n.value))
        return buf

    def load_window(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._chk(self._L.kb_load_window(self._h, buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes))

    def fill_free_window(self, tmax, n=None, mass=1.0):
        nn = None if n is None else _f64(n, self.V)
        self._chk(self._L.kb_fill_free_window(self._h, int(tmax), _ptr(nn), float(mass)))

    # ---- stepping -----------------------------------------------------
    def step(self, t):
        self._chk(self._L.kb_step(self._h, int(t)))

    def window(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._chk(self._L.kb_window(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def hartreemass2(self):
        m = ctypes.c_double()
        self._chk(self._L.kb_hartree(self._h, ctypes.byref(m)))
        return m.value

    def synchronize(self):
        self._chk(self._L.kb_synchronize(self._h))

    def use_graph(self, on=True):
        self._chk(self._L.kb_use_graph(self._h, 1 if on else 0))

    def set_stream(self, handle):
        self._chk(self._L.kb_set_stream(self._h, ctypes.c_void_p(handle) if handle else None))

    # ---- read-out -----------------------------------------------------
    def measure(self, t, modes=True):
        m = kb_meas()
        arrs = {k: np.empty(self.V) for k in ("F", "ddF", "r", "n", "omega", "c")} if modes else {}
        g = arrs.get
        self._chk(self._L.kb_measure(self._h, int(t), ctypes.byref(m), _ptr(g("F")), _ptr(g("ddF")),
                                     _ptr(g("r")), _ptr(g("n")), _ptr(g("omega")), _ptr(g("c"))))
        out = dict(time=m.time, E=m.E, E_pot=m.E_pot, E_kin=m.E_kin, E_mem=m.E_mem, M2=m.M2,
                   n_unstable=m.n_unstable)
        out.update(arrs)
        return out

    def measure_history(self, t):
        tmin, tmax = self.window()
        L = t - tmin
        bufs = [np.empty(max(L, 0)) for _ in range(4)]
        self._chk(self._L.kb_measure_history(self._h, int(t), *[_ptr(b) for b in bufs]))
        return dict(SigF=bufs[0], Sigr=bufs[1], Fonet=bufs[2], ronet=bufs[3])

    def get(self, which, i, j):
        out = np.empty(self.V)
        self._chk(self._L.kb_get_entry(self._h, 1 if which == "r" else 0, int(i), int(j), _ptr(out)))
        return out

    def set(self, which, i, j, arr):
        self._chk(self._L.kb_set_entry(self._h, 1 if which == "r" else 0, int(i), int(j),
                                       _ptr(_f64(arr, self.V))))

    def sigma(self, which, tp):
        out = np.empty(self.V)
        self._chk(self._L.kb_get_sigma(self._h, 1 if which == "r" else 0, int(tp), _ptr(out)))
        return out

    _NINV = {"F2kr2": 0, "Fr": 1, "Ir": 2, "IF_row": 3, "Ir_row": 4}

    def ninv(self, which, i, j):
        """NinverseNLO internals in momentum space: 'F2kr2' / 'Fr' bubbles of
        the pair (i, j), 'Ir' resummed kernel Ir(i, j) (i > j), 'IF_row' /
        'Ir_row' the last step's IF(t-1, j) / Ir(t-1, j) (i ignored)."""
        out = np.empty(self.V)
        self._chk(self._L.kb_ninv_get(self._h, self._NINV[which], int(i), int(j), _ptr(out)))
        return out

    def check_finite(self):
        self._chk(self._L.kb_check_finite(self._h))

    # ---- timing -------------------------------------------------------
    def enable_timing(self, on=True):
        self._chk(self._L.kb_enable_timing(self._h, 1 if on else 0))

    def timing(self, stage):
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        self._chk(self._L.kb_timing(self._h, int(stage), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def reset_timing(self):
        self._chk(self._L.kb_reset_timing(self._h))

    # ---- tile shapes --------------------------------------------------
    def tile_shapes(self):
        """[(rows per work item, rows per batch, threads)] of the memory-integral kernel."""
        out = []
        r, b, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        while self._L.kb_tile_shape(len(out), ctypes.byref(r), ctypes.byref(b), ctypes.byref(n)) == 0:
            out.append((r.value, b.value, n.value))
        return out

    def set_tiles(self, variant):
        self._chk(self._L.kb_set_tiles(self._h, int(variant)))

    def tiles(self):
        v = ctypes.c_int()
        self._chk(self._L.kb_get_tiles(self._h, ctypes.byref(v)))
        return v.value

    def autotune(self, reps=3):
        """Re-time the last step's memory integral in every tile shape and keep
        the fastest.  Returns {"variant", "rows", "batch", "ms": [...]}."""
        shapes = self.tile_shapes()
        ms = np.zeros(len(shapes))
        best = ctypes.c_int()
        self._chk(self._L.kb_autotune(self._h, int(reps), ctypes.byref(best), _ptr(ms)))
        r, b, n = shapes[best.value]
        return {"variant": best.value, "rows": r, "batch": b, "threads": n, "ms": ms.tolist()}

    def fft_shapes(self):
        """[threads per workgroup] of the self-energy plane kernels."""
        out = []
        n = ctypes.c_int()
        while self._L.kb_fft_shape(len(out), ctypes.byref(n)) == 0:
            out.append(n.value)
        return out

    def set_fft_tiles(self, variant):
        """A plane-kernel block size (index into fft_shapes()), or
        FFT_ORBIT_PRODUCT in an orbit-storage LambdaNLO context."""
        self._chk(self._L.kb_set_fft_tiles(self._h, int(variant)))

    def fft_tiles(self):
        v = ctypes.c_int()
        self._chk(self._L.kb_get_fft_tiles(self._h, ctypes.byref(v)))
        return v.value

    def autotune_fft(self, reps=3):
        """Re-time the last step's self-energy stage for every plane-kernel
        block size (and, in orbit storage, the orbit product, timed last) and
        keep the fastest.  Returns {"variant", "threads", "ms"}; threads is
        None when the orbit product wins."""
        shapes = self.fft_shapes()
        ms = np.zeros(len(shapes) + 1)
        best = ctypes.c_int()
        self._chk(self._L.kb_autotune_fft(self._h, int(reps), ctypes.byref(best), _ptr(ms)))
        if not self.orbit_product:
            ms = ms[:len(shapes)]
        b = best.value
        return {"variant": b, "threads": None if b == FFT_ORBIT_PRODUCT else shapes[b], "ms": ms.tolist()}

    # ---- adjoint ------------------------------------------------------
    def adj_record(self, checkpoint_every=0, max_steps=None, host_windows=None):
        """Record the coming kb_step calls for the reverse sweep (call right
        after initialisation).  checkpoint_every = 0 keeps a tape of every
        dropped column; C > 0 keeps the window every C steps and recomputes;
        -1 keeps the tape packed to one value per cubic orbit (the state must
        be cubic-symmetric).  With host_windows (C > 0), at most that many
        windows are kept, in pinned host memory, under the binomial
        checkpointing schedule (kb_adj_record_host)."""
        if max_steps is None:
            raise ValueError("max_steps is required")
        if host_windows is None:
            self._chk(self._L.kb_adj_record(self._h, int(checkpoint_every), int(max_steps)))
        else:
            self._chk(self._L.kb_adj_record_host(self._h, int(checkpoint_every), int(max_steps),
                                                 int(host_windows)))

    def adj_gradient(self, wgt, n0=True, F0=False):
        """Reverse sweep for L = sum wgt * n(tmax).  Returns a dict with
        'lambda' (dL/dLambda) and, on request, 'n0' (dL/dn(p,0) per site) and
        'F0' (3 x V: dL/dF(1,1), dL/dF(2,1), dL/dF(2,2))."""
        w = _f64(wgt, self.V)
        lam = ctypes.c_double()
        dn = np.empty(self.V) if n0 else None
        dF = np.empty(3 * self.V) if F0 else None
        self._chk(self._L.kb_adj_gradient(self._h, _ptr(w), ctypes.byref(lam), _ptr(dn), _ptr(dF)))
        out = {"lambda": lam.value}
        if n0:
            out["n0"] = dn
        if F0:
            out["F0"] = dF.reshape(3, self.V)
        return out

    def adj_info(self):
        b, r = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self._L.kb_adj_info(self._h, ctypes.byref(b), ctypes.byref(r)))
        return {"device_bytes": b.value, "recomputed_steps": r.value}

    def adj_stop(self):
        self._chk(self._L.kb_adj_stop(self._h))

    def adj_debug_stop(self, t_stop):
        self._chk(self._L.kb_adj_debug_stop(self._h, int(t_stop)))

    def adj_get(self, which, i, j):
        """Adjoint of F(i,j) ('F') or rho(i,j) ('r') after a stopped sweep."""
        w = {"F": 0, "r": 1, 0: 0, 1: 1}[which]
        out = np.empty(self.V)
        self._chk(self._L.kb_adj_get_entry(self._h, w, int(i), int(j), _ptr(out)))
        return out


class KBEnsemble:
    """Several KB runs advanced together (a kb_ens).  members[i] is a
    KBContext view of run i for initialisation and read-out.  A run whose
    dict has symmetric=True is stored in cubic orbits (as KBContext's
    symmetric=True)."""

    def __init__(self, runs, device=0):
        self._L = load_library()
        runs = [dict(r) for r in runs]
        sym = [bool(r.pop("symmetric", False)) for r in runs]
        arr = (kb_params * len(runs))(*[make_params(**dict(r, device=device)) for r in runs])
        h = ctypes.c_void_p()
        flags = (ctypes.c_int * len(runs))(*[1 if x else 0 for x in sym])
        rc = self._L.kb_ens_create_sym(arr, flags, len(runs), ctypes.byref(h))
        if rc != KB_OK:
            raise KBError(rc, "kb_ens_create_sym failed (invalid parameters, no HIP device, or out of memory)")
        self._h = h
        self.members = []
        for i, r in enumerate(runs):
            m = ctypes.c_void_p()
            self._chk(self._L.kb_ens_member(self._h, i, ctypes.byref(m)))
            member = KBContext(_handle=m, _owner=self, device=device, symmetric=sym[i], **r)
            # members share the default memory-integral shape (one launch over
            # all of them); the self-energy stage is per member
            _, f = committed_tiles(r["Nx"], r["sdim"], r["NstepsinMemory"], symmetric=sym[i])
            if f is not None and not member.orbit_product:
                member.set_fft_tiles(f)
            self.members.append(member)

    def _chk(self, rc):
        if rc != KB_OK:
            raise KBError(rc, self._L.kb_ens_last_error(self._h).decode())

    def step(self, t):
        self._chk(self._L.kb_ens_step(self._h, int(t)))

    def use_graph(self, on=True):
        self._chk(self._L.kb_ens_use_graph(self._h, 1 if on else 0))

    def set_tiles(self, variant):
        """Shape of the pooled memory-integral launch (a tile_shape variant)."""
        self._chk(self._L.kb_ens_set_tiles(self._h, int(variant)))

    def tiles(self):
        v = ctypes.c_int()
        self._chk(self._L.kb_ens_get_tiles(self._h, ctypes.byref(v)))
        return v.value

    def autotune(self, reps=3):
        """Re-time the last ensemble step's pooled memory integral in every
        tile shape and keep the fastest.  Returns {"variant", "rows", "batch",
        "threads", "ms": [...]} like KBContext.autotune."""
        shapes = self.members[0].tile_shapes()
        ms = np.zeros(len(shapes))
        best = ctypes.c_int()
        self._chk(self._L.kb_ens_autotune(self._h, int(reps), ctypes.byref(best), _ptr(ms)))
        r, b, n = shapes[best.value]
        return {"variant": best.value, "rows": r, "batch": b, "threads": n, "ms": ms.tolist()}

    def synchronize(self):
        self._chk(self._L.kb_ens_synchronize(self._h))

    def enable_timing(self, on=True):
        self._chk(self._L.kb_ens_enable_timing(self._h, 1 if on else 0))

    def timing(self):
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        self._chk(self._L.kb_ens_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            self._L.kb_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
