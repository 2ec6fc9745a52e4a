"""Checkpoint and restart (kb_window_bytes / kb_save_window / kb_load_window):
a run restored from a snapshot continues as the run it was taken from, in a
fresh context or in the same context rewound; snapshots of other parameters,
formats or windows are refused.

The memory integral adds its row and column sums in the order in which the
workgroup's waves take work items from a shared queue, so two runs of the
same step agree to rounding (1e-14), not bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


def window_rows(ctx):
    tmin, tmax = ctx.window()
    return {(w, tp): ctx.get(w, tmax, tp) for tp in range(tmin, tmax + 1) for w in "Fr"}


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / s if s > 0 else np.max(np.abs(a - b))


@pytest.mark.parametrize("kw,nmem,tol", [
    pytest.param({}, 9, 1e-14, id="lambda-NLO"),
    # a window long enough that each site block's work items span several
    # waves, so the accumulation order of the memory integral varies
    pytest.param({}, 72, 1e-14, id="lambda-NLO-Nmem72"),
    pytest.param({"symmetric": True}, 9, 1e-14, id="orbits"),
    pytest.param({"symmetric": True}, 72, 1e-14, id="orbits-Nmem72"),
    # the 1/N rows are rebuilt after a load: the same sums in another order
    pytest.param({"pexp": 3, "symmetric": True}, 9, 1e-12, id="1/N-orbits"),
    pytest.param({"pexp": 3}, 9, 1e-12, id="1/N"),
])
def test_restart_continues_the_run(kw, nmem, tol):
    from qftkb import KBContext
    if kw.get("pexp") == 3:
        args = (8, 3, nmem, 0.05, 2.0, 4, 1.0, 1.0)
    else:
        args = (8, 3, nmem, 0.1, 1.0, 1, 1.0, 1.0)
    t_snap = nmem + 11  # past N_mem: the window slides
    print(f"\n[case] window snapshot {kw} N_mem={nmem}", flush=True)
    a = KBContext(*args, **kw)
    a.init_particles(thermal(a.kL2()), 1.0)
    for t in range(3, t_snap + 1):
        a.step(t)
    snap = a.save_window()
    t_end = t_snap + 10
    for t in range(t_snap + 1, t_end + 1):
        a.step(t)
    ref = window_rows(a)
    e_ref = a.measure(t_end)["E"]
    # a fresh context with the same parameters
    b = KBContext(*args, **kw)
    b.load_window(snap)
    assert b.window() == (t_snap - nmem + 1, t_snap)
    for t in range(t_snap + 1, t_end + 1):
        b.step(t)
    got = window_rows(b)
    # the same context rewound
    a.load_window(snap)
    for t in range(t_snap + 1, t_end + 1):
        a.step(t)
    again = window_rows(a)
    for k, v in ref.items():
        assert rel(got[k], v) <= tol, k
        assert rel(again[k], v) <= tol, k
    assert abs(b.measure(t_end)["E"] - e_ref) <= tol * abs(e_ref)
    a.close()
    b.close()


def test_snapshots_of_other_runs_are_refused():
    from qftkb import KBContext, KBError
    args = (6, 3, 7, 0.1, 1.0, 1, 1.0, 1.0)
    a = KBContext(*args)
    with pytest.raises(KBError):
        a.save_window()  # not initialised
    a.init_particles(thermal(a.kL2()), 1.0)
    for t in range(3, 10):
        a.step(t)
    snap = a.save_window()
    for other in [(6, 3, 7, 0.1, 1.5, 1, 1.0, 1.0), (6, 3, 8, 0.1, 1.0, 1, 1.0, 1.0)]:
        c = KBContext(*other)
        with pytest.raises(KBError):
            c.load_window(snap)
        c.close()
    s = KBContext(*args, symmetric=True)
    with pytest.raises(KBError):
        s.load_window(snap)  # other storage (and size)
    s.close()
    b = KBContext(*args)
    with pytest.raises(KBError):
        b.load_window(snap[:-8])  # truncated
    bad = snap.copy()
    bad[0] ^= 0xFF
    with pytest.raises(KBError):
        b.load_window(bad)  # not this format
    # a header whose window is not the window of its last step
    from qftkb._abi import kb_params

    class Head(ctypes.Structure):
        _fields_ = [("magic", ctypes.c_char * 8), ("version", ctypes.c_int), ("sym", ctypes.c_int),
                    ("p", kb_params), ("V", ctypes.c_int), ("N", ctypes.c_int), ("NP", ctypes.c_longlong),
                    ("t", ctypes.c_int), ("tmin", ctypes.c_int), ("tmax", ctypes.c_int)]
    h = Head.from_buffer_copy(bytes(snap[:ctypes.sizeof(Head)]))
    assert (h.tmin, h.tmax) == a.window() == (3, 9)
    for field, value in (("tmin", 4), ("tmin", 2), ("tmax", 10), ("tmax", 1)):
        bad = snap.copy()
        off = getattr(Head, field).offset
        bad[off:off + 4] = np.frombuffer(np.int32(value).tobytes(), dtype=bad.dtype)
        with pytest.raises(KBError):
            b.load_window(bad)
    b.load_window(snap)
    b.step(10)
    a.step(10)
    assert rel(b.get("F", 10, 10), a.get("F", 10, 10)) <= 1e-14
    a.close()
    b.close()
