"""Checkpoint and restart (kb_window_bytes / kb_save_window / kb_load_window):
a run restored from a snapshot continues exactly as the run it was taken
from, in a fresh context or in the same context rewound; snapshots of other
parameters or formats are refused."""
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


@pytest.mark.parametrize("kw,exact", [
    pytest.param({}, True, id="lambda-NLO"),
    pytest.param({"symmetric": True}, True, id="orbits"),
    pytest.param({"pexp": 3, "symmetric": True}, False, id="1/N-orbits"),
    pytest.param({"pexp": 3}, False, id="1/N"),
])
def test_restart_continues_the_run(kw, exact):
    from qftkb import KBContext
    args = (8, 3, 9, 0.05, 2.0, 4, 1.0, 1.0) if kw.get("pexp") == 3 else (8, 3, 9, 0.1, 1.0, 1, 1.0, 1.0)
    print(f"\n[case] window snapshot {kw}", flush=True)
    a = KBContext(*args, **kw)
    a.init_particles(thermal(a.kL2()), 1.0)
    for t in range(3, 21):  # past N_mem: the window slides
        a.step(t)
    snap = a.save_window()
    for t in range(21, 31):
        a.step(t)
    ref = window_rows(a)
    e_ref = a.measure(30)["E"]
    # a fresh context with the same parameters
    b = KBContext(*args, **kw)
    b.load_window(snap)
    assert b.window() == (20 - 9 + 1, 20)
    for t in range(21, 31):
        b.step(t)
    got = window_rows(b)
    # the same context rewound
    a.load_window(snap)
    for t in range(21, 31):
        a.step(t)
    again = window_rows(a)
    # the 1/N rows are rebuilt after a load: the same sums, possibly in another order
    tol = 0.0 if exact else 1e-12
    for k, v in ref.items():
        assert rel(got[k], v) <= tol, k
        assert rel(again[k], v) <= tol, k
    assert abs(b.measure(30)["E"] - e_ref) <= max(tol, 1e-15) * abs(e_ref)
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
    b.load_window(snap)
    b.step(10)
    a.step(10)
    assert np.array_equal(b.get("F", 10, 10), a.get("F", 10, 10))
    a.close()
    b.close()
