"""Window snapshots at seeded random points of seeded random runs.

Each case runs a thermal state to a random time (the window filling or
sliding), saves the window (kb_save_window), loads it into a fresh context
with the same parameters (kb_load_window), and continues both.  The
restarted run must follow the uninterrupted one (1e-13; 1e-12 with 1/N rows,
which are rebuilt after a load: the same sums in another order) and the
oracle (1e-10, norm-wise relative per lattice).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

SEED = 20261206
NCASES = 16


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def draw_cases():
    rng = np.random.default_rng(SEED)
    cases = []
    for i in range(NCASES):
        sdim = int(rng.choice([1, 2, 3, 3]))
        Nx = int(rng.integers(2, {1: 40, 2: 14, 3: 10}[sdim] + 1))
        pexp = int((0, 1, 2, 3, 4)[i % 5])
        Nmem = int(rng.integers(3, 10 if pexp == 3 else 16))
        Mass = float(rng.uniform(0.7, 1.3))
        cases.append(dict(args=(Nx, sdim, Nmem, float(rng.choice([0.05, 0.1])), float(rng.uniform(0.3, 2.5)),
                                int(rng.integers(1, 5)), Mass, Mass * Mass * float(rng.uniform(0.7, 1.3))),
                          pexp=pexp, symmetric=bool(sdim == 3 and rng.uniform() < 0.5),
                          t_snap=int(rng.integers(2, Nmem + 12)), more=int(rng.integers(1, 10))))
    return cases


CASES = draw_cases()


def _id(c):
    a = c["args"]
    return f"{a[0]}^{a[1]}-N{a[2]}-p{c['pexp']}-snap{c['t_snap']}" + ("-orbits" if c["symmetric"] else "")


@pytest.mark.parametrize("c", CASES, ids=[_id(c) for c in CASES])
def test_random_restart_matches_run_and_oracle(c):
    from qftkb import KBContext
    print(f"\n[case] random restart {c}", flush=True)
    args, m = c["args"], c["args"][6]
    tol = 1e-12 if c["pexp"] == 3 else 1e-13
    o = Oracle(*args, pexp=c["pexp"], nthreads=4)
    k2 = o.kL2()
    n = 1.0 / (np.exp(np.sqrt(k2 + m * m) / (2.0 * m)) - 1.0)
    o.init_particles(n, m)
    a = KBContext(*args, pexp=c["pexp"], symmetric=c["symmetric"])
    a.init_particles(n, m)
    ts, te = c["t_snap"], c["t_snap"] + c["more"]
    for t in range(3, ts + 1):
        a.step(t)
    snap = a.save_window()
    b = KBContext(*args, pexp=c["pexp"], symmetric=c["symmetric"])
    b.load_window(snap)
    assert b.window() == a.window()
    for t in range(ts + 1, te + 1):
        a.step(t)
        b.step(t)
    for t in range(3, te + 1):
        o.evolve(t)
    assert a.window() == b.window() == o.window()
    tmin, tmax = o.window()
    for x in range(tmin, tmax + 1):
        for y in range(tmin, x + 1):
            for w in "Fr":
                ref = a.get(w, x, y)
                assert rel(b.get(w, x, y), ref) <= tol, (w, x, y)
                assert rel(b.get(w, x, y), o.get(w, x, y)) <= 1e-10, (w, x, y)
    ea, eb = a.measure(te)["E"], b.measure(te)["E"]
    assert abs(eb - ea) <= tol * abs(ea)
    a.close()
    b.close()
    o.close()
