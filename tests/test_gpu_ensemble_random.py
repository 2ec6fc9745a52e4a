"""Ensembles of seeded random members against one oracle run per member.

Each ensemble draws 2-7 members with their own lattice (1D-3D, ragged site
blocks, pencil-pass sizes), window, O(N) group, coupling, time step, masses
(Mass2 != Mass^2), expansion (all five) and storage (cubic orbits for some
3D members), and steps eagerly or graph-replayed.  Every member's window
and measurements are compared with its oracle at 1e-10 (norm-wise relative
per lattice).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10
SEED = 20261205
NENS = 8


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def draw_member(rng):
    sdim = int(rng.choice([1, 2, 3, 3]))
    Nx = int(rng.integers(1, {1: 70, 2: 16, 3: 10}[sdim] + 1))
    pexp = int(rng.choice([0, 1, 1, 2, 3, 4]))
    if pexp == 3 and Nx > 64:
        Nx = 64  # the 1/N stage runs on the plane kernels in full storage
    Mass = float(rng.uniform(0.7, 1.3))
    return dict(Nx=Nx, sdim=sdim, NstepsinMemory=int(rng.integers(3, 10 if pexp == 3 else 14)),
                dt=float(rng.choice([0.05, 0.1])), Lambda=float(rng.uniform(0.3, 2.5)),
                ONgroup=int(rng.integers(1, 5)), Mass=Mass, Mass2=Mass * Mass * float(rng.uniform(0.7, 1.3)),
                pexp=pexp, symmetric=bool(sdim == 3 and rng.uniform() < 0.4))


def draw_ensembles():
    rng = np.random.default_rng(SEED)
    out = []
    for _ in range(NENS):
        runs = [draw_member(rng) for _ in range(int(rng.integers(2, 8)))]
        out.append(dict(runs=runs, graph=bool(rng.uniform() < 0.5), T=int(rng.integers(8, 20)),
                        Tocc=[float(rng.uniform(0.8, 3.0)) for _ in runs]))
    return out


ENSEMBLES = draw_ensembles()


@pytest.mark.parametrize("k", range(NENS))
def test_random_ensemble_matches_oracles(k):
    from qftkb import KBEnsemble
    e = ENSEMBLES[k]
    print(f"\n[case] random ensemble {k}: graph={e['graph']} T={e['T']} members "
          f"{[(r['Nx'], r['sdim'], r['NstepsinMemory'], r['pexp'], r['symmetric']) for r in e['runs']]}",
          flush=True)
    ens = KBEnsemble(e["runs"])
    ens.use_graph(e["graph"])
    oracles = []
    for r, m, Tocc in zip(e["runs"], ens.members, e["Tocc"]):
        o = Oracle(r["Nx"], r["sdim"], r["NstepsinMemory"], r["dt"], r["Lambda"], r["ONgroup"], r["Mass"],
                   r["Mass2"], pexp=r["pexp"], nthreads=4)
        k2 = o.kL2()
        n = 1.0 / (np.exp(np.sqrt(k2 + r["Mass"] ** 2) / (Tocc * r["Mass"])) - 1.0)
        o.init_particles(n, r["Mass"])
        m.init_particles(n, r["Mass"])
        oracles.append(o)
    T = e["T"]
    for t in range(3, T + 1):
        ens.step(t)
        for o in oracles:
            o.evolve(t)
    for i, (o, m) in enumerate(zip(oracles, ens.members)):
        assert m.window() == o.window(), i
        tmin, tmax = o.window()
        for a in range(tmin, tmax + 1):
            for b in range(tmin, a + 1):
                for w in "Fr":
                    assert rel(m.get(w, a, b), o.get(w, a, b)) <= RTOL, (i, w, a, b)
        mo, mg = o.measure(T), m.measure(T)
        assert abs(mg["E"] - mo["E"]) <= RTOL * abs(mo["E"]), i
        assert abs(mg["M2"] - mo["M2"]) <= RTOL * abs(mo["M2"]), i
        assert rel(mg["n"], mo["n"]) <= RTOL, i
        o.close()
    ens.close()
