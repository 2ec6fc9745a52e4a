"""hipGraph-replayed steps and batched ensembles, through the C ABI, against
the oracle (tolerance 1e-10, norm-wise relative per lattice)."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


def check_window(o, g, tol=RTOL):
    assert o.window() == g.window()
    tmin, tmax = o.window()
    for a in range(tmin, tmax + 1):
        for b in range(tmin, a + 1):
            for w in "Fr":
                assert rel(g.get(w, a, b), o.get(w, a, b)) <= tol, (w, a, b)


def test_graph_replay_matches_oracle():
    """Steps replayed from one captured graph, with eager measurement and
    synchronisation between replays, follow the oracle step by step."""
    from qftkb import KBContext
    print("\n[case] graph replay 6^3 N_mem=5", flush=True)
    args = (6, 3, 5, 0.1, 1.2, 1, 1.0, 1.0)
    o = Oracle(*args)
    g = KBContext(*args)
    n = thermal(o.kL2())
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    g.use_graph(True)
    for t in range(3, 30):
        o.evolve(t)
        g.step(t)
        mo, mg = o.measure(t), g.measure(t)
        assert abs(mg["E"] - mo["E"]) <= RTOL * abs(mo["E"])
        assert rel(mg["n"], mo["n"]) <= RTOL
        g.synchronize()
    check_window(o, g)


def test_graph_and_eager_agree_at_32cubed():
    """At the benchmark lattice: the same 6 steps eager and from a graph agree
    to rounding (the memory integral's accumulation order is not fixed)."""
    from qftkb import KBContext
    print("\n[case] graph vs eager 32^3 N_mem=16", flush=True)
    ctxs = []
    for graph in (False, True):
        c = KBContext(32, 3, 16, 0.1, 1.0, 1, 1.0, 1.0)
        c.fill_free_window(16, thermal(c.kL2()), 1.0)
        c.use_graph(graph)
        for t in range(17, 23):
            c.step(t)
            c.measure(t, modes=False)
        ctxs.append(c)
    a, b = ctxs
    for tp in range(7, 23):
        for w in "Fr":
            assert rel(b.get(w, 22, tp), a.get(w, 22, tp)) <= 1e-14


RUNS = [  # uneven members: lattice, window, coupling, O(N), dt
    dict(Nx=4, sdim=3, NstepsinMemory=5, dt=0.1, Lambda=0.7, ONgroup=1, Mass=1.0, Mass2=1.0),
    dict(Nx=6, sdim=3, NstepsinMemory=9, dt=0.1, Lambda=1.6, ONgroup=1, Mass=1.0, Mass2=1.0),
    dict(Nx=8, sdim=3, NstepsinMemory=6, dt=0.05, Lambda=1.1, ONgroup=4, Mass=0.8, Mass2=0.64),
    dict(Nx=5, sdim=2, NstepsinMemory=7, dt=0.1, Lambda=2.0, ONgroup=2, Mass=1.0, Mass2=1.0),
    dict(Nx=1, sdim=3, NstepsinMemory=4, dt=0.1, Lambda=1.0, ONgroup=1, Mass=1.0, Mass2=1.0),
    # past the plane kernels: pencil passes, one of them a prime length
    dict(Nx=72, sdim=2, NstepsinMemory=5, dt=0.1, Lambda=1.2, ONgroup=1, Mass=1.0, Mass2=1.0),
    dict(Nx=67, sdim=1, NstepsinMemory=6, dt=0.1, Lambda=1.4, ONgroup=2, Mass=1.0, Mass2=1.0),
]


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_members_match_oracle(graph):
    from qftkb import KBEnsemble
    print(f"\n[case] ensemble of {len(RUNS)} uneven runs, graph={graph}", flush=True)
    ens = KBEnsemble(RUNS)
    ens.use_graph(graph)
    oracles = []
    for r, m in zip(RUNS, ens.members):
        o = Oracle(r["Nx"], r["sdim"], r["NstepsinMemory"], r["dt"], r["Lambda"], r["ONgroup"], r["Mass"],
                   r["Mass2"])
        n = thermal(o.kL2(), m=r["Mass"])
        o.init_particles(n, r["Mass"])
        m.init_particles(n, r["Mass"])
        oracles.append(o)
    for t in range(3, 26):
        ens.step(t)
        for o in oracles:
            o.evolve(t)
        if t % 5 == 0:
            for o, m in zip(oracles, ens.members):
                mo, mg = o.measure(t), m.measure(t)
                assert abs(mg["E"] - mo["E"]) <= RTOL * abs(mo["E"])
                assert rel(mg["n"], mo["n"]) <= RTOL
    for o, m in zip(oracles, ens.members):
        check_window(o, m)
    # members cannot be stepped on their own
    from qftkb import KBError
    with pytest.raises(KBError):
        ens.members[0].step(26)
    ens.close()


SYM_RUNS = [  # orbit-storage members beside a full-storage one; one 1/N member
    dict(Nx=6, sdim=3, NstepsinMemory=9, dt=0.1, Lambda=1.6, ONgroup=1, Mass=1.0, Mass2=1.0, symmetric=True),
    dict(Nx=8, sdim=3, NstepsinMemory=6, dt=0.05, Lambda=1.1, ONgroup=4, Mass=0.8, Mass2=0.64),
    dict(Nx=7, sdim=2, NstepsinMemory=7, dt=0.1, Lambda=2.0, ONgroup=2, Mass=1.0, Mass2=1.0, symmetric=True),
    dict(Nx=8, sdim=3, NstepsinMemory=9, dt=0.05, Lambda=2.0, ONgroup=4, Mass=1.0, Mass2=1.0, pexp=3,
         symmetric=True),
]


@pytest.mark.parametrize("graph", [False, True])
def test_orbit_storage_members(graph):
    """Members in cubic-orbit storage step with the others: the LambdaNLO
    members follow the oracle, and every member follows a standalone context
    of its own storage (whose memory integral runs in the split form: the
    same sums in another order)."""
    from qftkb import KBContext, KBEnsemble
    print(f"\n[case] ensemble with orbit-storage members, graph={graph}", flush=True)
    ens = KBEnsemble(SYM_RUNS)
    ens.use_graph(graph)
    alone, oracles = [], []
    for r, m in zip(SYM_RUNS, ens.members):
        assert m.symmetric == r.get("symmetric", False)
        c = KBContext(**r)
        n = thermal(c.kL2(), m=r["Mass"])
        m.init_particles(n, r["Mass"])
        c.init_particles(n, r["Mass"])
        alone.append(c)
        o = None
        if r.get("pexp", 2) != 3:
            o = Oracle(r["Nx"], r["sdim"], r["NstepsinMemory"], r["dt"], r["Lambda"], r["ONgroup"], r["Mass"],
                       r["Mass2"])
            o.init_particles(n, r["Mass"])
        oracles.append(o)
    for t in range(3, 24):
        ens.step(t)
        for c, o in zip(alone, oracles):
            c.step(t)
            if o is not None:
                o.evolve(t)
        if t % 5 == 0:
            for c, o, m in zip(alone, oracles, ens.members):
                mg, mc = m.measure(t), c.measure(t)
                assert abs(mg["E"] - mc["E"]) <= 1e-12 * abs(mc["E"])
                if o is not None:
                    mo = o.measure(t)
                    assert abs(mg["E"] - mo["E"]) <= RTOL * abs(mo["E"])
                    assert rel(mg["n"], mo["n"]) <= RTOL
    for c, o, m in zip(alone, oracles, ens.members):
        check_window(c, m, tol=1e-12)
        if o is not None:
            check_window(o, m)
        c.close()
    ens.close()


def test_orbit_storage_ensemble_at_32cubed():
    """Two 32^3 orbit-storage runs and a 24^3 one at N_mem = 40 from filled
    windows: each member matches its standalone context."""
    from qftkb import KBContext, KBEnsemble, lattice
    print("\n[case] orbit-storage ensemble 32^3, 32^3, 24^3 N_mem=40", flush=True)
    runs = [dict(Nx=nx, sdim=3, NstepsinMemory=40, dt=0.1, Lambda=lam, ONgroup=1, Mass=1.0, Mass2=1.0,
                 symmetric=True) for nx, lam in ((32, 1.0), (32, 1.7), (24, 0.6))]
    ens = KBEnsemble(runs)
    ens.use_graph(True)
    alone = []
    for r, m, (n0, Q) in zip(runs, ens.members, ((2.0, 1.0), (3.0, 0.7), (1.5, 1.3))):
        c = KBContext(**r)
        for x in (m, c):
            x.fill_free_window(40, lattice.n_box(x.kL2(), n0, Q), 1.0)
        alone.append(c)
    for t in range(41, 49):
        ens.step(t)
        for c in alone:
            c.step(t)
    for c, m in zip(alone, ens.members):
        for tp in range(9, 49):
            for w in "Fr":
                assert rel(m.get(w, 48, tp), c.get(w, 48, tp)) <= 1e-12, (w, tp)
        mg, mc = m.measure(48), c.measure(48)
        assert abs(mg["E"] - mc["E"]) <= 1e-12 * abs(mc["E"])
        c.close()
    ens.close()


def test_ensemble_tiles_and_autotune():
    """The pooled memory integral in every tile shape (members' item tables
    follow the shape), and the autotuner, keep the members on the oracle."""
    from qftkb import KBEnsemble, KBError
    print("\n[case] ensemble tile shapes + autotune", flush=True)
    ens = KBEnsemble(RUNS[:3] + SYM_RUNS[:1])
    oracles = []
    for r, m in zip(RUNS[:3] + SYM_RUNS[:1], ens.members):
        o = Oracle(r["Nx"], r["sdim"], r["NstepsinMemory"], r["dt"], r["Lambda"], r["ONgroup"], r["Mass"],
                   r["Mass2"])
        n = thermal(o.kL2(), m=r["Mass"])
        o.init_particles(n, r["Mass"])
        m.init_particles(n, r["Mass"])
        oracles.append(o)
    with pytest.raises(KBError):
        ens.autotune(1)  # no step yet
    nshapes = len(ens.members[0].tile_shapes())
    t = 3
    for v in list(range(nshapes)) + [None]:
        if v is None:
            best = ens.autotune(2)
            assert ens.tiles() == best["variant"] and len(best["ms"]) == nshapes
            assert all(x > 0 for x in best["ms"])
        else:
            ens.set_tiles(v)
            assert ens.tiles() == v
        ens.use_graph(v is None or v % 2 == 1)
        for _ in range(3):
            ens.step(t)
            for o in oracles:
                o.evolve(t)
            t += 1
    for o, m in zip(oracles, ens.members):
        check_window(o, m)
        mo, mg = o.measure(t - 1), m.measure(t - 1)
        assert abs(mg["E"] - mo["E"]) <= RTOL * abs(mo["E"])
    ens.close()


@pytest.mark.parametrize("graph", [False, True])
def test_all_orbit_ensemble_with_1n_members(graph):
    """Every member in orbit storage (each lane runs its own memory integral
    after its stages), two of them in the 1/N expansion (split resummation
    sweeps on a side stream): each member follows a standalone context."""
    from qftkb import KBContext, KBEnsemble
    print(f"\n[case] all-orbit ensemble with 1/N members, graph={graph}", flush=True)
    runs = [
        dict(Nx=8, sdim=3, NstepsinMemory=9, dt=0.05, Lambda=2.0, ONgroup=4, Mass=1.0, Mass2=1.0, pexp=3,
             symmetric=True),
        dict(Nx=6, sdim=3, NstepsinMemory=9, dt=0.1, Lambda=1.6, ONgroup=1, Mass=1.0, Mass2=1.0, symmetric=True),
        dict(Nx=9, sdim=2, NstepsinMemory=7, dt=0.05, Lambda=1.5, ONgroup=3, Mass=1.0, Mass2=1.0, pexp=3,
             symmetric=True),
        dict(Nx=10, sdim=3, NstepsinMemory=5, dt=0.1, Lambda=0.9, ONgroup=1, Mass=1.0, Mass2=1.0, symmetric=True),
        dict(Nx=4, sdim=3, NstepsinMemory=6, dt=0.1, Lambda=1.2, ONgroup=2, Mass=1.0, Mass2=1.0, symmetric=True),
    ]
    ens = KBEnsemble(runs)
    ens.use_graph(graph)
    alone = []
    for r, m in zip(runs, ens.members):
        c = KBContext(**r)
        n = thermal(c.kL2(), m=r["Mass"])
        m.init_particles(n, r["Mass"])
        c.init_particles(n, r["Mass"])
        alone.append(c)
    for t in range(3, 22):
        ens.step(t)
        for c in alone:
            c.step(t)
        if t % 6 == 0:
            for c, m in zip(alone, ens.members):
                mg, mc = m.measure(t), c.measure(t)
                assert abs(mg["E"] - mc["E"]) <= 1e-12 * abs(mc["E"])
    for c, m in zip(alone, ens.members):
        check_window(c, m, tol=1e-12)
        c.close()
    ens.close()
