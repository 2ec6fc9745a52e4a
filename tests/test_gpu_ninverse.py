"""Parity of the 1/N expansion (NinverseLO, NinverseNLO) on the HIP path with
the oracle, through the C ABI.

The HIP path computes the resummed kernels Ir, IF in momentum space by three
sweeps over slot triangles (bubbles and a stored Ir history); the oracle
restates the reference's CPUfull code, which transforms every term of the
time sums (TwoPIScalar_Evolution.jl:170-206, 354-436, 683-698).  Tolerance as
for the lambda expansion: 1e-10 relative, norm-wise per lattice.
"""
import numpy as np
import pytest

from oracle.kb_numpy import NumpyKB
from oracle.oracle import Oracle

from test_gpu_parity import RTOL, compare_meas, compare_window, rel

pytestmark = pytest.mark.gpu

NINV_NLO = 3
NINV_LO = 2


def tophat(kL2, sdim):
    from qftkb import lattice
    return lattice.n_tophat(kL2, 1, sdim, 1.0)


def make_pair(Nx, sdim, Nmem, dt=0.05, lam=10.0, ON=8, pexp=NINV_NLO, symmetric=False):
    from qftkb import KBContext
    o = Oracle(Nx, sdim, Nmem, dt, lam, ON, 1.0, 1.0, pexp=pexp)
    g = KBContext(Nx, sdim, Nmem, dt, lam, ON, 1.0, 1.0, pexp=pexp, symmetric=symmetric)
    return o, g


def compare_sigma(o, g, tol=RTOL):
    tmin, tmax = o.window()
    scale = max(max(np.abs(o.sigma(w, tp)).max() for w in "Fr") for tp in range(tmin, tmax))
    for tp in range(tmin, tmax):
        for w in "Fr":
            d = np.abs(g.sigma(w, tp) - o.sigma(w, tp)).max()
            assert d <= tol * scale, f"Sigma{w}({tp}) {d / scale:.3e}"


CASES = [
    # Nx, sdim, Nmem, steps: full memory, sliding windows, 1D/2D/3D, odd sizes
    (4, 3, 6, 16),
    (8, 3, 9, 14),
    (5, 3, 20, 12),
    (6, 2, 5, 18),
    (7, 2, 12, 14),
    (16, 1, 4, 20),
    (1, 3, 6, 12),
]


# 1D lattices past 64 run on the plane kernels as one line: the reference's 1/N test lattice
# (TestTwoPIScalar_CPUred(512, 1, 4, "NinverseNLO", 50), test/runtests.jl), a prime and the largest
LARGE_1D = [(512, 1, 5, 10), (67, 1, 6, 12), (2048, 1, 4, 8)]
# 2D and 3D lattices past 64 run the bubble and chain passes as pencil passes (a prime among them)
PENCILS = [(72, 2, 5, 10), (67, 2, 5, 9), (130, 2, 4, 8), (66, 3, 3, 5)]


@pytest.mark.parametrize("Nx,sdim,Nmem,steps", CASES + [(32, 3, 5, 8), (40, 3, 4, 7), (48, 3, 4, 6), (64, 3, 3, 5)]
                         + LARGE_1D + PENCILS)
def test_ninverse_nlo_trajectory(Nx, sdim, Nmem, steps):
    print(f"\n[case] 1/N NLO trajectory Nx={Nx} sdim={sdim} Nmem={Nmem} steps={steps}", flush=True)
    o, g = make_pair(Nx, sdim, Nmem)
    n = tophat(o.kL2(), sdim)
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    for t in range(3, 3 + steps):
        o.evolve(t)
        g.step(t)
        if t % 4 == 0 or t == 2 + steps:
            compare_sigma(o, g)
            compare_meas(o, g, t)
    compare_window(o, g)
    g.check_finite()


def test_ninverse_internals_match_momentum_twin():
    """Bubbles, the stored Ir history and the last (IF, Ir) row against the
    numpy restatement of the reference's momentum-space (CPUred2) recursion."""
    Nx, sdim, Nmem = 4, 3, 7
    args = (Nx, sdim, Nmem, 0.05, 10.0, 8, 1.0, 1.0)
    nk = NumpyKB(*args, pexp=NINV_NLO)
    o, g = make_pair(Nx, sdim, Nmem)
    n = tophat(o.kL2(), sdim)
    nk.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    rows = {}
    for t in range(3, 15):
        nk.evolve(t)
        g.step(t)
        rows[t - 1] = dict(nk.Irk)
    T = 13
    tmin, tmax = g.window()
    for tp in range(tmin, T + 1):
        assert rel(g.ninv("IF_row", 0, tp), nk.IFk[tp]) <= RTOL, ("IF", tp)
        ir = g.ninv("Ir_row", 0, tp)
        if tp < T:
            assert rel(ir, nk.Irk[tp]) <= RTOL, ("Ir", tp)
        else:
            assert np.abs(ir).max() == 0.0
    nk._bub = {}
    for a in range(tmin, T + 1):
        for b in range(tmin, a + 1):
            PF, PR = nk.bubbles(a, b)
            assert rel(g.ninv("F2kr2", a, b), PF) <= RTOL, ("F2kr2", a, b)
            if a > b:
                assert rel(g.ninv("Fr", a, b), PR) <= RTOL, ("Fr", a, b)
                assert rel(g.ninv("Fr", b, a), -PR) <= RTOL
                assert rel(g.ninv("Ir", a, b), rows[a][b]) <= RTOL, ("Ir", a, b)


@pytest.mark.parametrize("Nx,sdim", [(4, 3), (72, 2)])  # plane kernels; pencil passes
def test_ninverse_window_written_from_outside(Nx, sdim):
    """Writes into the window (set_entry) and a free-propagator window
    (fill_free_window) invalidate the stored bubbles and Ir rows; the next
    step rebuilds them row by row."""
    print(f"\n[case] 1/N rows rebuilt after outside writes, {Nx}^{sdim}", flush=True)
    o, g = make_pair(Nx, sdim, 8)
    n = tophat(o.kL2(), 3)
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    for t in range(3, 9):
        o.evolve(t)
        g.step(t)
    # perturbations constant on momentum shells (createrandomF of the
    # reference's test): the lattices stay inversion-even, as every state of
    # the dynamics is; for odd parts the reference's fast and simple 1/N
    # variants differ from each other, and the HIP path keeps the even part
    from qftkb import lattice
    _, shells = lattice.levels(o.kL2())
    rng = np.random.default_rng(5)

    # the site of -p: rounding can put p and -p in different shells, so the
    # noise is made even explicitly
    ix = np.indices((Nx,) * sdim)[::-1]  # (x, y, z) index grids, x fastest in the flat order
    refl = np.ravel_multi_index(tuple(((Nx - c) % Nx) for c in ix[::-1]), (Nx,) * sdim).reshape(-1)

    def shell_noise():
        u = np.empty(o.V)
        for s in shells:
            u[s] = rng.random()
        u = 0.5 * (u + u[refl])
        return 1 + 0.1 * u

    tmin, tmax = o.window()
    for (a, b) in ((tmax - 1, tmin + 1), (tmax, tmax - 2)):
        x = o.get("F", a, b) * shell_noise()
        o.set("F", a, b, x)
        g.set("F", a, b, x)
        y = o.get("r", a, b) * shell_noise()
        o.set("r", a, b, y)
        g.set("r", a, b, y)
    for t in range(tmax + 1, tmax + 6):
        o.evolve(t)
        g.step(t)
    compare_sigma(o, g)
    compare_window(o, g)

    o2, g2 = make_pair(Nx, sdim, 10)
    o2.fill_free_window(10, n, 1.0)
    g2.fill_free_window(10, n, 1.0)
    for t in range(11, 16):
        o2.evolve(t)
        g2.step(t)
    compare_sigma(o2, g2)
    compare_window(o2, g2)


def test_ninverse_graph_replay_equals_eager():
    from qftkb import KBContext
    args = (6, 3, 8, 0.05, 10.0, 8, 1.0, 1.0)
    a = KBContext(*args, pexp=NINV_NLO)
    b = KBContext(*args, pexp=NINV_NLO)
    n = tophat(a.kL2(), 3)
    a.init_particles(n, 1.0)
    b.init_particles(n, 1.0)
    b.use_graph(True)
    for t in range(3, 20):
        a.step(t)
        b.step(t)
        if t % 5 == 0:
            b.synchronize()
            assert rel(b.get("F", t, t), a.get("F", t, t)) == 0.0
    tmin, tmax = a.window()
    for tp in range(tmin, tmax + 1):
        assert np.array_equal(a.get("F", tmax, tp), b.get("F", tmax, tp))
        assert np.array_equal(a.get("r", tmax, tp), b.get("r", tmax, tp))


def test_ninverse_lo_matches_oracle():
    o, g = make_pair(6, 3, 7, pexp=NINV_LO)
    n = tophat(o.kL2(), 3)
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    assert abs(g.hartreemass2() - o.hartreemass2) <= RTOL * abs(o.hartreemass2)
    for t in range(3, 18):
        o.evolve(t)
        g.step(t)
    compare_meas(o, g, 17)
    compare_window(o, g)


def test_ninverse_orbit_storage_equals_full_storage():
    _, full = make_pair(8, 3, 9)
    _, sym = make_pair(8, 3, 9, symmetric=True)
    n = tophat(full.kL2(), 3)
    full.init_particles(n, 1.0)
    sym.init_particles(n, 1.0)
    for t in range(3, 16):
        full.step(t)
        sym.step(t)
    tmin, tmax = full.window()
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(sym.get(w, tmax, tp), full.get(w, tmax, tp)) <= 1e-12
    # Sigma against the ring-wide scale: Sigr(T, T) is zero up to roundoff
    ref = {(w, tp): full.sigma(w, tp) for tp in range(tmin, tmax) for w in "Fr"}
    scale = max(np.abs(v).max() for v in ref.values())
    for (w, tp), b in ref.items():
        assert np.abs(sym.sigma(w, tp) - b).max() <= 1e-12 * scale, (w, tp)


@pytest.mark.parametrize("Nx,Nmem,fill", [(8, 9, False), (12, 10, True), (7, 6, False)])
def test_ninverse_orbit_product_equals_plane_kernels(Nx, Nmem, fill):
    """Orbit storage, NinverseNLO: the bubble and chain passes as products
    with the orbit cosine matrix (the default) and as plane FFTs over the
    expanded lattice give one trajectory to 1e-12; with a filled window the
    bubble rows are first rebuilt one row at a time through the same path.
    The product context is graph-replayed."""
    from qftkb import KBContext
    from qftkb._abi import FFT_ORBIT_PRODUCT
    print(f"\n[case] 1/N orbit product vs plane kernels Nx={Nx} Nmem={Nmem} fill={fill}", flush=True)
    args = (Nx, 3, Nmem, 0.05, 10.0, 8, 1.0, 1.0)
    a = KBContext(*args, pexp=NINV_NLO, symmetric=True)
    b = KBContext(*args, pexp=NINV_NLO, symmetric=True, tiles=None)
    assert a.orbit_product and a.fft_tiles() == FFT_ORBIT_PRODUCT
    b.set_fft_tiles(0)
    a.use_graph(True)
    n = tophat(a.kL2(), 3)
    if fill:
        a.fill_free_window(Nmem, n, 1.0)
        b.fill_free_window(Nmem, n, 1.0)
        t0 = Nmem + 1
    else:
        a.init_particles(n, 1.0)
        b.init_particles(n, 1.0)
        t0 = 3
    for t in range(t0, t0 + 2 * Nmem):
        a.step(t)
        b.step(t)
    tmin, tmax = b.window()
    assert a.window() == (tmin, tmax)
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(a.get(w, tmax, tp), b.get(w, tmax, tp)) <= 1e-12, (w, tp)
    ref = {(w, tp): b.sigma(w, tp) for tp in range(tmin, tmax) for w in "Fr"}
    scale = max(np.abs(v).max() for v in ref.values())
    for (w, tp), v in ref.items():
        assert np.abs(a.sigma(w, tp) - v).max() <= 1e-12 * scale, (w, tp)
    for which in ("F2kr2", "Fr", "Ir"):
        x, y = a.ninv(which, tmax - 1, tmin + 1), b.ninv(which, tmax - 1, tmin + 1)
        assert rel(x, y) <= 1e-12, which
    a.close()
    b.close()


def test_ninverse_unsupported_combinations_are_rejected():
    from qftkb import KBContext
    from qftkb._abi import KBError
    g = KBContext(4, 3, 5, 0.05, 1.0, 8, 1.0, 1.0, pexp=NINV_NLO)
    g.init_particles(np.zeros(64), 1.0)
    with pytest.raises(KBError):
        g.adj_record(max_steps=10)


def test_ninverse_parameter_dict_run_matches_oracle():
    """The reference's 1/N test problem (TwoPIScalar_Tests_NinverseNLO_
    checkfastimplementation.jl:29-66: Mass 1, dt 0.01, lambda 10, N 8,
    TopHatT1, RNone, full memory) run through the parameter-Dict driver; every
    measurement's energies and n(p) against the oracle."""
    from qftkb import TwoPIScalarfile, run
    print("\n[case] 1/N parameter-Dict run 6^3, 12 steps", flush=True)
    parameters = {"Mass": 1, "Nx": 6, "sdim": 3, "dt": 0.01, "Nsteps": 12, "NstepsinMemory": 0, "Nmeas": 0,
                  "Mod": "phi4", "Lambda": 10, "ONgroup": 8, "init": "TopHatT1", "xi": 0.4, "eta": 5.0, "sig": 1.0,
                  "T": 5, "n": 2, "Reno": "RNone", "Pexp": "NinverseNLO", "Num": "GPUfull", "Nchunks": 0,
                  "tag": "testing", "ov": True}
    sol = run(TwoPIScalarfile(), dict(parameters))
    N = sol.problem.disc.NstepsinMemory
    o = Oracle(6, 3, N, 0.01, 10.0, 8, 1.0, 1.0, pexp=NINV_NLO)
    o.init_particles(tophat(o.kL2(), 3), 1.0)
    ref = {2: o.measure(2)}
    for t in range(3, 15):
        o.evolve(t)
        ref[t] = o.measure(t)
    got = [m for m in sol.measurearray if m is not None]
    assert len(got) == len(ref)
    for m, t in zip(got, sorted(ref)):
        mo = ref[t]
        for k in ("E", "E_kin", "E_pot", "M2"):
            assert abs(getattr(m, k) - mo[k]) <= RTOL * abs(mo[k]), (t, k)
        assert abs(m.E_mem - mo["E_mem"]) <= RTOL * abs(mo["E"]), (t, "E_mem")
        assert rel(m.n, mo["n"]) <= RTOL, (t, "n")


def test_ninverse_in_mixed_ensemble():
    """An ensemble of runs in different expansions and sizes (1/N NLO, 1/N LO,
    lambda NLO), graph-replayed, each member against its own oracle run."""
    from qftkb import KBEnsemble
    runs = [dict(Nx=6, sdim=3, NstepsinMemory=8, dt=0.05, Lambda=10.0, ONgroup=8, Mass=1.0, Mass2=1.0, pexp=NINV_NLO),
            dict(Nx=4, sdim=3, NstepsinMemory=20, dt=0.05, Lambda=5.0, ONgroup=4, Mass=1.0, Mass2=1.0, pexp=NINV_NLO),
            dict(Nx=8, sdim=2, NstepsinMemory=6, dt=0.05, Lambda=10.0, ONgroup=8, Mass=1.0, Mass2=1.0, pexp=NINV_LO),
            dict(Nx=5, sdim=3, NstepsinMemory=7, dt=0.05, Lambda=1.0, ONgroup=1, Mass=1.0, Mass2=1.0, pexp=1),
            # 1/N on the pencil passes and on a 1D line past 64
            dict(Nx=72, sdim=2, NstepsinMemory=5, dt=0.05, Lambda=10.0, ONgroup=8, Mass=1.0, Mass2=1.0, pexp=NINV_NLO),
            dict(Nx=130, sdim=1, NstepsinMemory=6, dt=0.05, Lambda=10.0, ONgroup=8, Mass=1.0, Mass2=1.0,
                 pexp=NINV_NLO)]
    print(f"\n[case] mixed-expansion ensemble, {len(runs)} runs", flush=True)
    e = KBEnsemble(runs)
    e.use_graph(True)
    oracles = []
    for r, m in zip(runs, e.members):
        o = Oracle(r["Nx"], r["sdim"], r["NstepsinMemory"], r["dt"], r["Lambda"], r["ONgroup"], r["Mass"],
                   r["Mass2"], pexp=r["pexp"])
        n = tophat(o.kL2(), r["sdim"])
        o.init_particles(n, 1.0)
        m.init_particles(n, 1.0)
        oracles.append(o)
    for t in range(3, 18):
        e.step(t)
        for o in oracles:
            o.evolve(t)
    e.synchronize()
    for o, m in zip(oracles, e.members):
        compare_meas(o, m, 17)
        compare_window(o, m)
