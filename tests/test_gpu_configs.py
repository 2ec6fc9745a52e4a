"""GPU parity at the lattice sizes and states of the larger workloads, and
for states whose bare mass differs from the discretisation mass.

* 40^3 (thermal) and 48^3 (box n0 = 4, Q = 1, lambda = 0.5, dt = 0.05): the
  plane kernels compiled for Nx = 40 and 48, in every block size, and the
  cubic-orbit storage of the same states, against the oracle at a short
  memory window; 64^3 (thermal), the largest lattice of the plane kernels,
  in every block size.
* The seeded 8-run ensemble {16,16,24,24,32,32,40,40}^3 with box
  occupations, graph-replayed, at N_mem = 8 against one oracle run per
  member.
* Mass2 != Mass^2: the RMass counterterm (ModelExpansionRenormalization.jl:
  23-47), the tachyonic model (:30-36) and the gap-mass initialisation
  (Initialization.jl:250-266, 289-320), through the reference-shaped driver
  loop, in full and orbit storage.
* The physics pins of the memory integrals on the HIP path: energy
  conservation converges with dt, and the spectral function is damped.

Tolerance 1e-10, norm-wise relative per lattice, as everywhere else.
"""
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
                e = rel(g.get(w, a, b), o.get(w, a, b))
                assert e <= tol, f"{w}({a},{b}) rel err {e:.3e}"


def check_meas(o, g, t, tol=RTOL):
    mo, mg = o.measure(t), g.measure(t)
    for k in ("E", "E_kin", "E_pot", "M2"):
        assert abs(mg[k] - mo[k]) <= tol * abs(mo[k]), f"{k}: {mg[k]} vs {mo[k]}"
    assert abs(mg["E_mem"] - mo["E_mem"]) <= tol * max(abs(mo["E"]), abs(mo["E_mem"]))
    assert mg["n_unstable"] == mo["nstab"]
    for k in ("F", "n", "ddF"):
        assert rel(mg[k], mo[k]) <= tol, f"per-site {k}"
    return mo, mg


def run_pair(o, g, t_end, sigma=True):
    for t in range(3, t_end + 1):
        o.evolve(t)
        g.step(t)
        if sigma:
            for tp in range(o.window()[0], t):
                for w in "Fr":
                    e = rel(g.sigma(w, tp), o.sigma(w, tp))
                    assert e <= RTOL, f"Sigma{w}({t - 1},{tp}) rel err {e:.3e}"
        check_meas(o, g, t)
    check_window(o, g)


# --- 40^3 and 48^3 -----------------------------------------------------------
LARGE = {
    # the ensemble's largest lattice: thermal T = 2, lambda = 1, dt = 0.1
    40: dict(dt=0.1, lam=1.0, occ=lambda k2: thermal(k2)),
    # the largest single-GPU workload's lattice and state: box n0 = 4, Q = 1,
    # lambda = 0.5, dt = 0.05
    48: dict(dt=0.05, lam=0.5, occ=lambda k2: np.where(k2 <= 1.0 * (1 + 1e-12), 4.0, 0.0)),
    # the largest lattice the plane kernels take: thermal T = 2, lambda = 1
    64: dict(dt=0.1, lam=1.0, occ=lambda k2: thermal(k2)),
}


def _fft_variants():
    from qftkb import KBContext
    c = KBContext(8, 3, 4, 0.1, 1.0, 1, 1.0, 1.0)
    v = list(range(len(c.fft_shapes())))
    c.close()
    return v


@pytest.mark.parametrize("Nx", [40, 48, 64])
def test_large_lattice_every_plane_shape(Nx):
    """The plane kernels for Nx = 40 / 48 / 64 in every block size: 10
    steps through a sliding N_mem = 6 window, against the oracle."""
    from qftkb import KBContext
    cfg = LARGE[Nx]
    for v in _fft_variants():
        print(f"\n[case] {Nx}^3 N_mem=6 plane shape {v}", flush=True)
        o = Oracle(Nx, 3, 6, cfg["dt"], cfg["lam"], 1, 1.0, 1.0)
        g = KBContext(Nx, 3, 6, cfg["dt"], cfg["lam"], 1, 1.0, 1.0)
        g.set_fft_tiles(v)
        n = cfg["occ"](o.kL2())
        o.init_particles(n, 1.0)
        g.init_particles(n, 1.0)
        run_pair(o, g, 12)
        g.close()
        o.close()


@pytest.mark.parametrize("Nx", [40, 48])
def test_large_lattice_orbit_storage(Nx):
    """The same states in cubic-orbit storage (orbit products, then the
    plane kernels over the expanded lattice), against the oracle."""
    from qftkb import FFT_ORBIT_PRODUCT, KBContext
    cfg = LARGE[Nx]
    for form in (FFT_ORBIT_PRODUCT, 0):
        print(f"\n[case] {Nx}^3 N_mem=7 orbit storage, transform form {form}", flush=True)
        o = Oracle(Nx, 3, 7, cfg["dt"], cfg["lam"], 1, 1.0, 1.0)
        g = KBContext(Nx, 3, 7, cfg["dt"], cfg["lam"], 1, 1.0, 1.0, symmetric=True)
        g.set_fft_tiles(form)
        n = cfg["occ"](o.kL2())
        o.init_particles(n, 1.0)
        g.init_particles(n, 1.0)
        run_pair(o, g, 11)
        g.close()
        o.close()


# --- the seeded 8-run ensemble ------------------------------------------------
@pytest.mark.parametrize("symmetric", [False, True])
def test_seeded_ensemble_members_match_oracle(symmetric):
    """The uneven ensemble {16,16,24,24,32,32,40,40}^3 with its seeded
    couplings and box occupations (the benchmark's own draws), every step
    replayed from one captured graph, at N_mem = 8 for 12 steps: every member
    against its own oracle run."""
    import bench
    from qftkb import KBEnsemble, lattice
    runs, occ = bench.ensemble_runs(nmem=8)
    assert [r["Nx"] for r in runs] == [16, 16, 24, 24, 32, 32, 40, 40]
    if symmetric:
        runs = [dict(r, symmetric=True) for r in runs]
    print(f"\n[case] seeded 8-run ensemble, N_mem=8, orbit storage={symmetric}", flush=True)
    ens = KBEnsemble(runs)
    ens.use_graph(True)
    oracles = []
    for r, m, oc in zip(runs, ens.members, occ):
        o = Oracle(r["Nx"], r["sdim"], r["NstepsinMemory"], r["dt"], r["Lambda"], r["ONgroup"], r["Mass"],
                   r["Mass2"])
        n = lattice.n_box(o.kL2(), oc["n0"], oc["Q"])
        o.init_particles(n, r["Mass"])
        m.init_particles(n, r["Mass"])
        oracles.append(o)
    for t in range(3, 15):
        ens.step(t)
        for o in oracles:
            o.evolve(t)
        if t % 4 == 2 or t == 14:
            for o, m in zip(oracles, ens.members):
                check_meas(o, m, t)
        ens.synchronize()
    for i, (o, m) in enumerate(zip(oracles, ens.members)):
        print(f"[case] member {i} ({runs[i]['Nx']}^3) window", flush=True)
        check_window(o, m)
    ens.close()


# --- Mass2 != Mass^2 -----------------------------------------------------------
def _params(**kw):
    p = {"Mass": 1.0, "Nx": 6, "sdim": 3, "dt": 0.1, "Nsteps": 12, "NstepsinMemory": 8, "Nmeas": 12,
         "Mod": "phi4", "Lambda": 1.5, "ONgroup": 1, "init": "Thermal", "T": 2.0, "n": 1.0, "Reno": "RNone",
         "Pexp": "LambdaNLO", "Num": "GPUfull", "Nchunks": 0, "tag": "", "ov": True}
    p.update(kw)
    return p


MASS_CASES = {
    # bare Mass2 = Mass^2 + counterterm, thermal init with Mass
    "rmass_thermal": dict(Reno="RMass"),
    # tachyonic: Mass2 = -Mass^2 + counterterm (negative), init with Mass
    "tachyonic_rmass": dict(Mod="phi4tachyonic", Reno="RMass", init="Pnr", n=1.0, Lambda=1.0),
    # constant occupation: initial mass is the LO gap mass
    "pnr_gapmass": dict(init="Pnr", n=0.5),
    # gap mass from the renormalised Mass2
    "pnr_gapmass_rmass": dict(init="Pnr", n=0.5, Reno="RMass", ONgroup=4),
    # TopHatT2 initialises at the gap mass as well
    "tophat2_gapmass": dict(init="TopHatT2", Nx=8, NstepsinMemory=6, Pexp="LambdaNLO", dt=0.05),
}


@pytest.mark.parametrize("num", ["GPUfull", "GPUorbits"])
@pytest.mark.parametrize("case", sorted(MASS_CASES))
def test_bare_mass_differs_from_mass(case, num):
    """Driver-loop runs (problem -> initialize -> evolve/measure) whose
    evolution Mass2 differs from Mass^2, against the oracle given the same
    Mass2 and initial mass: E, M2 and n(p,t) at every step."""
    from qftkb import QFTdynamicsProblem, QFTdynamicsSolution, TwoPIScalarfile, evolve, initialize, measure
    from qftkb.problem import getGapMass, getparticlenr
    p = _params(**MASS_CASES[case], Num=num)
    print(f"\n[case] {case} {num}", flush=True)
    prob = QFTdynamicsProblem(TwoPIScalarfile(), dict(p))
    Mass2 = prob.model.Mass2
    if prob.model.tachyonic or p["init"] in ("Thermal", "Box", "TopHatT1"):
        m0 = p["Mass"]
    else:
        m0 = getGapMass(prob)
    # every case evolves with Mass2 != Mass^2 or starts at a gap mass != Mass
    assert abs(Mass2 - p["Mass"] ** 2) > 1e-3 or abs(m0 - p["Mass"]) > 1e-3
    sol, tmp = QFTdynamicsSolution(TwoPIScalarfile(), prob)
    initialize(sol, tmp)
    d = prob.disc
    o = Oracle(d.Nx, d.sdim, d.NstepsinMemory, d.dt, prob.model.Lambda, prob.model.ONgroup, d.Mass, Mass2)
    o.init_particles(getparticlenr(prob.init, d), m0)
    g = sol.simdata.ctx
    assert abs(g.hartreemass2() - o.hartreemass2) <= RTOL * abs(o.hartreemass2)
    for t in range(3, prob.simsetup.Nsteps + 3):
        evolve(sol, tmp, t)
        measure(sol, t)
        o.evolve(t)
        mo = o.measure(t)
        m = sol.measurearray[t - 2]
        assert abs(m.E - mo["E"]) <= RTOL * abs(mo["E"]), (t, m.E, mo["E"])
        assert abs(m.M2 - mo["M2"]) <= RTOL * abs(mo["M2"]), (t, m.M2, mo["M2"])
        assert abs(m.E_pot - mo["E_pot"]) <= RTOL * abs(mo["E_pot"])
        assert rel(m.n, mo["n"]) <= RTOL, t
    check_window(o, g)


# --- physics pins of the memory integrals on the HIP path -------------------------
def _gpu_factory(Nx, sdim, ON, lam=10.0):
    from qftkb import KBContext

    class Ctx:
        def __init__(self, N, dt):
            self.c = KBContext(Nx, sdim, N, dt, lam, ON, 1.0, 1.0)

        def kL2(self):
            return self.c.kL2()

        def init_particles(self, n, m):
            self.c.init_particles(n, m)

        def evolve(self, t):
            self.c.step(t)

        def measure(self, t):
            return self.c.measure(t, modes=False)

        def get(self, w, a, b):
            return self.c.get(w, a, b)

    return Ctx


@pytest.mark.parametrize("Nx,sdim,ON", [(4, 3, 1), (6, 2, 4)])
def test_energy_conservation_and_damping_on_gpu(Nx, sdim, ON):
    """lambda = 10, full memory: the HIP path's energy drift shrinks with dt
    like the oracle's (>= 2.8x, then >= 2.4x per halving, <= 0.8% at
    dt = 0.025), and the spectral function rho(t, 2) is damped over t <= 15."""
    from test_oracle import ECONS_DTS, energy_converges, energy_drift, rho_envelope, spectral_damped
    print(f"\n[case] energy conservation {Nx}^{sdim} O({ON}) lambda=10", flush=True)
    fac = _gpu_factory(Nx, sdim, ON)
    d = [energy_drift(fac, dt) for dt in ECONS_DTS]
    assert energy_converges(d), d
    od = [energy_drift(lambda N, dt: Oracle(Nx, sdim, N, dt, 10.0, ON, 1.0, 1.0), dt) for dt in ECONS_DTS]
    assert np.allclose(d, od, rtol=1e-6), (d, od)
    steps = 150
    g = fac(steps + 2, 0.1)
    g.init_particles(thermal(g.kL2()), 1.0)
    for t in range(3, steps + 3):
        g.evolve(t)
    env = rho_envelope(g, steps + 2)
    assert spectral_damped(env), env
