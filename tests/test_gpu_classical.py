"""LambdaNLOclassical (TwoPIScalar_Evolution.jl:107-120): the sunset without
its rho^2 terms, on the HIP path through the C ABI against the oracle (1e-10,
norm-wise relative per lattice): full storage on the plane kernels and the
pencil passes past Nx = 64, cubic-orbit storage on the orbit products and on
the planes, graph replay, an ensemble member, the driver loop from a
parameter Dict, kb_sunset, and the physics pins (energy conservation,
damping).  The adjoint's classical cases are in test_gpu_adjoint.py."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10
CLASSICAL = 4


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


def compare(o, g, t):
    mo, mg = o.measure(t), g.measure(t)
    for k in ("E", "E_kin", "E_pot", "M2"):
        assert abs(mg[k] - mo[k]) <= RTOL * abs(mo[k]), (t, k, mg[k], mo[k])
    assert abs(mg["E_mem"] - mo["E_mem"]) <= RTOL * max(abs(mo["E"]), abs(mo["E_mem"]))
    assert rel(mg["n"], mo["n"]) <= RTOL, t


def check_window(o, g):
    assert o.window() == g.window()
    tmin, tmax = o.window()
    for a in range(tmin, tmax + 1):
        for b in range(tmin, a + 1):
            for w in "Fr":
                assert rel(g.get(w, a, b), o.get(w, a, b)) <= RTOL, (w, a, b)


CASES = [
    # Nx, sdim, Nmem, steps, orbit storage, transform form (None: default)
    pytest.param(6, 3, 7, 14, False, None, id="6^3"),
    pytest.param(5, 3, 4, 14, False, None, id="5^3-odd"),
    pytest.param(8, 2, 6, 20, False, None, id="8^2"),
    pytest.param(12, 1, 5, 20, False, None, id="12-1d"),
    pytest.param(72, 2, 4, 8, False, None, id="72^2-pencils"),
    pytest.param(8, 3, 9, 16, True, None, id="8^3-orbit-product"),
    pytest.param(8, 3, 9, 16, True, 0, id="8^3-orbit-planes"),
]


@pytest.mark.parametrize("Nx,sdim,Nmem,steps,sym,form", CASES)
def test_classical_trajectory_matches_oracle(Nx, sdim, Nmem, steps, sym, form):
    from qftkb import KBContext
    print(f"\n[case] classical {Nx}^{sdim} N_mem={Nmem} orbits={sym} form={form}", flush=True)
    args = (Nx, sdim, Nmem, 0.1, 1.5, 1, 1.0, 1.0)
    o = Oracle(*args, pexp=CLASSICAL)
    g = KBContext(*args, pexp=CLASSICAL, symmetric=sym)
    if form is not None:
        g.set_fft_tiles(form)
    n = thermal(o.kL2())
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    for t in range(3, steps + 1):
        o.evolve(t)
        g.step(t)
        for tp in range(o.window()[0], t):
            for w in "Fr":
                assert rel(g.sigma(w, tp), o.sigma(w, tp)) <= RTOL, (t, w, tp)
        compare(o, g, t)
    check_window(o, g)
    # and it is not the quantum sunset: the two trajectories differ
    q = Oracle(*args, pexp=1)
    q.init_particles(n, 1.0)
    for t in range(3, steps + 1):
        q.evolve(t)
    tmin = o.window()[0]
    assert rel(g.get("F", steps, tmin), q.get("F", steps, tmin)) > 1e-7
    g.close()


def test_classical_graph_replay_and_ensemble_member():
    from qftkb import KBContext, KBEnsemble
    print("\n[case] classical graph replay and ensemble", flush=True)
    args = (6, 3, 6, 0.1, 1.4, 2, 1.0, 1.0)
    o = Oracle(*args, pexp=CLASSICAL)
    g = KBContext(*args, pexp=CLASSICAL)
    g.use_graph(True)
    runs = [dict(Nx=6, sdim=3, NstepsinMemory=6, dt=0.1, Lambda=1.4, ONgroup=2, Mass=1.0, Mass2=1.0,
                 pexp=CLASSICAL),
            dict(Nx=8, sdim=3, NstepsinMemory=5, dt=0.1, Lambda=1.0, ONgroup=1, Mass=1.0, Mass2=1.0)]
    ens = KBEnsemble(runs)
    ens.use_graph(True)
    n = thermal(o.kL2())
    for c in (o, g, ens.members[0]):
        c.init_particles(n, 1.0)
    ens.members[1].init_particles(thermal(ens.members[1].kL2()), 1.0)
    for t in range(3, 20):
        o.evolve(t)
        g.step(t)
        ens.step(t)
        compare(o, g, t)
        compare(o, ens.members[0], t)
        g.synchronize()
    check_window(o, g)
    check_window(o, ens.members[0])
    ens.close()


@pytest.mark.parametrize("num", ["GPUfull", "GPUorbits"])
def test_classical_parameter_dict_run(num):
    """"Pexp" => "LambdaNLOclassical" through the driver loop."""
    from qftkb import TwoPIScalarfile, run
    print(f"\n[case] classical parameter Dict, {num}", flush=True)
    p = {"Mass": 1.0, "Nx": 6, "sdim": 3, "dt": 0.1, "Nsteps": 12, "NstepsinMemory": 8, "Nmeas": 6,
         "Mod": "phi4", "Lambda": 1.5, "ONgroup": 1, "init": "Thermal", "T": 2.0, "Reno": "RNone",
         "Pexp": "LambdaNLOclassical", "Num": num, "Nchunks": 0, "tag": "", "ov": True}
    sol = run(TwoPIScalarfile(), dict(p))
    o = Oracle(6, 3, 8, 0.1, 1.5, 1, 1.0, 1.0, pexp=CLASSICAL)
    o.init_particles(thermal(o.kL2()), 1.0)
    for t in range(3, 15):
        o.evolve(t)
        if (t - 2) % 2 == 0:
            mo = o.measure(t)
            m = sol.measurearray[(t - 2) // 2]
            assert abs(m.E - mo["E"]) <= RTOL * abs(mo["E"])
            assert rel(m.n, mo["n"]) <= RTOL


@pytest.mark.parametrize("Nx,sdim", [(8, 3), (7, 2), (40, 3)])
def test_classical_sunset_of_device_lattices(Nx, sdim):
    from hipmem import DeviceArray
    from qftkb._abi import sunset, sunset_workspace_bytes
    print(f"\n[case] kb_sunset classical {Nx}^{sdim}", flush=True)
    rng = np.random.default_rng(Nx)
    N = 5
    o = Oracle(Nx, sdim, N, 0.1, 1.2, 3, 1.0, 1.0, pexp=CLASSICAL)
    o.init_particles(np.zeros(o.V), 1.0)
    for _ in range(N - 2):
        o.expand()
    for a in range(1, N + 1):
        for b in range(1, a + 1):
            o.set("F", a, b, rng.standard_normal(o.V))
            o.set("r", a, b, rng.standard_normal(o.V))
    tm = N - 1
    o.selfenergies(tm)
    tps = range(1, tm + 1)
    F = np.array([o.get("F", tm, tp) for tp in tps])
    R = np.array([o.get("r", tm, tp) for tp in tps])
    nrows, V = F.shape
    dF, dR = DeviceArray.from_host(F), DeviceArray.from_host(R)
    SF, SR = DeviceArray(nrows * V), DeviceArray(nrows * V)
    nb = sunset_workspace_bytes(Nx, sdim, nrows)
    ws = DeviceArray(nb, np.uint8)
    sunset(Nx, sdim, 1.2, 3, nrows, dF.ptr, dR.ptr, SF.ptr, SR.ptr, ws.ptr, nb, classical=True)
    gF, gR = SF.host().reshape(nrows, V), SR.host().reshape(nrows, V)
    for q, tp in enumerate(tps):
        assert rel(gF[q], o.sigma("F", tp)) <= RTOL
        assert rel(gR[q], o.sigma("r", tp)) <= RTOL


def test_classical_energy_conservation_and_damping_on_gpu():
    from qftkb import KBContext
    from test_oracle import ECONS_DTS, energy_converges, energy_drift, rho_envelope, spectral_damped
    print("\n[case] classical energy conservation 4^3 lambda=10", flush=True)

    class Ctx:
        def __init__(self, N, dt):
            self.c = KBContext(4, 3, N, dt, 10.0, 1, 1.0, 1.0, pexp=CLASSICAL)

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

    d = [energy_drift(Ctx, dt) for dt in ECONS_DTS]
    assert energy_converges(d), d
    g = Ctx(152, 0.1)
    g.init_particles(thermal(g.kL2()), 1.0)
    for t in range(3, 153):
        g.evolve(t)
    assert spectral_damped(rho_envelope(g, 152))
