"""Cubic-orbit storage (kb_create_sym): one stored value per orbit of the
lattice's cubic group.  For cubic-symmetric initial data the KB step keeps
the state symmetric, so orbit storage must reproduce full storage (and the
oracle) exactly up to FP64 rounding: tolerance 1e-12 against full storage,
1e-10 against the oracle."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


@pytest.mark.parametrize("Nx,sdim,orbits", [(16, 3, 165), (32, 3, 969), (16, 2, 45), (16, 1, 9), (5, 3, 10)])
def test_orbit_counts(Nx, sdim, orbits):
    from qftkb import KBContext
    c = KBContext(Nx, sdim, 4, 0.1, 1.0, 1, 1.0, 1.0, symmetric=True)
    assert c.storage_info() == (orbits, Nx ** sdim)
    c.close()


CASES = [
    # (Nx, sdim, Nmem, dt, lambda, ON, Mass, Mass2), pexp, steps
    pytest.param((8, 3, 10, 0.1, 1.3, 1, 1.0, 1.0), 1, 40, id="8^3-sliding"),
    pytest.param((6, 2, 7, 0.08, 1.7, 4, 0.9, 0.81), 1, 30, id="6^2-O(4)"),
    pytest.param((10, 1, 6, 0.1, 2.0, 2, 1.0, 1.0), 1, 30, id="10-1d"),
    pytest.param((16, 3, 20, 0.1, 1.0, 1, 1.0, 1.0), 1, 45, id="16^3"),
    pytest.param((8, 3, 10, 0.1, 1.0, 1, 1.0, 1.0), 0, 30, id="8^3-LO"),
]


@pytest.mark.parametrize("args,pexp,T", CASES)
def test_orbit_storage_equals_full_storage(args, pexp, T):
    from qftkb import KBContext
    print(f"\n[case] orbit storage {args} pexp={pexp} T={T}", flush=True)
    full = KBContext(*args, pexp=pexp)
    red = KBContext(*args, pexp=pexp, symmetric=True)
    assert red.storage_info()[0] < full.storage_info()[0]
    n = thermal(full.kL2(), m=args[6])
    # the same terms summed in each site's own axis order: equal to rounding
    np.testing.assert_allclose(red.kL2(), full.kL2(), rtol=1e-14, atol=1e-14)
    full.init_particles(n, args[6])
    red.init_particles(n, args[6])
    for t in range(3, T + 1):
        full.step(t)
        red.step(t)
        if t % 5 == 0 or t == T:
            mf, mr = full.measure(t), red.measure(t)
            for k in ("E", "E_kin", "E_pot", "E_mem", "M2"):
                assert abs(mr[k] - mf[k]) <= 1e-12 * max(abs(mf[k]), 1e-300), (t, k, mr[k], mf[k])
            assert rel(mr["n"], mf["n"]) <= 1e-12
    tmin, tmax = full.window()
    assert red.window() == (tmin, tmax)
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(red.get(w, tmax, tp), full.get(w, tmax, tp)) <= 1e-12, (w, tp)
    if pexp:
        # Sigma(t-1, t') over the window, against the scale of the whole ring
        # (Sigr(t-1, t-1) is zero up to transform rounding in both storages)
        for w in ("F", "r"):
            sf = np.array([full.sigma(w, tp) for tp in range(tmin, tmax)])
            sr = np.array([red.sigma(w, tp) for tp in range(tmin, tmax)])
            assert np.max(np.abs(sr - sf)) <= 1e-12 * np.max(np.abs(sf)), w
    full.close()
    red.close()


def test_orbit_storage_config1_box_against_oracle():
    """The 16^3 box state (60 steps) in orbit storage against
    the oracle, n(p,t) and E at every 10th step."""
    from qftkb import KBContext
    from qftkb.lattice import n_box
    print("\n[case] 16^3 box, orbit storage vs oracle", flush=True)
    args = (16, 3, 62, 0.1, 1.0, 1, 1.0, 1.0)
    o = Oracle(*args)
    g = KBContext(*args, symmetric=True)
    n = n_box(o.kL2(), 2.0, 1.0)
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    for t in range(3, 61):
        o.evolve(t)
        g.step(t)
        if t % 10 == 0:
            mo, mg = o.measure(t), g.measure(t)
            assert abs(mg["E"] - mo["E"]) <= 1e-10 * abs(mo["E"])
            assert rel(mg["n"], mo["n"]) <= 1e-10
    g.close()


def _thermal_run(ctx, N, steps):
    n = thermal(ctx.kL2())
    ctx.fill_free_window(N, n, 1.0)
    for t in range(N + 1, N + 1 + steps):
        ctx.step(t)
    return N + steps


@pytest.mark.parametrize("args,filled", [
    pytest.param((16, 3, 20, 0.1, 1.0, 1, 1.0, 1.0), False, id="16^3-N20"),
    pytest.param((12, 2, 9, 0.1, 1.5, 2, 1.0, 1.0), False, id="12^2-O(2)"),
    pytest.param((11, 3, 6, 0.1, 1.0, 1, 1.0, 1.0), False, id="11^3-odd"),
    pytest.param((32, 3, 400, 0.1, 1.0, 1, 1.0, 1.0), True, id="32^3-N400-filled"),
])
def test_orbit_product_equals_plane_kernels(args, filled):
    """Orbit storage, LambdaNLO: the self-energy as products with the orbit
    cosine matrix (the default) and as plane FFTs of the expanded lattice are
    the same sums in another order: one trajectory to 1e-12."""
    from qftkb import KBContext
    from qftkb._abi import FFT_ORBIT_PRODUCT
    print(f"\n[case] orbit product vs plane kernels {args}", flush=True)
    a = KBContext(*args, symmetric=True)
    b = KBContext(*args, symmetric=True, tiles=None)
    assert a.orbit_product and a.fft_tiles() == FFT_ORBIT_PRODUCT
    b.set_fft_tiles(0)
    assert b.fft_tiles() == 0
    if filled:
        T = _thermal_run(a, args[2], 4)
        assert _thermal_run(b, args[2], 4) == T
    else:
        n = thermal(a.kL2())
        a.init_particles(n, 1.0)
        b.init_particles(n, 1.0)
        T = 3 * args[2]
        for t in range(3, T + 1):
            a.step(t)
            b.step(t)
    tmin, tmax = a.window()
    assert tmax == T
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(a.get(w, tmax, tp), b.get(w, tmax, tp)) <= 1e-12, (w, tp)
    for w in ("F", "r"):
        sa = np.array([a.sigma(w, tp) for tp in range(tmin, tmax)])
        sb = np.array([b.sigma(w, tp) for tp in range(tmin, tmax)])
        assert np.max(np.abs(sa - sb)) <= 1e-12 * np.max(np.abs(sb)), w
    ma, mb = a.measure(T), b.measure(T)
    assert abs(ma["E"] - mb["E"]) <= 1e-12 * abs(mb["E"])
    # the autotuner times the plane shapes and the product (last entry)
    r = a.autotune_fft(reps=2)
    assert len(r["ms"]) == len(a.fft_shapes()) + 1 and all(x > 0 for x in r["ms"])
    assert r["variant"] == FFT_ORBIT_PRODUCT or 0 <= r["variant"] < len(a.fft_shapes())
    a.close()
    b.close()


def test_orbit_product_only_for_orbit_storage_transform_stages():
    """The orbit product exists in orbit-storage contexts with a transform
    stage (LambdaNLO sunset, NinverseNLO bubble and chain), not in full
    storage or LambdaLO."""
    from qftkb import KBContext, KBError
    from qftkb._abi import FFT_ORBIT_PRODUCT
    args = (8, 3, 10, 0.1, 1.0, 1, 1.0, 1.0)
    for kw in ({}, {"pexp": 3}, {"symmetric": True, "pexp": 0}):
        c = KBContext(*args, **kw)
        assert not c.orbit_product
        with pytest.raises(KBError):
            c.set_fft_tiles(FFT_ORBIT_PRODUCT)
        c.close()
    for pexp in (1, 3):
        c = KBContext(*args, symmetric=True, pexp=pexp)
        assert c.orbit_product
        c.set_fft_tiles(0)
        c.set_fft_tiles(FFT_ORBIT_PRODUCT)
        assert c.fft_tiles() == FFT_ORBIT_PRODUCT
        c.close()


def test_orbit_storage_graph_replay_equals_eager():
    """Orbit storage (orbit product, split memory integral) replayed from a
    captured hipGraph gives the eager trajectory: the same kernels in the same
    order, so equal to rounding, checked after many replays while the window
    fills and slides, with eager measurements between the replays."""
    from qftkb import KBContext
    args = (16, 3, 12, 0.1, 1.0, 1, 1.0, 1.0)
    print("\n[case] orbit storage, graph replay vs eager", flush=True)
    a = KBContext(*args, symmetric=True)
    b = KBContext(*args, symmetric=True)
    assert a.orbit_product and b.orbit_product
    b.use_graph(True)
    n = thermal(a.kL2())
    a.init_particles(n, 1.0)
    b.init_particles(n, 1.0)
    for t in range(3, 41):
        a.step(t)
        b.step(t)
        if t % 4 == 0 or t == 40:
            tmin, tmax = a.window()
            assert b.window() == (tmin, tmax)
            for tp in range(tmin, tmax + 1):
                for w in "Fr":
                    assert rel(b.get(w, tmax, tp), a.get(w, tmax, tp)) <= 1e-14, (t, w, tp)
            ma, mb = a.measure(t), b.measure(t)
            assert abs(ma["E"] - mb["E"]) <= 1e-14 * abs(ma["E"]), t
    a.close()
    b.close()


def test_orbit_storage_above_product_limit_uses_plane_kernels():
    """Past 4096 orbits (56^3 has 4495) orbit storage keeps the plane-FFT
    self-energy; it still equals full storage."""
    from qftkb import KBContext
    args = (56, 3, 4, 0.1, 1.0, 1, 1.0, 1.0)
    print("\n[case] 56^3 orbit storage without the orbit product", flush=True)
    full = KBContext(*args)
    red = KBContext(*args, symmetric=True)
    assert red.storage_info()[0] == 4495 and not red.orbit_product
    n = thermal(full.kL2())
    full.init_particles(n, 1.0)
    red.init_particles(n, 1.0)
    for t in range(3, 9):
        full.step(t)
        red.step(t)
    tmin, tmax = full.window()
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(red.get(w, tmax, tp), full.get(w, tmax, tp)) <= 1e-12, (w, tp)
    full.close()
    red.close()


@pytest.mark.parametrize("args,pexp,T", [
    pytest.param((72, 2, 6, 0.1, 1.0, 1, 1.0, 1.0), 1, 20, id="72^2-NLO"),
    pytest.param((72, 3, 4, 0.1, 1.0, 1, 1.0, 1.0), 0, 8, id="72^3-LO"),
])
def test_orbit_storage_past_the_plane_kernels(args, pexp, T):
    """Nx > 64: orbit storage runs LambdaNLO on the orbit product alone (the
    plane kernels stop at 64; full storage uses its pencil passes there) and
    LambdaLO without transforms; both equal full storage."""
    from qftkb import KBContext, KBError
    print(f"\n[case] orbit storage past the plane kernels {args} pexp={pexp}", flush=True)
    full = KBContext(*args, pexp=pexp)
    red = KBContext(*args, pexp=pexp, symmetric=True)
    assert red.orbit_product == (pexp == 1)
    if pexp == 1:
        with pytest.raises(KBError):
            red.set_fft_tiles(0)
    n = thermal(full.kL2())
    full.init_particles(n, 1.0)
    red.init_particles(n, 1.0)
    for t in range(3, T + 1):
        full.step(t)
        red.step(t)
    tmin, tmax = full.window()
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(red.get(w, tmax, tp), full.get(w, tmax, tp)) <= 1e-12, (w, tp)
    mf, mr = full.measure(T), red.measure(T)
    assert abs(mr["E"] - mf["E"]) <= 1e-12 * abs(mf["E"])
    full.close()
    red.close()


def test_orbit_storage_ninverse_past_the_plane_kernels_against_oracle():
    """NinverseNLO at 72^2 exists only in orbit storage (full storage needs the
    plane kernels, Nx <= 64): trajectory against the oracle."""
    from qftkb import KBContext
    args = (72, 2, 5, 0.05, 10.0, 8, 1.0, 1.0)
    print("\n[case] 1/N orbit storage 72^2 vs oracle", flush=True)
    o = Oracle(*args, pexp=3)
    g = KBContext(*args, pexp=3, symmetric=True)
    assert g.orbit_product
    n = thermal(o.kL2())
    o.init_particles(n, 1.0)
    g.init_particles(n, 1.0)
    for t in range(3, 13):
        o.evolve(t)
        g.step(t)
    tmin, tmax = o.window()
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(g.get(w, tmax, tp), o.get(w, tmax, tp)) <= 1e-10, (w, tp)
    scale = max(np.abs(o.sigma(w, tp)).max() for w in "Fr" for tp in range(tmin, tmax))
    for tp in range(tmin, tmax):
        for w in "Fr":
            assert np.abs(g.sigma(w, tp) - o.sigma(w, tp)).max() <= 1e-10 * scale, (w, tp)
    g.close()


def test_orbit_storage_every_tile_shape():
    """The split memory integral runs the context's tile shape: its item
    height (32, 64 or 128 slot rows) sets the item table, and its rows per
    batch and threads the split kernel.  Every shape the autotuner can pick
    gives the default trajectory."""
    from qftkb import KBContext
    args = (16, 3, 20, 0.1, 1.0, 1, 1.0, 1.0)
    ref = KBContext(*args, symmetric=True, tiles=None)
    shapes = ref.tile_shapes()
    n = thermal(ref.kL2())
    ref.init_particles(n, 1.0)
    for t in range(3, 31):
        ref.step(t)
    tmin, tmax = ref.window()
    assert len(shapes) == 13
    for v in range(len(shapes)):
        print(f"\n[case] orbit storage, tile {shapes[v]}", flush=True)
        g = KBContext(*args, symmetric=True, tiles=None)
        g.set_tiles(v)
        g.init_particles(n, 1.0)
        for t in range(3, 31):
            g.step(t)
        for tp in range(tmin, tmax + 1):
            for w in "Fr":
                assert rel(g.get(w, tmax, tp), ref.get(w, tmax, tp)) <= 1e-12, (shapes[v], w, tp)
        g.close()
    # the autotuner re-times the split form mid-run and the run continues
    g = KBContext(*args, symmetric=True, tiles=None)
    g.init_particles(n, 1.0)
    for t in range(3, 31):
        g.step(t)
        if t == 12:
            r = g.autotune(reps=2)
            assert all(x > 0 for x in r["ms"]) and g.tiles() == r["variant"]
    for tp in range(tmin, tmax + 1):
        for w in "Fr":
            assert rel(g.get(w, tmax, tp), ref.get(w, tmax, tp)) <= 1e-12, ("autotune", w, tp)
    g.close()
    ref.close()


@pytest.mark.parametrize("init", ["Thermal", "Box", "TopHatT1", "Gauss"])
def test_parameter_dict_run_with_orbit_numerics(init):
    """"Num" => "GPUorbits" runs the reference-shaped driver loop on orbit
    storage; every initial state of the reference is cubic-symmetric, so the
    measurements equal those of "GPUfull" to rounding."""
    from qftkb import TwoPIScalarfile, run
    print(f"\n[case] parameter-Dict run, GPUorbits vs GPUfull, init={init}", flush=True)
    base = {"Mass": 1.0, "Nx": 8, "sdim": 3, "dt": 0.1, "Nsteps": 30, "NstepsinMemory": 12, "Nmeas": 10,
            "Mod": "phi4", "Lambda": 1.0, "ONgroup": 1, "init": init, "T": 2.0, "n0": 2.0, "Q": 1.0,
            "xi": 0.4, "eta": 5.0, "sig": 1.0, "Reno": "RNone", "Pexp": "LambdaNLO", "Nchunks": 0, "tag": "",
            "ov": True}
    full = run(TwoPIScalarfile(), dict(base, Num="GPUfull"))
    orb = run(TwoPIScalarfile(), dict(base, Num="GPUorbits"))
    assert orb.simdata.ctx.storage_info()[0] < full.simdata.ctx.storage_info()[0]
    for mf, mo in zip(full.measurearray, orb.measurearray):
        assert mf is not None and mo is not None
        for k in ("E", "E_kin", "E_pot", "E_mem"):
            a, b = getattr(mo, k), getattr(mf, k)
            assert abs(a - b) <= 1e-12 * max(abs(b), abs(mf.E)), (k, a, b)
        np.testing.assert_allclose(mo.n, mf.n, rtol=1e-11, atol=1e-13)


def test_plane_workspace_only_where_a_plane_stage_runs():
    """The plane kernels' workspace (half a complex lattice per slot: the
    planes z in [0, Nx/2], plus the 1/N transforms, a whole lattice per slot)
    is held only by contexts that run a plane stage: not by LambdaLO, not by
    orbit-storage contexts on the orbit products, until a plane form is
    selected."""
    from qftkb import KBContext
    Nx, N = 16, 12
    Vl = Nx ** 3
    W = N * Nx * Nx * (Nx // 2 + 1) * 16
    Xn = N * Vl * 16
    lo_full = KBContext(Nx, 3, N, 0.1, 1.0, 1, 1.0, 1.0, pexp=0)
    nlo_full = KBContext(Nx, 3, N, 0.1, 1.0, 1, 1.0, 1.0)
    assert nlo_full.device_bytes() - lo_full.device_bytes() == W
    sym = KBContext(Nx, 3, N, 0.1, 1.0, 1, 1.0, 1.0, symmetric=True)
    b0 = sym.device_bytes()
    sym.set_fft_tiles(0)  # the plane kernels over the expanded lattice
    assert sym.device_bytes() - b0 == W
    sym.set_fft_tiles(-1)
    sym.set_fft_tiles(1)
    assert sym.device_bytes() - b0 == W  # allocated once
    ninv = KBContext(Nx, 3, N, 0.05, 2.0, 4, 1.0, 1.0, pexp=3, symmetric=True)
    c0 = ninv.device_bytes()
    ninv.set_fft_tiles(0)
    assert ninv.device_bytes() - c0 == W + Xn  # W and the kept transforms
    # the plane form still runs after the late allocation: a step against the
    # orbit-product form of the same state
    k2 = sym.kL2()
    n = 1.0 / (np.exp(np.sqrt(k2 + 1.0) / 2.0) - 1.0)
    ref = KBContext(Nx, 3, N, 0.1, 1.0, 1, 1.0, 1.0, symmetric=True)
    for c in (sym, ref):
        c.init_particles(n, 1.0)
        for t in range(3, 8):
            c.step(t)
    a, b = sym.get("F", 7, 5), ref.get("F", 7, 5)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    for c in (lo_full, nlo_full, sym, ninv, ref):
        c.close()


@pytest.mark.parametrize("args,pexp,T", [
    # the reference's own test lattices (test/runtests.jl): 1024 in 1D, 512 in 1D in the 1/N
    # expansion (TestTwoPIScalar_CPUred(512, 1, 4, "NinverseNLO", 50)); past 128 in 2D and 3D
    pytest.param((1024, 1, 5, 0.1, 1.0, 1, 1.0, 1.0), 1, 12, id="1024-1d-NLO"),
    pytest.param((512, 1, 5, 0.05, 2.0, 4, 1.0, 1.0), 3, 10, id="512-1d-1/N"),
    pytest.param((130, 2, 4, 0.1, 1.2, 2, 1.0, 0.8), 1, 8, id="130^2-NLO"),
    pytest.param((130, 3, 4, 0.1, 1.0, 1, 1.0, 1.0), 0, 8, id="130^3-LO"),
])
def test_orbit_storage_large_lattices_match_oracle(args, pexp, T):
    """Orbit storage past the plane kernels on large lattices (orbit products
    up to 4096 orbits; LambdaLO without transforms), against the oracle."""
    from qftkb import KBContext
    print(f"\n[case] orbit storage, large lattice {args} pexp={pexp}", flush=True)
    o = Oracle(*args, pexp=pexp, nthreads=8)
    g = KBContext(*args, pexp=pexp, symmetric=True)
    n = thermal(o.kL2(), m=args[6])
    o.init_particles(n, args[6])
    g.init_particles(n, args[6])
    for t in range(3, T + 1):
        o.evolve(t)
        g.step(t)
    tmin, tmax = o.window()
    assert g.window() == (tmin, tmax)
    for a in range(tmin, tmax + 1):
        for b in range(tmin, a + 1):
            for w in "Fr":
                assert rel(g.get(w, a, b), o.get(w, a, b)) <= 1e-10, (w, a, b)
    mo, mg = o.measure(T), g.measure(T)
    assert abs(mg["E"] - mo["E"]) <= 1e-10 * abs(mo["E"])
    assert rel(mg["n"], mo["n"]) <= 1e-10
    g.close()
    o.close()


def test_lattice_limits_are_refused():
    """Nx up to 2048 and at most 2^24 sites; beyond either, kb_create refuses."""
    from qftkb import KBContext, KBError
    for bad in [(2049, 1), (257, 3), (4097, 2)]:
        with pytest.raises(KBError):
            KBContext(bad[0], bad[1], 3, 0.1, 1.0, 1, 1.0, 1.0, pexp=0)
    c = KBContext(256, 3, 3, 0.1, 1.0, 1, 1.0, 1.0, pexp=0)  # 2^24 sites, no transforms
    c.close()
