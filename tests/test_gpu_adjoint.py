"""Hand-written adjoint kernels through the C ABI: gradients of
L = sum_p wgt(p) n(p, T) against the numpy reverse-mode twin
(tests/adjoint_twin.py, itself pinned against oracle finite differences)
and against central finite differences (h = 1e-5) of the C oracle.

Tolerances: adjoint vs twin 1e-10 (both exact, FP64 rounding only);
adjoint vs finite differences 1e-6 relative to the largest component."""
import numpy as np
import pytest

from oracle.kb_numpy import kL2_table
from oracle.oracle import Oracle
from tests.adjoint_twin import gradient, run_forward, reverse

pytestmark = pytest.mark.gpu

H = 1e-5
FD_TOL = 1e-6
TWIN_TOL = 1e-10


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / s if s > 0 else np.max(np.abs(a - b))


def gpu_gradient(args, n0, m, T, wgt, C=0):
    from qftkb import KBContext
    c = KBContext(*args[:8], pexp=args[8] if len(args) > 8 else 1)
    c.init_particles(n0, m)
    c.adj_record(C, T)
    for t in range(3, T + 1):
        c.step(t)
    L = float((wgt * c.measure(T)["n"]).sum())
    g = c.adj_gradient(wgt, n0=True, F0=True)
    assert c.window() == (1, 2)
    return c, L, g


def oracle_loss(args, n0, m, T, wgt):
    o = Oracle(*args)
    o.init_particles(np.ascontiguousarray(n0, dtype=np.float64), m)
    for t in range(3, T + 1):
        o.evolve(t)
    L = float((wgt * o.measure(T)["n"]).sum())
    o.close()
    return L


def with_lambda(args, lam):
    a = list(args)
    a[4] = lam
    return tuple(a)


def setup(args):
    k2 = kL2_table(args[0], args[1])
    m = args[6]
    return thermal(k2, m=m), m, np.exp(-k2)


CASES = [
    # (Nx, sdim, Nmem, dt, lambda, ON, Mass, Mass2, pexp), T
    pytest.param((4, 3, 5, 0.1, 1.3, 1, 1.0, 1.0, 1), 14, id="4^3-sliding"),
    pytest.param((4, 3, 16, 0.1, 1.0, 1, 1.0, 1.0, 1), 12, id="4^3-full-memory"),
    pytest.param((6, 2, 6, 0.08, 1.7, 4, 0.9, 0.81, 1), 15, id="6^2-O(4)"),
    pytest.param((5, 1, 4, 0.1, 2.0, 2, 1.0, 1.0, 1), 11, id="5-1d"),
    pytest.param((5, 3, 7, 0.1, 1.2, 1, 1.0, 1.0, 1), 25, id="5^3-ragged-blocks"),
    pytest.param((4, 3, 5, 0.1, 1.0, 1, 1.0, 1.0, 0), 14, id="4^3-LO"),
    pytest.param((4, 3, 6, 0.1, 1.6, 1, 1.0, 1.0, 4), 13, id="4^3-classical"),
    pytest.param((6, 3, 5, 0.1, 1.4, 2, 1.0, 1.0, 4), 12, id="6^3-classical-O(2)"),
    # a 1D lattice past 64 (the plane kernels' one-line form)
    pytest.param((256, 1, 5, 0.1, 1.2, 2, 1.0, 0.9, 1), 12, id="256-1d"),
]


@pytest.mark.parametrize("args,T", CASES)
def test_adjoint_matches_twin_and_oracle_fd(args, T):
    print(f"\n[case] adjoint {args} T={T}", flush=True)
    n0, m, wgt = setup(args)
    c, L, g = gpu_gradient(args, n0, m, T, wgt)
    Lt, lt, dnt, dFt = gradient(args, n0, m, T, wgt)
    assert L == pytest.approx(Lt, rel=1e-12)
    assert g["lambda"] == pytest.approx(lt, rel=TWIN_TOL, abs=TWIN_TOL * abs(Lt))
    assert rel(g["n0"], dnt) <= TWIN_TOL
    for k in range(3):
        assert rel(g["F0"][k], dFt[k]) <= TWIN_TOL, k
    lam = args[4]
    fd = (oracle_loss(with_lambda(args, lam + H), n0, m, T, wgt)
          - oracle_loss(with_lambda(args, lam - H), n0, m, T, wgt)) / (2 * H)
    assert abs(g["lambda"] - fd) <= FD_TOL * max(abs(fd), 1e-3 * abs(L)), (g["lambda"], fd)
    for p in (0, 1, n0.size // 2, n0.size - 1):
        e = np.zeros_like(n0)
        e[p] = H
        fdp = (oracle_loss(args, n0 + e, m, T, wgt) - oracle_loss(args, n0 - e, m, T, wgt)) / (2 * H)
        assert abs(g["n0"][p] - fdp) <= FD_TOL * np.max(np.abs(g["n0"])), (p, g["n0"][p], fdp)
    c.close()


@pytest.mark.parametrize("args,T", [
    pytest.param((4, 3, 61, 0.1, 1.3, 1, 1.0, 1.0, 1), 80, id="4^3-N61-three-chunks"),
    pytest.param((4, 3, 30, 0.1, 1.1, 1, 1.0, 1.0, 1), 75, id="4^3-N30-two-chunks-wrapped"),
    pytest.param((5, 2, 33, 0.08, 1.6, 2, 0.9, 0.81, 1), 70, id="5^2-N33-ragged-blocks"),
])
def test_adjoint_matches_twin_across_apply_chunks(args, T):
    """Windows longer than one chunk of the deferred-update kernel (28 slot
    rows), not a multiple of its 8-column panels, sliding so that the ring
    wraps: the gradient against the numpy twin."""
    print(f"\n[case] adjoint chunks {args} T={T}", flush=True)
    n0, m, wgt = setup(args)
    c, L, g = gpu_gradient(args, n0, m, T, wgt)
    Lt, lt, dnt, dFt = gradient(args, n0, m, T, wgt)
    assert L == pytest.approx(Lt, rel=1e-12)
    assert g["lambda"] == pytest.approx(lt, rel=TWIN_TOL, abs=TWIN_TOL * abs(Lt))
    assert rel(g["n0"], dnt) <= TWIN_TOL
    for k in range(3):
        assert rel(g["F0"][k], dFt[k]) <= TWIN_TOL, k
    c.close()


@pytest.mark.parametrize("args,T", [
    pytest.param((32, 3, 5, 0.1, 1.0, 1, 1.0, 1.0, 1), 12, id="32^3"),
    pytest.param((48, 3, 4, 0.05, 0.5, 1, 1.0, 1.0, 1), 10, id="48^3"),
    pytest.param((64, 3, 3, 0.1, 1.0, 1, 1.0, 1.0, 1), 8, id="64^3"),
])
def test_adjoint_matches_twin_on_large_lattices(args, T):
    """The adjoint self-energy planes compiled for the larger lattices (up to
    the planes' largest, 64^3), with short windows: the gradient against
    the numpy twin."""
    print(f"\n[case] adjoint large lattice {args} T={T}", flush=True)
    n0, m, wgt = setup(args)
    c, L, g = gpu_gradient(args, n0, m, T, wgt)
    Lt, lt, dnt, dFt = gradient(args, n0, m, T, wgt)
    assert L == pytest.approx(Lt, rel=1e-12)
    assert g["lambda"] == pytest.approx(lt, rel=TWIN_TOL, abs=TWIN_TOL * abs(Lt))
    assert rel(g["n0"], dnt) <= TWIN_TOL
    for k in range(3):
        assert rel(g["F0"][k], dFt[k]) <= TWIN_TOL, k
    c.close()


def test_adjoint_entries_after_partial_sweep():
    """Stop the sweep after a few reverse steps and compare every adjoint
    entry of the window with the twin's (the Hartree adjoint of the last
    reversed step is applied one step later on the device)."""
    from qftkb import KBContext
    args, T, stop = (4, 3, 5, 0.1, 1.3, 1, 1.0, 1.0, 1), 14, 11
    print(f"\n[case] partial sweep {args} T={T} stop={stop}", flush=True)
    n0, m, wgt = setup(args)
    kb = run_forward(args, n0, m, T)
    rec = {}
    Fb, rb = reverse(kb, T, wgt, m, record=rec, stop=stop)
    c = KBContext(*args[:8])
    c.init_particles(n0, m)
    c.adj_record(0, T)
    for t in range(3, T + 1):
        c.step(t)
    c.adj_debug_stop(stop)
    c.adj_gradient(wgt)
    tmin, tmax = c.window()
    assert tmax == stop - 1
    hart = (1 + 2) / 6.0 * args[4] / n0.size * rec[stop]["M2b"]
    for a in range(tmin, tmax + 1):
        for b in range(tmin, a + 1):
            want = Fb.get((a, b), np.zeros(n0.size))
            if (a, b) == (stop - 1, stop - 1):
                want = want - hart
            assert rel(c.adj_get("F", a, b), want) <= TWIN_TOL, ("F", a, b)
            if a > b + 1:
                assert rel(c.adj_get("r", a, b), rb.get((a, b), np.zeros(n0.size))) <= TWIN_TOL, ("r", a, b)
    c.close()


@pytest.mark.parametrize("C", [1, 4, 7])
def test_checkpointed_sweep_equals_taped_sweep(C):
    args, T = (6, 3, 8, 0.1, 1.4, 1, 1.0, 1.0, 1), 40
    print(f"\n[case] checkpoints every {C}: {args} T={T}", flush=True)
    n0, m, wgt = setup(args)
    c0, L0, g0 = gpu_gradient(args, n0, m, T, wgt, C=0)
    c1, L1, g1 = gpu_gradient(args, n0, m, T, wgt, C=C)
    assert L1 == pytest.approx(L0, rel=1e-13)
    assert g1["lambda"] == pytest.approx(g0["lambda"], rel=1e-11)
    assert rel(g1["n0"], g0["n0"]) <= 1e-11
    info = c1.adj_info()
    assert info["recomputed_steps"] > 0
    c0.close()
    c1.close()


@pytest.mark.parametrize("args,T", [
    pytest.param((6, 3, 8, 0.1, 1.4, 1, 1.0, 1.0, 1), 40, id="6^3-sliding"),
    pytest.param((5, 3, 7, 0.1, 1.2, 1, 1.0, 1.0, 1), 25, id="5^3-ragged-blocks"),
    pytest.param((8, 2, 6, 0.08, 1.7, 4, 0.9, 0.81, 1), 30, id="8^2-O(4)"),
])
def test_packed_tape_equals_plain_tape(args, T):
    """checkpoint_every = -1 keeps one tape value per cubic orbit; for a
    cubic-symmetric state the gradient equals the plain tape's."""
    print(f"\n[case] packed tape {args} T={T}", flush=True)
    n0, m, wgt = setup(args)
    c0, L0, g0 = gpu_gradient(args, n0, m, T, wgt, C=0)
    c1, L1, g1 = gpu_gradient(args, n0, m, T, wgt, C=-1)
    assert L1 == pytest.approx(L0, rel=1e-13)
    assert g1["lambda"] == pytest.approx(g0["lambda"], rel=1e-11)
    assert rel(g1["n0"], g0["n0"]) <= 1e-11
    assert rel(g1["F0"], g0["F0"]) <= 1e-11
    assert c1.adj_info()["device_bytes"] < c0.adj_info()["device_bytes"]
    c0.close()
    c1.close()


def test_packed_tape_rejects_a_state_without_cubic_symmetry():
    from qftkb import KBContext
    from qftkb._abi import KBError
    args = (6, 3, 8, 0.1, 1.4, 1, 1.0, 1.0)
    n0, m, _ = setup(args + (1,))
    n0 = n0.copy()
    n0[1] += 0.1  # site (1, 0, 0) only
    c = KBContext(*args)
    c.init_particles(n0, m)
    with pytest.raises(KBError):
        c.adj_record(-1, 20)
    c.adj_record(0, 20)  # the plain tape still records
    c.close()


def test_second_forward_and_gradient_repeat():
    """After a gradient the context is back at the initial state; the same
    forward run and sweep give the same numbers."""
    from qftkb import KBContext
    args, T = (6, 3, 8, 0.1, 1.4, 1, 1.0, 1.0, 1), 30
    print(f"\n[case] repeat {args} T={T}", flush=True)
    n0, m, wgt = setup(args)
    c = KBContext(*args[:8])
    c.init_particles(n0, m)
    c.adj_record(5, T)
    out = []
    for _ in range(2):
        for t in range(3, T + 1):
            c.step(t)
        out.append(c.adj_gradient(wgt))
    assert out[1]["lambda"] == pytest.approx(out[0]["lambda"], rel=1e-12)
    assert rel(out[1]["n0"], out[0]["n0"]) <= 1e-12
    c.close()


def test_adjoint_vs_oracle_fd_at_16cubed():
    """A larger lattice with a sliding window: GPU adjoint vs oracle FD."""
    args, T = (16, 3, 12, 0.1, 1.0, 1, 1.0, 1.0, 1), 40
    print(f"\n[case] 16^3 N_mem=12 T={T}", flush=True)
    n0, m, wgt = setup(args)
    c, L, g = gpu_gradient(args, n0, m, T, wgt, C=10)
    lam = args[4]
    fd = (oracle_loss(with_lambda(args, lam + H), n0, m, T, wgt)
          - oracle_loss(with_lambda(args, lam - H), n0, m, T, wgt)) / (2 * H)
    assert abs(g["lambda"] - fd) <= FD_TOL * abs(fd), (g["lambda"], fd)
    # one radial shell of occupations moved together
    k2 = kL2_table(16, 3)
    shell = (k2 > 1.0) & (k2 <= 1.5)
    e = np.where(shell, H, 0.0)
    fds = (oracle_loss(args, n0 + e, m, T, wgt) - oracle_loss(args, n0 - e, m, T, wgt)) / (2 * H)
    ad = float(g["n0"][shell].sum())
    assert abs(ad - fds) <= FD_TOL * abs(fds), (ad, fds)
    c.close()


def test_differentiation_workload_vs_oracle_fd_fixture():
    """24^3, N_mem = 200, 500 steps, checkpoints every 50: the adjoint
    gradient against the oracle's central finite differences committed in
    tests/golden/adjoint_24c_n200_fd.json (made by tests/golden/make_adjoint_fd.py)."""
    import json
    import os
    from qftkb import KBContext
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "adjoint_24c_n200_fd.json")))
    print("\n[case] differentiation workload 24^3 N_mem=200 500 steps, C=50", flush=True)
    c = KBContext(gold["Nx"], gold["sdim"], gold["N_mem"], gold["dt"], gold["Lambda"], gold["ONgroup"],
                  gold["Mass"], gold["Mass"] ** 2)
    k2 = c.kL2()
    n0 = thermal(k2, m=gold["Mass"])
    wgt = np.exp(-k2)
    T = gold["T_final"]
    c.init_particles(n0, gold["Mass"])
    c.adj_record(50, T)
    for t in range(3, T + 1):
        c.step(t)
    L = float((wgt * c.measure(T)["n"]).sum())
    assert L == pytest.approx(0.5 * (gold["L_plus"] + gold["L_minus"]), rel=1e-9)
    g = c.adj_gradient(wgt)
    fd = gold["dL_dlambda_fd"]
    assert abs(g["lambda"] - fd) <= FD_TOL * abs(fd), (g["lambda"], fd)
    fdb = gold.get("dL_dn0_bins_fd")
    if fdb and all(v is not None for v in fdb):
        edges = np.linspace(0.0, np.sqrt(12.0), 17)
        b = np.minimum(np.searchsorted(edges, np.sqrt(k2), side="right") - 1, 15)
        ad = np.array([g["n0"][b == i].sum() for i in range(16)])
        assert np.max(np.abs(ad - np.array(fdb))) <= FD_TOL * np.max(np.abs(fdb))
    c.close()


@pytest.mark.parametrize("C", [0, 25, -1], ids=["taped", "checkpoints-25", "packed-tape"])
def test_largest_lattice_gradient_vs_oracle_fd_fixture(C):
    """The 48^3 lattice with the box state and coupling of the 4000-step
    N_mem = 300 differentiation (n0 = 4, Q = 1, lambda = 0.5, dt = 0.05) at
    a window the oracle can run: N_mem = 20, 100 steps.  The gradient, taped,
    checkpointed or on the orbit-packed tape, against the oracle's central
    finite differences with Richardson extrapolation, R(h) = (4 D(h) - D(2h))/3
    at h = 5e-4 (h = 1e-3 for the radial bins), committed in
    tests/golden/adjoint_48c_n20_fd.json (tests/golden/make_adjoint_fd_48.py),
    to 1e-6.  (D(1e-5) alone is limited to about 1e-6 by the rounding of L.)"""
    import json
    import os
    from qftkb import KBContext, lattice
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "adjoint_48c_n20_fd.json")))
    print(f"\n[case] 48^3 N_mem=20 100 steps, recording mode {C}", flush=True)
    c = KBContext(gold["Nx"], gold["sdim"], gold["N_mem"], gold["dt"], gold["Lambda"], gold["ONgroup"],
                  gold["Mass"], gold["Mass"] ** 2)
    k2 = c.kL2()
    n0 = lattice.n_box(lattice.kL2_table(gold["Nx"], gold["sdim"]), 4.0, 1.0)
    wgt = np.exp(-k2)
    T = gold["T_final"]
    c.init_particles(n0, gold["Mass"])
    c.adj_record(C, T)
    for t in range(3, T + 1):
        c.step(t)
    L = float((wgt * c.measure(T)["n"]).sum())
    assert L == pytest.approx(0.5 * (gold["L_plus_1e-05"] + gold["L_minus_1e-05"]), rel=1e-9)
    g = c.adj_gradient(wgt)
    assert gold["richardson_spread"] <= 2e-7  # R(5e-4) and R(1e-3) agree: the fixture is good to ~1e-7
    fd = gold["dL_dlambda_richardson"]
    assert abs(g["lambda"] - fd) <= FD_TOL * abs(fd), (g["lambda"], fd)
    fdb = gold.get("dL_dn0_bins_fd")
    if fdb and all(v is not None for v in fdb):
        b = lattice.radial_bins(k2, gold["sdim"], 16)
        ad = np.array([g["n0"][b == i].sum() for i in range(16)])
        assert np.max(np.abs(ad - np.array(fdb))) <= FD_TOL * np.max(np.abs(fdb))
    c.close()


@pytest.mark.parametrize("args,T,C", [
    pytest.param((72, 2, 5, 0.1, 1.2, 2, 1.0, 0.9, 1), 12, 0, id="72^2-taped"),
    pytest.param((67, 2, 4, 0.1, 1.5, 1, 1.0, 1.0, 4), 10, 3, id="67^2-classical-checkpointed"),
    pytest.param((66, 3, 3, 0.1, 1.0, 1, 1.0, 1.0, 1), 7, 0, id="66^3"),
])
def test_adjoint_on_pencil_passes_matches_twin(args, T, C):
    """2D and 3D lattices past the plane kernels: the recompute keeps its
    transforms and the adjoint sunset runs as pencil passes; against the
    reverse-mode twin at 1e-10."""
    print(f"\n[case] adjoint on pencil passes {args} T={T} C={C}", flush=True)
    n0, m, wgt = setup(args)
    c, L, g = gpu_gradient(args, n0, m, T, wgt, C=C)
    Lt, lt, dnt, dFt = gradient(args, n0, m, T, wgt)
    assert L == pytest.approx(Lt, rel=1e-12)
    assert g["lambda"] == pytest.approx(lt, rel=TWIN_TOL, abs=TWIN_TOL * abs(Lt))
    assert rel(g["n0"], dnt) <= TWIN_TOL
    for k in range(3):
        assert rel(g["F0"][k], dFt[k]) <= TWIN_TOL, k
    c.close()
