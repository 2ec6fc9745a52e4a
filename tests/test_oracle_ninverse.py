"""Pins of the oracle's 1/N expansion (NinverseLO, NinverseNLO).

The reference's only active test of this path compares its fast
calcIrs!/calcIFs! with calcIrssimple!/calcIFssimple! on random, lattice-
symmetric F and rho after ten steps (Tests/TwoPIScalar_Tests_NinverseNLO_
checkfastimplementation.jl); that comparison is repeated here on the
oracle's restatements of both.  The restatement is further pinned by

  * a second formulation: the momentum-space recursion of the reference's
    CPUred2 back-end, restated in numpy (oracle/kb_numpy.py), evolved beside
    the C oracle's position-space CPUfull restatement;
  * the small-coupling limit: to leading order in lambda the resummed
    kernels reduce to single bubbles and the self-energies to the LambdaNLO
    sunset times N/(N+2) (the large-N part of its (N+2)/N^2 prefactor), with
    corrections of relative order lambda;
  * a closed form: with time-independent bubbles the Ir recursion is a
    geometric series;
  * NinverseLO differs from LambdaLO only by the Hartree prefactor lambda/6
    (Initialization.jl:246), so it equals LambdaLO at coupling
    lambda N/(N+2).
"""
import numpy as np
import pytest

from oracle.kb_numpy import NumpyKB
from oracle.oracle import Oracle
from qftkb import lattice


def _tophat(o, sdim):
    return lattice.n_tophat(o.kL2(), 1, sdim, 1.0)


@pytest.mark.parametrize("Nx,sdim", [(8, 1), (4, 2), (4, 3)])
def test_fast_equals_simple_on_random_symmetric_data(Nx, sdim):
    """TwoPIScalar_Tests_NinverseNLO_checkfastimplementation.jl: Mass 1,
    dt 0.01, lambda 10, N 8, TopHatT1, full memory; evolve to step 10, fill
    F(t, tp) and r(t, tp) for t in 3..10 with random values constant on each
    momentum shell (createrandomF), expand, compare Ir and IF at tmone = 10
    elementwise."""
    o = Oracle(Nx, sdim, 14, 0.01, 10.0, 8, 1.0, 1.0, pexp=3, nthreads=1)
    o.init_particles(_tophat(o, sdim), 1.0)
    for t in range(3, 11):
        o.evolve(t)
    o.expand()
    _, shells = lattice.levels(o.kL2())
    rng = np.random.default_rng(Nx + 10 * sdim)

    def createrandomF():
        a = np.empty(o.V)
        for g in shells:
            a[g] = rng.random() - 0.5
        return a

    tmin, _ = o.window()
    for t in range(3, 11):
        for tp in range(tmin, t + 1):
            o.set("F", t, tp, createrandomF())
            o.set("r", t, tp, createrandomF())
    o.ninv_resum(10, simple=False)
    fast = {(w, tp): o.ninv(w, tp) for w in "Fr" for tp in range(tmin, 11)}
    o.ninv_resum(10, simple=True)
    for (w, tp), a in fast.items():
        b = o.ninv(w, tp)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(b).max()), (w, tp)
    assert np.abs(fast[("F", 10)]).max() > 1e-3 and np.abs(fast[("r", 1)]).max() > 1e-3


@pytest.mark.parametrize("Nx,sdim,Nmem", [(8, 1, 6), (4, 2, 5), (4, 3, 6), (3, 3, 20)])
def test_position_space_equals_momentum_space_recursion(Nx, sdim, Nmem):
    """C oracle (CPUfull: a transform pair per term of the time sums) against
    the numpy restatement of the CPUred2 momentum-space recursion, over a
    window that fills and slides."""
    args = (Nx, sdim, Nmem, 0.05, 10.0, 8, 1.0, 1.0)
    o = Oracle(*args, pexp=3, nthreads=1)
    nk = NumpyKB(*args, pexp=3)
    n = _tophat(o, sdim)
    o.init_particles(n, 1.0)
    nk.init_particles(n, 1.0)
    for t in range(3, 16):
        o.evolve(t)
        nk.evolve(t)
        tmin, _ = o.window()
        for tp in range(tmin, t):
            for w, d in (("r", nk.Sigr), ("F", nk.SigF)):
                ref = d[tp]
                assert np.abs(o.sigma(w, tp) - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-30), (t, tp, w)
        for tp in range(tmin, t + 1):
            ref = nk.getF(t, tp)
            assert np.abs(o.get("F", t, tp) - ref).max() <= 1e-12 * np.abs(ref).max()
            ref = nk.getr(t, tp)
            assert np.abs(o.get("r", t, tp) - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-30)
    # the resummed kernels themselves: Ir, IF of the last step in momentum
    # space are the transforms of the oracle's position-space Irx, IFx
    T = o.window()[1] - 1
    tmin = o.window()[0]
    shape = (Nx,) * sdim
    for tp in range(tmin, T + 1):
        for w, d in (("r", nk.Irk), ("F", nk.IFk)):
            x = o.ninv(w, tp).real
            k = np.fft.fftn(x.reshape(shape[::-1])).real.reshape(-1)
            assert np.abs(k - d[tp]).max() <= 1e-12 * max(np.abs(d[tp]).max(), 1e-30), (w, tp)


def _sigmas(pexp, lam, ON):
    o = Oracle(4, 3, 8, 0.05, lam, ON, 1.0, 1.0, pexp=pexp, nthreads=1)
    o.init_particles(_tophat(o, 3), 1.0)
    for t in range(3, 12):
        o.evolve(t)
    tmin, tmax = o.window()
    return [(o.sigma("r", tp), o.sigma("F", tp)) for tp in range(tmin, tmax - 1)]


def test_small_coupling_limit_is_large_N_sunset():
    ON = 4
    dev = []
    for lam in (1e-3, 1e-4):
        worst = 0.0
        for (r1, F1), (r3, F3) in zip(_sigmas(1, lam, ON), _sigmas(3, lam, ON)):
            for x1, x3 in ((r1, r3), (F1, F3)):
                x1 = x1 * ON / (ON + 2)
                worst = max(worst, np.abs(x3 - x1).max() / np.abs(x1).max())
        dev.append(worst)
    # leading order agrees; the deviation is first order in lambda
    assert dev[0] < 1e-4
    assert 8.0 < dev[0] / dev[1] < 12.0


def test_Ir_geometric_series_for_constant_bubbles():
    """F(z,tp) = f and rho(z,tp) = g sgn(z - tp) at every momentum make every
    bubble Fr(z,tp) = f g, so Ir(tmone, tp) = (lambda/3) f g q^(tmone-1-tp)
    at x = 0 (zero elsewhere) with q = 1 - dt (lambda/3) f g."""
    lam, dt, f, g = 7.0, 0.05, 0.8, 0.6
    o = Oracle(4, 3, 12, dt, lam, 8, 1.0, 1.0, pexp=3, nthreads=1)
    o.init_particles(np.zeros(o.V), 1.0)
    for t in range(3, 11):
        o.evolve(t)
    tmin, tmax = o.window()
    for a in range(tmin, tmax + 1):
        for b in range(tmin, a + 1):
            o.set("F", a, b, np.full(o.V, f))
            o.set("r", a, b, np.full(o.V, g if a > b else 0.0))
    T = tmax
    o.ninv_resum(T)
    a3 = lam / 3.0
    q = 1.0 - dt * a3 * f * g
    for tp in range(tmin, T):
        x = o.ninv("r", tp)
        want = a3 * f * g * q ** (T - 1 - tp)
        assert abs(x[0].real - want) <= 1e-13 * abs(want), tp
        assert np.abs(x[1:]).max() <= 1e-13 * abs(want)
    assert np.abs(o.ninv("r", T)).max() == 0.0


def test_ninverse_lo_is_lambda_lo_at_rescaled_coupling():
    lam, ON = 3.0, 5
    a = Oracle(4, 3, 6, 0.05, lam, ON, 1.0, 1.0, pexp=2, nthreads=1)
    b = Oracle(4, 3, 6, 0.05, lam * ON / (ON + 2), ON, 1.0, 1.0, pexp=0, nthreads=1)
    n = _tophat(a, 3)
    a.init_particles(n, 1.0)
    b.init_particles(n, 1.0)
    assert a.hartreemass2 == pytest.approx(b.hartreemass2, rel=1e-14)
    for t in range(3, 12):
        a.evolve(t)
        b.evolve(t)
    tmin, tmax = a.window()
    for tp in range(tmin, tmax + 1):
        assert np.allclose(a.get("F", tmax, tp), b.get("F", tmax, tp), rtol=1e-13, atol=0)
    ma, mb = a.measure(tmax), b.measure(tmax)
    assert ma["E"] == pytest.approx(mb["E"], rel=1e-13)
    # and the prefactor is not the lambda-expansion one
    c = Oracle(4, 3, 6, 0.05, lam, ON, 1.0, 1.0, pexp=0, nthreads=1)
    c.init_particles(n, 1.0)
    assert abs(c.hartreemass2 - a.hartreemass2) > 1e-3
