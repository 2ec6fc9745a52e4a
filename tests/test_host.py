"""Host-side mirror of the reference's parameter / set-up API (no GPU).

Expected values below are derived by hand from the cited reference lines.
"""
import math

import numpy as np
import pytest

from qftkb import lattice
from qftkb.problem import (TwoPIScalarfile, _make_problem, check_parameters, getParameterstring, splitter,
                           createSimSetup)


def base_params(**kw):
    p = {"Mass": 1.0, "Nx": 8, "sdim": 3, "dt": 0.1, "Nsteps": 300, "NstepsinMemory": 0, "Nmeas": 0,
         "Mod": "phi4", "Lambda": 1.0, "ONgroup": 1, "init": "Thermal", "xi": 0.4, "eta": 5.0, "sig": 1.0,
         "T": 2, "n": 2, "n0": 2.0, "Q": 1.0, "Reno": "RNone", "Pexp": "LambdaNLO", "Num": "GPUfull",
         "Nchunks": 0, "tag": "testing", "ov": True}
    p.update(kw)
    return p


def test_check_parameters_defaults():
    """DefAbstractSimSetup.jl:21-31."""
    p = base_params()
    check_parameters(p)
    assert p["NstepsinMemory"] == 302 and p["Nmeas"] == 300
    p = base_params(NstepsinMemory=400, Nmeas=7)
    check_parameters(p)
    assert p["NstepsinMemory"] == 400 and p["Nmeas"] == 7


def test_measurement_grid():
    """NstepsbetweenNmeas = floor(Nsteps/Nmeas), Nsteps rounded down to a
    multiple (DefAbstractSimSetup.jl:105-107)."""
    p = base_params(Nsteps=2000, Nmeas=300, NstepsinMemory=400)
    ss = createSimSetup(TwoPIScalarfile(), p)
    assert ss.NstepsbetweenNmeas == 6 and ss.Nsteps == 1800 and ss.Nmeas == 300
    # measured times t satisfy (t-2) % 6 == 0 and land at index (t-2)/6 (Measurement.jl:227)
    ts = [t for t in range(3, ss.Nsteps + 3) if (t - 2) % ss.NstepsbetweenNmeas == 0]
    assert [(t - 2) // 6 for t in ts] == list(range(1, 301))


def test_splitter():
    """Int.(ceil.(range(0, n, length=k+1))) chunks (DefAbstractSimSetup.jl:7-10)."""
    chunks = splitter(5, 10, 3)
    # range(0,10,length=4) = 0, 3.33, 6.67, 10 -> ceil 0, 4, 7, 10
    assert [list(c) for c in chunks] == [[5, 6, 7, 8], [9, 10, 11], [12, 13, 14]]
    assert sum(sum(c) for c in splitter(1, 399, 28)) == sum(range(1, 400))


def test_parameterstring():
    """TwoPIScalar_Modelfile.jl:15-44."""
    p = base_params(NstepsinMemory=400, Nsteps=2000, Nx=32, Mass=1.0, dt=0.1, T=2)
    s = getParameterstring(TwoPIScalarfile(), p)
    assert s == ("phi4LambdaNLO_L100_ON1_RNone_T20_Nx32_sdim3_Nsteps2000_dt100_M100_NstepsinMemory400"
                 "_NumGPUfull_testing")
    p = base_params(init="Pnr", n=2, Lambda=6, ONgroup=8, Mass=0.7, dt=0.05, tag="")
    assert getParameterstring(TwoPIScalarfile(), p).startswith("phi4LambdaNLO_L600_ON8_RNone_n20_Nx8")
    with pytest.raises(ValueError):  # Int64(0.123*1000) is an InexactError in Julia
        getParameterstring(TwoPIScalarfile(), base_params(dt=0.1234))


def test_problem_factory_and_errors():
    prob = _make_problem(TwoPIScalarfile(), base_params())
    assert prob.disc.vol == 512 and prob.disc.ivol == 1 / 512
    assert prob.model.Mass2 == 1.0 and prob.simsetup.NstepsinMemory == 302
    with pytest.raises(ValueError):
        _make_problem(TwoPIScalarfile(), base_params(Num="CPUfull"))
    # the two device numerics: every site, or one value per cubic orbit
    for num in ("GPUfull", "GPUorbits"):
        assert _make_problem(TwoPIScalarfile(), base_params(Num=num)).num == num
    with pytest.raises(ValueError):
        _make_problem(TwoPIScalarfile(), base_params(Num="CPUred"))
    with pytest.raises(NotImplementedError):
        _make_problem(TwoPIScalarfile(), base_params(init="Quench"))
    with pytest.raises(ValueError):
        _make_problem(TwoPIScalarfile(), base_params(Pexp="LambdaNNLO"))
    for pexp in ("NinverseLO", "NinverseNLO"):
        assert _make_problem(TwoPIScalarfile(), base_params(Pexp=pexp)).pexp == pexp
    # Hartreeprefactor per expansion (Initialization.jl:245-247) in the RMass counterterm
    lo = _make_problem(TwoPIScalarfile(), base_params(Pexp="NinverseLO", Reno="RMass", Lambda=3.0, ONgroup=4, Nx=4))
    nlo = _make_problem(TwoPIScalarfile(), base_params(Pexp="NinverseNLO", Reno="RMass", Lambda=3.0, ONgroup=4, Nx=4))
    assert (lo.model.Mass2 - 1.0) / (nlo.model.Mass2 - 1.0) == pytest.approx((1 / 6) / (6 / 24), rel=1e-12)


def test_counterterm_rmass():
    """getct: ct = -sum_p 1/(2 omega_p) * H * ivol (ModelExpansionRenormalization.jl:38-47)."""
    prob = _make_problem(TwoPIScalarfile(), base_params(Reno="RMass", Lambda=3.0, ONgroup=4, Nx=4))
    k2 = lattice.kL2_table(4, 3)
    H = 3.0 * 6 / 24
    ct = -np.sum(0.5 / np.sqrt(k2 + 1.0)) * H / 64
    assert prob.model.Mass2 == pytest.approx(1.0 + ct, rel=1e-15)
    tach = _make_problem(TwoPIScalarfile(), base_params(Reno="RMass", Mod="phi4tachyonic", Nx=4))
    assert tach.model.Mass2 == pytest.approx(-1.0 + lattice.counterterm(k2, 1.0, 0.5), rel=1e-15)


def test_gap_mass_solves_gap_equation():
    """getGapMass (Initialization.jl:250-266): residual of the gap equation."""
    k2 = lattice.kL2_table(8, 3)
    n = np.full(k2.size, 2.0)
    H = lattice.hartreeprefactor(6.0, 8)
    m = lattice.gap_mass(k2, n, 0.49, H)
    res = np.sum((n + 0.5) / np.sqrt(k2 + m * m)) / k2.size * H + 0.49 - m * m
    assert abs(res) < 1e-13 and m > 0.7


def test_occupations():
    k2 = lattice.kL2_table(16, 3)
    th = lattice.n_thermal(k2, 1.0, 2.0)
    assert th[0] == pytest.approx(1 / (math.exp(0.5) - 1))
    box = lattice.n_box(k2, 2.0, 1.0)
    assert set(np.unique(box)) == {0.0, 2.0}
    assert box[0] == 2.0 and box[np.argmax(k2)] == 0.0
    # the Nx = 48 site with lev2 = 4 sin^2(pi/6) = 1 sits on the edge and is occupied
    k48 = lattice.kL2_table(48, 1)
    assert lattice.n_box(k48, 4.0, 1.0)[8] == 4.0 and lattice.n_box(k48, 4.0, 1.0)[9] == 0.0
    lev, groups = lattice.levels(k2)
    assert np.all(np.diff(lev) > 0) and sum(len(g) for g in groups) == k2.size
    t1 = lattice.n_tophat(k2, 1, 3, 1.0)
    occ = k2[t1 > 0]
    assert occ.min() >= lev[np.argmin(np.abs(lev - 0.68 * 3))] - 1e-12


def test_kL2_matches_oracle():
    from oracle.oracle import Oracle
    for Nx, sdim in [(8, 3), (5, 2), (12, 1)]:
        o = Oracle(Nx, sdim, 3, 0.1, 1.0, 1, 1.0, 1.0, nthreads=1)
        np.testing.assert_allclose(o.kL2(), lattice.kL2_table(Nx, sdim), rtol=4e-16, atol=4e-16)


def test_radial_bins():
    k2 = lattice.kL2_table(24, 3)
    b = lattice.radial_bins(k2, 3, 16)
    assert b.min() == 0 and b.max() == 15 and b[0] == 0
