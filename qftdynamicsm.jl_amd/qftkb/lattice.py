"""Lattice tables and initial occupations for the TwoPIScalar model.

Restates, per lattice site, what the reference computes per FFTWhelper
level.  Site order is column-major (x fastest), i.e. the memory order of a
Julia Array{Float64,sdim}.

lev2 (the reference's FFTWhelper level value, SimDataStruct.jl:673-679) is
the lattice momentum squared sum_i 4 sin^2(pi n_i / Nx).  FFTWhelper itself
is not vendored with the reference; the level grouping used for TopHat
windows and for per-level presentation below groups sites into shells of
equal lev2 sorted ascending.
"""
from __future__ import annotations

import math

import numpy as np


def kL2_table(Nx: int, sdim: int) -> np.ndarray:
    n = np.arange(Nx)
    s = np.sin(np.pi * n / Nx)
    k1 = 4.0 * s * s
    if sdim == 1:
        return k1.copy()
    if sdim == 2:
        return (k1[None, :] + k1[:, None]).reshape(-1)
    if sdim == 3:
        return (k1[None, None, :] + k1[None, :, None] + k1[:, None, None]).reshape(-1)
    raise ValueError("sdim must be 1, 2 or 3")


def levels(kL2: np.ndarray, decimals: int = 12):
    """Group sites into shells of equal lev2 (ascending).  Returns
    (lev2 values, list of site-index arrays); the first index of each shell
    is its representative site (Measurement.jl:25 takes ind[1])."""
    key = np.round(kL2, decimals)
    vals, inv = np.unique(key, return_inverse=True)
    groups = [np.nonzero(inv == i)[0] for i in range(len(vals))]
    return np.array([kL2[g[0]] for g in groups]), groups


def hartreeprefactor(Lambda: float, ONgroup: int, pexp: str = "LambdaNLO") -> float:
    """Initialization.jl:245-247 (every lambda expansion, LambdaNLOclassical
    included, and NinverseNLO: lambda (N+2)/(6N))."""
    if pexp == "NinverseLO":
        return Lambda * (1 / 6)
    return Lambda * (ONgroup + 2) / (6 * ONgroup)


# ---- occupations (getparticlenr, Initialization.jl:133-203) -----------------
def n_particle(kL2, n):
    """TwoPIScalarParticle: constant n on every level (:133-135)."""
    return np.full(kL2.shape, float(n))


def n_thermal(kL2, Mass, T):
    """TwoPIScalarThermal: 1/(exp(sqrt(lev2+M^2)/(T M)) - 1) (:137-140)."""
    return 1.0 / (np.exp(np.sqrt(kL2 + Mass ** 2) / (T * Mass)) - 1.0)


def _tophat(kL2, k2min, k2max, eta):
    lev, groups = levels(kL2)
    imin = int(np.argmin(np.abs(lev - k2min)))
    imax = int(np.argmin(np.abs(lev - k2max)))
    n = np.zeros(kL2.shape)
    for i in range(imin, imax + 1):
        n[groups[i]] = eta
    return n


def n_tophat(kL2, variant, sdim, Mass):
    """TwoPIScalarTopHatT1/T2/T3 on the lattice (:142-191).  The level
    window follows the reference's findmin over level values; the level
    ordering is that of `levels` (parity unpinned: FFTWhelper is absent)."""
    if variant == 1:
        return _tophat(kL2, 0.68 * sdim * Mass ** 2, 2.04 * sdim * Mass ** 2, 2.0)
    if variant == 2:
        return _tophat(kL2, 0.0 * sdim * Mass ** 2, 1.9 * sdim * Mass ** 2, 1.85)
    if variant == 3:
        return _tophat(kL2, 2.04 * sdim * Mass ** 2, 2.72 * sdim * Mass ** 2, 1.6)
    raise ValueError("TopHat variant must be 1, 2 or 3")


BOX_EDGE_RTOL = 1e-12


def n_box(kL2, n0, Q):
    """Over-occupied box n(p) = n0 * Theta(Q - |p|) with |p| = sqrt(lev2).

    The reference defines Box only for its continuum discretisation
    (Initialization.jl:194-203); this lattice form is an extension.  Sites
    with lev2 <= Q^2 (1 + 1e-12) are inside, so a site exactly on the edge
    (e.g. 4 sin^2(pi 8/48) = 1 at Nx = 48, Q = 1) counts as occupied."""
    return np.where(kL2 <= Q * Q * (1.0 + BOX_EDGE_RTOL), float(n0), 0.0)


def gauss_rows(V, dt, xi, eta, sig):
    """F(1,1), F(2,1), F(2,2) of the Gauss initial state (Initialization.jl:
    95-112)."""
    F11 = np.full(V, xi ** 2)
    F21 = xi * eta * dt + F11
    F22 = (eta ** 2 + sig ** 2 / (4 * xi ** 2)) * dt ** 2 - F11 + 2 * F21
    return F11, F21, F22


def gauss_energy(xi, eta, sig, Mass2, Lambda, ONgroup):
    """Conserved energy printed by the Gauss init (Initialization.jl:109)."""
    return (0.5 * eta ** 2 + sig ** 2 / (8 * xi ** 2) + 0.5 * Mass2 * xi ** 2
            + (1 / 24) * Lambda * (ONgroup + 2) / ONgroup * xi ** 4)


def gap_mass(kL2, n, Mass2, H, x0=0.5, maxiter=100):
    """getGapMass (Initialization.jl:250-266): Newton root of
    sum_p (n_p + 1/2)/sqrt(lev2 + x^2) * ivol * H + Mass2 - x^2 = 0
    from x0 = 0.5, with the analytic derivative in place of ForwardDiff."""
    ivol = 1.0 / kL2.size
    x = float(x0)
    for _ in range(maxiter):
        w = np.sqrt(kL2 + x * x)
        f = np.sum((n + 0.5) / w) * ivol * H + Mass2 - x * x
        df = -np.sum((n + 0.5) * x / w ** 3) * ivol * H - 2 * x
        dx = f / df
        x -= dx
        if abs(dx) <= 4 * np.finfo(float).eps * max(abs(x), 1.0):
            break
    return x


def counterterm(kL2, Mass, H):
    """getct (ModelExpansionRenormalization.jl:38-47):
    -(sum_p 1/(2 sqrt(lev2 + Mass^2))) * H * ivol."""
    ct = -np.sum(0.5 / np.sqrt(kL2 + Mass ** 2))
    return ct * H * (1.0 / kL2.size)


def radial_bins(kL2, sdim, nbins=16):
    """Radial |p| = sqrt(lev2) bins over [0, sqrt(4 sdim)] (closed at the
    top): returns the bin index of every site."""
    edge = math.sqrt(4.0 * sdim)
    b = np.floor(np.sqrt(kL2) / edge * nbins).astype(np.int64)
    return np.clip(b, 0, nbins - 1)
