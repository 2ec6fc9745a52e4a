"""The mathematics behind the orbit products (CPU, no library): on functions
constant on the orbits of the lattice's cubic group, the unnormalised cosine
transform C X(k) = sum_x X(x) cos(k.x) is the orbit matrix
Mo[o][b] = sum_{x in orbit b} cos(k_o . x), it equals numpy's FFT of the
expanded lattice at the representatives, and Mo Mo = V I (the transform of
even functions is its own inverse up to V).  The sunset built from two such
products equals the FFT form used by the reference (TwoPIScalar_Evolution.jl:
25-31)."""
import itertools

import numpy as np
import pytest


def orbits(Nx, sdim):
    V = Nx ** sdim
    orb = -np.ones(V, dtype=int)
    reps = []
    for s in range(V):
        if orb[s] >= 0:
            continue
        c = [(s // Nx ** d) % Nx for d in range(sdim)]
        o = len(reps)
        reps.append(s)
        for perm in itertools.permutations(range(sdim)):
            for flips in itertools.product([0, 1], repeat=sdim):
                cc = [c[perm[d]] for d in range(sdim)]
                cc = [(Nx - x) % Nx if f else x for x, f in zip(cc, flips)]
                orb[sum(x * Nx ** d for d, x in enumerate(cc))] = o
    return orb, np.array(reps)


def orbit_matrix(Nx, sdim, orb, reps):
    V = Nx ** sdim
    xc = np.array([[(x // Nx ** d) % Nx for d in range(sdim)] for x in range(V)])
    ctab = np.cos(2 * np.pi * np.arange(Nx) / Nx)
    Mo = np.zeros((len(reps), len(reps)))
    for o, r in enumerate(reps):
        m = (xc @ xc[r]) % Nx
        np.add.at(Mo[o], orb, ctab[m])
    return Mo


@pytest.mark.parametrize("Nx,sdim,norb", [(6, 3, 20), (7, 3, 20), (8, 2, 15), (10, 1, 6), (5, 2, 6)])
def test_orbit_matrix_is_the_cosine_transform(Nx, sdim, norb):
    rng = np.random.default_rng(Nx * 10 + sdim)
    orb, reps = orbits(Nx, sdim)
    assert len(reps) == norb
    V = Nx ** sdim
    Mo = orbit_matrix(Nx, sdim, orb, reps)
    for _ in range(3):
        xo = rng.standard_normal(len(reps))
        lat = xo[orb].reshape([Nx] * sdim, order="F")
        ft = np.fft.fftn(lat).reshape(-1, order="F")
        np.testing.assert_allclose(ft.imag, 0.0, atol=1e-10 * np.abs(ft).max())
        np.testing.assert_allclose(Mo @ xo, ft.real[reps], rtol=1e-12, atol=1e-12 * np.abs(ft).max())
    np.testing.assert_allclose(Mo @ Mo, V * np.eye(len(reps)), atol=1e-9 * V)


def test_sunset_from_two_orbit_products_equals_fft_form():
    Nx, sdim = 6, 3
    rng = np.random.default_rng(7)
    orb, reps = orbits(Nx, sdim)
    Mo = orbit_matrix(Nx, sdim, orb, reps)
    F, r = rng.standard_normal(len(reps)), rng.standard_normal(len(reps))
    # FFT form: Fx = C F, rx = C rho, Sigma = C[Fx (Fx^2 - 3/4 rx^2)], C[rx (Fx^2 - rx^2/12)]
    lat = lambda v: v[orb].reshape([Nx] * sdim, order="F")
    Fx, rx = np.fft.fftn(lat(F)).real, np.fft.fftn(lat(r)).real
    SF = np.fft.fftn(Fx * (Fx ** 2 - 0.75 * rx ** 2)).real.reshape(-1, order="F")[reps]
    SR = np.fft.fftn(rx * (Fx ** 2 - rx ** 2 / 12)).real.reshape(-1, order="F")[reps]
    fo, ro = Mo @ F, Mo @ r
    np.testing.assert_allclose(Mo @ (fo * (fo ** 2 - 0.75 * ro ** 2)), SF, rtol=1e-11, atol=1e-11 * np.abs(SF).max())
    np.testing.assert_allclose(Mo @ (ro * (fo ** 2 - ro ** 2 / 12)), SR, rtol=1e-11, atol=1e-11 * np.abs(SR).max())
