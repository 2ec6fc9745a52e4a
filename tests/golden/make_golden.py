"""Writes the golden fixtures under tests/golden/ from closed forms.

Nothing here runs the reference (it is Julia and cannot run in this image);
every value below follows from the mathematics of the KB step:

free_theory.npz
    lambda = 0: the sunset and the Hartree shift vanish and evolve! reduces
    to the leapfrog recurrence x_t = 2c x_{t-1} - x_{t-2}, c = 1 - Omega/2,
    Omega = (kL2 + Mass^2) dt^2 (TwoPIScalar_Evolution.jl:890-921).  Its
    exact solution from the initial rows (Initialization.jl:8-29) is
        F(t,t') = K(t)^T M K(t'),  M = [[F11, F21], [F21, F22]],
        K_1(t) = -U_{t-3}(c),  K_2(t) = U_{t-2}(c),
        rho(t,t') = dt U_{t-t'-1}(c),
    with U_n the Chebyshev polynomials of the second kind (U_{-1} = 0,
    U_{-2} = -1).  Evaluated in 50-digit arithmetic with mpmath from the
    float64 inputs.

gauss_energy.json
    Conserved energy of the Gauss initial state at t = 2 printed by the
    reference (Initialization.jl:109), 0.5 eta^2 + sig^2/(8 xi^2)
    + 0.5 m^2 xi^2 + lambda (N+2) xi^4 / (24 N), for a few parameter sets.

Run: python tests/golden/make_golden.py
"""
import json
import os

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mpmath.mp.dps = 50


def cheb_u(n, c):
    """U_n(c) for n >= -2 by the three-term recurrence in mp precision."""
    if n == -2:
        return mpmath.mpf(-1)
    if n == -1:
        return mpmath.mpf(0)
    u0, u1 = mpmath.mpf(0), mpmath.mpf(1)  # U_{-1}, U_0
    for _ in range(n):
        u0, u1 = u1, 2 * c * u1 - u0
    return u1


def free_theory():
    # a small lattice: 4^3 thermal occupations, dt = 0.1, Mass = 1, N_mem = 7,
    # evolved to t = 40 (the window wraps several times)
    Nx, sdim, dt, Mass, T, Nmem, tend = 4, 3, 0.1, 1.0, 2.0, 7, 40
    n = np.arange(Nx)
    k1 = 4.0 * np.sin(np.pi * n / Nx) ** 2
    kL2 = (k1[None, None, :] + k1[None, :, None] + k1[:, None, None]).reshape(-1)
    om = np.sqrt(kL2 + Mass * Mass)
    occ = 1.0 / (np.exp(om / (T * Mass)) - 1.0)
    F11 = (1.0 / om) * (occ + 0.5)
    F21 = (1.0 / om) * (occ + 0.5)
    F22 = om * (occ + 0.5) * (dt * dt) + (1.0 / om) * (occ + 0.5)
    tmin = tend - Nmem + 1
    pairs = [(a, b) for a in range(tmin, tend + 1) for b in range(tmin, a + 1)]
    Fx = np.zeros((len(pairs), kL2.size))
    rx = np.zeros((len(pairs), kL2.size))
    for p in range(kL2.size):
        Om = (mpmath.mpf(kL2[p]) + mpmath.mpf(Mass) * mpmath.mpf(Mass)) * mpmath.mpf(dt) ** 2
        c = 1 - Om / 2
        U = {m: cheb_u(m, c) for m in range(-2, tend + 1)}
        K = {t: (-U[t - 3], U[t - 2]) for t in range(1, tend + 1)}
        M = [[mpmath.mpf(F11[p]), mpmath.mpf(F21[p])], [mpmath.mpf(F21[p]), mpmath.mpf(F22[p])]]
        for i, (a, b) in enumerate(pairs):
            Ka, Kb = K[a], K[b]
            v = sum(Ka[x] * M[x][y] * Kb[y] for x in range(2) for y in range(2))
            Fx[i, p] = float(v)
            rx[i, p] = float(mpmath.mpf(dt) * U[a - b - 1])
    np.savez(os.path.join(HERE, "free_theory.npz"), Nx=Nx, sdim=sdim, dt=dt, Mass=Mass, T=T, Nmem=Nmem,
             tend=tend, pairs=np.array(pairs), F=Fx, rho=rx, occ=occ)


def gauss():
    out = []
    for xi, eta, sig, m2, lam, N in [(0.4, 5.0, 1.0, 1.0, 1.0, 1), (0.7, 1.5, 0.5, 0.49, 6.0, 8),
                                     (1.3, 0.0, 2.0, 2.25, 0.5, 4)]:
        x, e, s, mm, l = (mpmath.mpf(v) for v in (xi, eta, sig, m2, lam))
        E = e ** 2 / 2 + s ** 2 / (8 * x ** 2) + mm * x ** 2 / 2 + l * (N + 2) * x ** 4 / (24 * N)
        out.append(dict(xi=xi, eta=eta, sig=sig, Mass2=m2, Lambda=lam, ONgroup=N, E=float(E)))
    with open(os.path.join(HERE, "gauss_energy.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    free_theory()
    gauss()
    print("wrote", os.listdir(HERE))
