"""Reverse-mode twin of the KB step, in numpy, for tiny lattices.

TEST INFRASTRUCTURE ONLY.  It states the adjoint of one KB step (the
forward step is the numpy twin of the oracle, oracle/kb_numpy.py, which
restates TwoPIScalar_Evolution.jl:924-981) entry by entry in physical time
indices, with no ring slots, sign folds or fusion, so that

  * the adjoint mathematics is pinned on the CPU against central finite
    differences of the C oracle (tests/test_adjoint_twin.py), and
  * the fused, slot-space HIP reverse kernels can be compared with it
    quantity by quantity.

Loss: L = sum_p wgt(p) n(p, T), n measured at time T as createMeasurement
does (TwoPIScalar_Measurement.jl:38-59).  Gradients: dL/dlambda and
dL/dn(p, 0) per site (initial occupations of initSimData!,
TwoPIScalar_Initialization.jl:8-29), and dL/dF for the three initial rows.
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np

from oracle.kb_numpy import NumpyKB


class TapeKB(NumpyKB):
    """NumpyKB that keeps every entry ever written (the window still
    restricts reads) and records the per-step Hartree mass."""

    def expand(self):
        if self.tmax + 1 > self.N:
            self.tmin += 1
            self.tmax += 1
        else:
            self.tmin = 1
            self.tmax += 1

    def evolve(self, t):
        super().evolve(t)
        self.m2_at[t] = self.hartreemass2
        self.tmin_at[t] = self.tmin


def cosT(x, shape):
    """sum_x X(x) cos(k.x) on the full grid (self-adjoint)."""
    return np.fft.fftn(x.reshape(shape)).real.reshape(-1)


def seed_from_n(kb, T, wgt):
    """dL/dF of F(T-1,T-1), F(T,T-1), F(T,T) for L = sum wgt n(T)."""
    dt = kb.dt
    Ftt, Fmt, F = kb.getF(T, T), kb.getF(T - 1, T), kb.getF(T - 1, T - 1)
    ddF = (Ftt - 2 * Fmt + F) / dt ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        om2 = np.where(ddF / F > 0, ddF / F, 0.0)
    stable = 0.25 * dt ** 2 * om2 < 1
    pos = ddF * F > 0
    D = np.where(stable, ddF * F - 0.25 * dt ** 2 * ddF ** 2, ddF * F)
    sq = np.sqrt(np.where(pos & (D > 0), D, 1.0))
    dn_dF = np.where(pos, ddF / (2 * sq), 0.0)
    dn_dd = np.where(pos, np.where(stable, (F - 0.5 * dt ** 2 * ddF) / (2 * sq), F / (2 * sq)), 0.0)
    g = wgt * dn_dd / dt ** 2
    return {(T, T): g, (T, T - 1): -2 * g, (T - 1, T - 1): wgt * dn_dF + g}


def run_forward(args, n0, initmass, T):
    kb = TapeKB(*args)
    kb.m2_at, kb.tmin_at = {}, {}
    kb.init_particles(n0, initmass)
    for t in range(3, T + 1):
        kb.evolve(t)
    return kb


def loss(kb, T, wgt):
    return float((wgt * kb.measure(T)["n"]).sum())


def reverse(kb, T, wgt, initmass, record=None, stop=None):
    """Reverse sweep over steps T .. 3.  Returns (dL/dlambda, dL/dn0,
    (dF11, dF21, dF22)).  `record`, if a dict, receives per-step
    intermediates keyed by t for kernel-level comparison.  With `stop`, the
    sweep ends after reversing step `stop` and returns the adjoint entries
    accumulated so far as (Fb, rb) dicts keyed (a, b), a >= b."""
    V, dt, lam = kb.V, kb.dt, kb.lam
    dt3 = dt ** 3
    H = kb.H()
    shape = kb.shape[::-1]
    ivol = 1.0 / V
    pr = ivol ** 3 * (-lam ** 2 * (kb.ON + 2) / (6 * kb.ON ** 2))
    pF = ivol ** 3 * (-lam ** 2 * (kb.ON + 2) / (18 * kb.ON ** 2))
    Fb = defaultdict(lambda: np.zeros(V))  # dL/dF(a,b), key a >= b
    rb = defaultdict(lambda: np.zeros(V))  # dL/drho(a,b), key a > b

    def addF(a, b, v):
        Fb[(a, b) if a >= b else (b, a)] += v

    def addr(a, b, v):  # adjoint of physical rho(a, b)
        if a > b:
            rb[(a, b)] += v
        elif a < b:
            rb[(b, a)] -= v

    def getrb(a, b):
        return rb[(a, b)] if a > b else -rb[(b, a)]

    for k, v in seed_from_n(kb, T, wgt).items():
        Fb[k] += v
    lam_bar = 0.0
    # rho^2 terms of the sunset: quantum (Evolution.jl:25-31), or none for the
    # classical sunset (pexp 4, :107-120)
    qF, qr = (0.0, 0.0) if kb.pexp == 4 else (0.75, 1.0 / 12.0)
    kb.tmax = T
    for t in range(T, 2, -1):
        tmin = kb.tmin_at[t]
        kb.tmin, kb.tmax = tmin, t
        F = kb.getF
        r = kb.getr
        # recompute the step's self-energies and Hartree mass
        SF, Sr, Fx, rx = {}, {}, {}, {}
        for z in range(tmin, t):
            Fx[z] = cosT(F(t - 1, z), shape)
            rx[z] = cosT(r(t - 1, z), shape)
            SF[z] = pF * cosT(Fx[z] * (Fx[z] ** 2 - qF * rx[z] ** 2), shape) if kb.pexp else np.zeros(V)
            Sr[z] = pr * cosT(rx[z] * (Fx[z] ** 2 - qr * rx[z] ** 2), shape) if kb.pexp else np.zeros(V)
        M2 = kb.m2_at[t]
        Om = (kb.kL2 + M2) * dt ** 2
        w = {z: (0.5 if z == tmin else 1.0) for z in range(tmin, t)}
        m = {z: (1.0 if z <= t - 2 else 0.0) for z in range(tmin, t)}
        SFb = defaultdict(lambda: np.zeros(V))
        Srb = defaultdict(lambda: np.zeros(V))
        # diagonal F(t,t) = (2-Om) F(t,t-1) - F(t,t-2) + dt^3 R_F(t)
        Fbtt = Fb[(t, t)].copy()
        RFt = dt3 * Fbtt
        Om_bar = -F(t, t - 1) * Fbtt
        fbar, rbarp = {}, {}
        for b in range(tmin, t):
            fb = Fb[(t, b)].copy()
            if b == t - 1:
                fb += (2 - Om) * Fbtt
            if b == t - 2:
                fb -= Fbtt
            if kb.pexp:
                # R_F(t) = -sum_{z<=t-2} w Sr F(z,t) + sum_{z<=t-1} w SF rho(z,t)
                fb -= m[b] * w[b] * Sr[b] * RFt
                SFb[b] += w[b] * r(b, t) * RFt
                Srb[b] -= m[b] * w[b] * F(b, t) * RFt
            fbar[b] = fb
            if b <= t - 2:
                rp = getrb(t, b).copy()
                if kb.pexp:
                    rp -= w[b] * SF[b] * RFt  # rho(b,t) = -rho(t,b)
                rbarp[b] = rp
        # row t: F(t,b) = (2-Om) F(t-1,b) - F(t-2,b) + dt^3 R_F(b), same for rho
        RFb = {b: dt3 * fbar[b] for b in fbar}
        Rrb = {b: dt3 * rbarp[b] for b in rbarp}
        for b in range(tmin, t):
            addF(t - 1, b, (2 - Om) * fbar[b])
            addF(t - 2, b, -fbar[b])
            Om_bar -= F(t - 1, b) * fbar[b]
        for b in range(tmin, t - 1):
            if b <= t - 3:
                addr(t - 1, b, (2 - Om) * rbarp[b])
                addr(t - 2, b, -rbarp[b])
            Om_bar -= r(t - 1, b) * rbarp[b]
        if kb.pexp:
            # R_F(b) = -sum_{z=tmin}^{t-2} w Sr(z) F(z,b) + sum_{z=tmin}^{b-1} w SF(z) rho(z,b)
            for b in range(tmin, t):
                for z in range(tmin, t - 1):
                    addF(z, b, -w[z] * Sr[z] * RFb[b])
                    Srb[z] -= w[z] * F(z, b) * RFb[b]
                for z in range(tmin, b):
                    addr(z, b, w[z] * SF[z] * RFb[b])
                    SFb[z] += w[z] * r(z, b) * RFb[b]
            # R_r(b) = -sum_{z=b+1}^{t-2} Sr(z) rho(z,b)
            for b in range(tmin, t - 1):
                for z in range(b + 1, t - 1):
                    addr(z, b, -Sr[z] * Rrb[b])
                    Srb[z] -= r(z, b) * Rrb[b]
            # sunset
            for z in range(tmin, t):
                gF = pF * cosT(SFb[z], shape)
                gr = pr * cosT(Srb[z], shape)
                Fxb = gF * (3 * Fx[z] ** 2 - qF * rx[z] ** 2) + gr * 2 * rx[z] * Fx[z]
                rxb = gF * (-2 * qF * Fx[z] * rx[z]) + gr * (Fx[z] ** 2 - 3 * qr * rx[z] ** 2)
                addF(t - 1, z, cosT(Fxb, shape))
                if z != t - 1:
                    addr(t - 1, z, cosT(rxb, shape))
                if lam != 0:
                    lam_bar += 2.0 / lam * float((SF[z] * SFb[z] + Sr[z] * Srb[z]).sum())
        # Hartree M2 = Mass2 + H/V sum F(t-1,t-1); Om = (kL2 + M2) dt^2
        M2b = dt ** 2 * float(Om_bar.sum())
        addF(t - 1, t - 1, np.full(V, H / V * M2b))
        lam_bar += (kb.ON + 2) / (6.0 * kb.ON) / V * float(F(t - 1, t - 1).sum()) * M2b
        if record is not None:
            record[t] = dict(RFb=RFb, Rrb=Rrb, SFb=dict(SFb), Srb=dict(Srb), M2b=M2b, Om_bar=Om_bar)
        for b in range(tmin, t + 1):
            Fb.pop((t, b), None)
            rb.pop((t, b), None)
        if stop is not None and t == stop:
            return dict(Fb), dict(rb)
    dF11, dF21, dF22 = Fb[(1, 1)], Fb[(2, 1)], Fb[(2, 2)]
    om = np.sqrt(kb.kL2 + initmass ** 2)
    dn0 = dF11 / om + dF21 / om + dF22 * (om * dt ** 2 + 1.0 / om)
    return lam_bar, dn0, (dF11, dF21, dF22)


def gradient(args, n0, initmass, T, wgt, record=None):
    """(L, dL/dlambda, dL/dn0) with args = (Nx, sdim, Nmem, dt, lam, ON,
    Mass, Mass2[, pexp])."""
    kb = run_forward(args, n0, initmass, T)
    L = loss(kb, T, wgt)
    lb, dn0, dF = reverse(kb, T, wgt, initmass, record)
    return L, lb, dn0, dF
