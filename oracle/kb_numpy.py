"""Independent numpy restatement of the reference KB step, for tiny lattices.

TEST INFRASTRUCTURE ONLY (a second checker that pins the C oracle).

It differs from the C oracle on purpose wherever the two could share a
mistake:
  * storage is a dict keyed by physical time pairs (a, b) with a >= b holding
    physical F and rho -- no ring slots, no sign fold;
  * the sunset uses numpy's rfftn / irfftn exactly like the reference's
    real(rfft(.)) / real(brfft(.)) (TwoPIScalar_Evolution.jl:25-31), halving
    the x axis as Julia's rfft does;
  * the memory integrals are written as the mathematics of
    TwoPIScalar_Evolution.jl:766-782 (getRHS_F!) and :846-854 (getRHS_r!)
    with rho antisymmetric and F symmetric;
  * the 1/N expansion (pexp = 3) follows the reference's momentum-space
    formulation (the CPUred2 back-end: bubbles Fr, F2kr2 of calc_Fr_F2kr2!,
    TwoPIScalar_SimDataStruct.jl:726-750; resummations Irk, IFk of
    TwoPIScalar_Evolution.jl:252-292 and 539-597; self-energies :721-735),
    a per-momentum recursion with no transform inside the time sums, where
    the C oracle restates the CPUfull back-end (a transform pair per term).

pexp: 0 LambdaLO, 1 LambdaNLO, 2 NinverseLO, 3 NinverseNLO, 4 LambdaNLO
classical (the sunset without its rho^2 terms, TwoPIScalar_Evolution.jl:
107-120).

Negative controls: `plant` names one deliberate mistake in the memory
integrals or the memory energy (see PLANTS), so that tests can show that a
check fails on a wrong restatement.  It is None (no mistake) unless a test
sets it.
"""
from __future__ import annotations

import numpy as np


def kL2_table(Nx: int, sdim: int) -> np.ndarray:
    """lev2 per site, sum_i 4 sin^2(pi n_i/Nx), site order x fastest."""
    n = np.arange(Nx)
    k1 = 4.0 * np.sin(np.pi * n / Nx) ** 2
    if sdim == 1:
        return k1.copy()
    if sdim == 2:
        return (k1[None, :] + k1[:, None]).reshape(-1)
    return (k1[None, None, :] + k1[None, :, None] + k1[:, None, None]).reshape(-1)


# deliberate mistakes a test may plant (each one term of getRHS_F!,
# getRHS_r! or getE_mem changed)
PLANTS = (
    "RF_sigr_sign",   # R_F: +w Sigr(z) F(z,t') instead of -w Sigr(z) F(z,t')
    "RF_sigf_sign",   # R_F: -w SigF(z) rho(z,t') instead of +
    "RF_tmin_weight",  # R_F: weight 1 instead of 1/2 at tmin
    "Rr_sign",        # R_rho: + instead of -
    "Rr_upper",       # R_rho: the sum stops at t-3 instead of t-2
    "Emem_sign",      # E_mem with the opposite sign
    "Emem_tmin_weight",  # E_mem: dt/4 instead of dt/8 at tmin
)


class NumpyKB:
    plant = None

    def __init__(self, Nx, sdim, Nmem, dt, lam, ON, Mass, Mass2, pexp=1):
        self.Nx, self.sdim, self.N = Nx, sdim, Nmem
        self.V = Nx ** sdim
        self.shape = (Nx,) * sdim  # numpy index order [z, y, x]
        self.dt, self.lam, self.ON, self.Mass, self.Mass2, self.pexp = dt, lam, ON, Mass, Mass2, pexp
        self.kL2 = kL2_table(Nx, sdim)
        self.F = {}
        self.r = {}
        self.SigF = {}
        self.Sigr = {}
        self.tmin = self.tmax = 0
        self.hartreemass2 = 0.0

    # storage ---------------------------------------------------------
    def getF(self, a, b):
        assert self.tmin <= a <= self.tmax and self.tmin <= b <= self.tmax
        return self.F[(a, b)] if a >= b else self.F[(b, a)]

    def getr(self, a, b):
        assert self.tmin <= a <= self.tmax and self.tmin <= b <= self.tmax
        return self.r[(a, b)] if a >= b else -self.r[(b, a)]

    def setF(self, a, b, v):
        self.F[(a, b) if a >= b else (b, a)] = np.array(v, dtype=np.float64)

    def setr(self, a, b, v):
        v = np.array(v, dtype=np.float64)
        if a >= b:
            self.r[(a, b)] = v
        else:
            self.r[(b, a)] = -v

    def expand(self):
        if self.tmax + 1 > self.N:
            self.tmin += 1
            self.tmax += 1
        else:
            self.tmin = 1
            self.tmax += 1
        for d in (self.F, self.r):
            for k in [k for k in d if k[1] < self.tmin]:
                del d[k]

    # physics ---------------------------------------------------------
    def H(self):
        """Hartreeprefactor, Initialization.jl:245-247."""
        if self.pexp == 2:
            return self.lam / 6.0
        return self.lam * (self.ON + 2) / (6.0 * self.ON)

    @property
    def nlo(self):
        return self.pexp in (1, 3, 4)

    def hartree(self, a, b):
        return self.Mass2 + self.H() / self.V * self.getF(a, b).sum()

    def init_particles(self, n, m):
        self.expand()
        self.expand()
        om = np.sqrt(self.kL2 + m * m)
        self.setF(1, 1, (n + 0.5) / om)
        self.setF(2, 1, (n + 0.5) / om)
        self.setF(2, 2, om * (n + 0.5) * self.dt ** 2 + (n + 0.5) / om)
        z = np.zeros(self.V)
        self.setr(1, 1, z)
        self.setr(2, 1, z + self.dt)
        self.setr(2, 2, z)
        self.hartreemass2 = self.hartree(2, 2)

    def _sunset(self, Fl, rl):
        ax = tuple(range(self.sdim))
        Fx = np.fft.rfftn(Fl.reshape(self.shape[::-1]), axes=ax).real
        rx = np.fft.rfftn(rl.reshape(self.shape[::-1]), axes=ax).real
        ivol = 1.0 / self.V
        lam, ON = self.lam, self.ON
        pr = ivol ** 3 * (-lam ** 2 * (ON + 2) / (6 * ON ** 2))
        pF = ivol ** 3 * (-lam ** 2 * (ON + 2) / (18 * ON ** 2))
        s = self.shape[::-1]
        if self.pexp == 4:
            gr = np.fft.irfftn(rx * Fx ** 2, s=s, axes=ax) * self.V
            gF = np.fft.irfftn(Fx ** 3, s=s, axes=ax) * self.V
        else:
            gr = np.fft.irfftn(rx * (Fx ** 2 - rx ** 2 / 12.0), s=s, axes=ax) * self.V
            gF = np.fft.irfftn(Fx * (Fx ** 2 - 0.75 * rx ** 2), s=s, axes=ax) * self.V
        return pF * gF.reshape(-1), pr * gr.reshape(-1)

    def _cos(self, x):
        """real(rfft(x)) on the full grid (the transform of the even part)."""
        return np.fft.fftn(x.reshape(self.shape[::-1])).real.reshape(-1)

    def bubbles(self, a, b):
        """(F2kr2, Fr) of the pair (a, b): ivol^2 C[Fx^2 - rx^2/4] and
        ivol^2 C[Fx rx] with Fx = C F(a,b), rx = C rho(a,b) (physical)."""
        key = (a, b)
        if key not in self._bub:
            Fx, rx = self._cos(self.getF(a, b)), self._cos(self.getr(a, b))
            iv2 = 1.0 / self.V ** 2
            self._bub[key] = (iv2 * self._cos(Fx * Fx - 0.25 * rx * rx), iv2 * self._cos(Fx * rx))
        return self._bub[key]

    def resum(self, T):
        """Momentum-space Ir(T, tp), IF(T, tp) for tp in [tmin, T]."""
        self._bub = {}
        tmin, lam, dt = self.tmin, self.lam, self.dt
        Ir = {T: np.zeros(self.V)}
        for tp in range(T - 1, tmin - 1, -1):
            acc = np.zeros(self.V)
            for z in range(tp + 1, T):
                acc += self.bubbles(z, tp)[1] * Ir[z]
            Ir[tp] = lam / 3.0 * (self.bubbles(T, tp)[1] - dt * acc)
        IF = {}
        for tp in range(tmin, T + 1):
            acc1 = np.zeros(self.V)
            for z in range(tmin, T):
                w = 0.5 if z == tmin else 1.0
                acc1 += w * Ir[z] * self.bubbles(z, tp)[0]
            acc2 = np.zeros(self.V)
            for z in range(tmin, tp):
                w = 0.5 if z == tmin else 1.0
                acc2 += w * IF[z] * self.bubbles(z, tp)[1]
            IF[tp] = lam / 6.0 * (self.bubbles(T, tp)[0] - dt * acc1) + lam / 3.0 * dt * acc2
        self.Irk, self.IFk = Ir, IF
        return Ir, IF

    def _ninv_sigma(self, T):
        Ir, IF = self.resum(T)
        ivol = 1.0 / self.V
        pre = -self.lam / (3.0 * self.ON) * ivol
        for tp in range(self.tmin, T + 1):
            A, B = self._cos(self.getF(T, tp)), self._cos(self.getr(T, tp))
            Irx, IFx = ivol * self._cos(Ir[tp]), ivol * self._cos(IF[tp])
            self.Sigr[tp] = pre * self._cos(A * Irx + B * IFx)
            self.SigF[tp] = pre * self._cos(A * IFx - 0.25 * B * Irx)

    def evolve(self, t):
        assert t == self.tmax + 1
        self.expand()
        tmin = self.tmin
        if self.pexp in (1, 4):
            for tp in range(tmin, t):
                self.SigF[tp], self.Sigr[tp] = self._sunset(self.getF(t - 1, tp), self.getr(t - 1, tp))
        elif self.pexp == 3:
            self._ninv_sigma(t - 1)
        self.hartreemass2 = self.hartree(t - 1, t - 1)
        Om = (self.kL2 + self.hartreemass2) * self.dt ** 2
        dt3 = self.dt ** 3
        newF, newr = {}, {}
        for tp in range(tmin, t):
            RF = np.zeros(self.V)
            Rr = np.zeros(self.V)
            if self.nlo:
                pl = self.plant
                for z in range(tp + 1, t - 2 if pl == "Rr_upper" else t - 1):
                    Rr -= self.Sigr[z] * self.getr(z, tp)
                for z in range(tmin, t - 1):
                    w = 0.5 if z == tmin and pl != "RF_tmin_weight" else 1.0
                    RF -= w * self.Sigr[z] * self.getF(z, tp) * (-1.0 if pl == "RF_sigr_sign" else 1.0)
                for z in range(tmin, tp):
                    w = 0.5 if z == tmin and pl != "RF_tmin_weight" else 1.0
                    RF += w * self.SigF[z] * self.getr(z, tp) * (-1.0 if pl == "RF_sigf_sign" else 1.0)
                if pl == "Rr_sign":
                    Rr = -Rr
            if tp == t - 1:
                newr[tp] = np.full(self.V, self.dt)
            else:
                newr[tp] = (2 - Om) * self.getr(t - 1, tp) - self.getr(t - 2, tp) + dt3 * Rr
            newF[tp] = (2 - Om) * self.getF(t - 1, tp) - self.getF(t - 2, tp) + dt3 * RF
        for tp in range(tmin, t):
            self.setF(t, tp, newF[tp])
            self.setr(t, tp, newr[tp])
        self.setr(t, t, np.zeros(self.V))
        RF = np.zeros(self.V)
        if self.nlo:
            pl = self.plant
            for z in range(tmin, t - 1):
                w = 0.5 if z == tmin and pl != "RF_tmin_weight" else 1.0
                RF -= w * self.Sigr[z] * self.getF(z, t) * (-1.0 if pl == "RF_sigr_sign" else 1.0)
            for z in range(tmin, t):
                w = 0.5 if z == tmin and pl != "RF_tmin_weight" else 1.0
                RF += w * self.SigF[z] * self.getr(z, t) * (-1.0 if pl == "RF_sigf_sign" else 1.0)
        self.setF(t, t, (2 - Om) * self.getF(t - 1, t) - self.getF(t - 2, t) + dt3 * RF)

    def measure(self, t):
        dt = self.dt
        ddF = (self.getF(t, t) - 2 * self.getF(t - 1, t) + self.getF(t - 1, t - 1)) / dt ** 2
        F = self.getF(t - 1, t - 1)
        with np.errstate(invalid="ignore", divide="ignore"):
            om = np.where(ddF / F > 0, np.sqrt(np.abs(ddF / F)), 0.0)
        arg = 0.25 * dt ** 2 * om ** 2
        c = np.where(arg < 1, np.sqrt(np.abs(1 - arg)), 1.0)
        n = np.where(ddF * F > 0, c * np.sqrt(np.abs(ddF * F)) - 0.5, 0.0)
        Ekin = 0.5 * ddF.sum()
        hs = F.sum()
        Epot = (0.5 * (self.Mass2 + self.kL2 + 0.5 * self.H() / self.V * hs) * F).sum()
        Emem = 0.0
        if t != 2 and self.nlo:
            tmin = self.tmin
            w0 = 0.25 if self.plant == "Emem_tmin_weight" else 0.125
            Emem = w0 * dt * (self.Sigr[tmin] * self.getF(tmin, t - 1)
                              - self.SigF[tmin] * self.getr(tmin, t - 1)).sum()
            for tp in range(tmin + 1, t - 1):
                Emem += 0.25 * dt * (self.Sigr[tp] * self.getF(tp, t - 1)
                                     - self.SigF[tp] * self.getr(tp, t - 1)).sum()
            if self.plant == "Emem_sign":
                Emem = -Emem
        return dict(time=(t - 1) * dt * self.Mass, E=Ekin + Epot + Emem, E_kin=Ekin, E_pot=Epot,
                    E_mem=Emem, M2=self.hartreemass2, n=n, ddF=ddF, F=F, omega=om, c=c)
