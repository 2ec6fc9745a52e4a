"""Central finite differences of the oracle for the largest single-GPU
differentiation workload at a window the host can run: the 48^3 lattice,
box n0 = 4, Q = 1, lambda = 0.5, dt = 0.05, O(1), m = 1 (the state and
coupling of the 4000-step N_mem = 300 run), with N_mem = 20 and 100 steps
(t = 3 .. 102); loss L = sum_p exp(-kL2) n(p, 102).

Writes tests/golden/adjoint_48c_n20_fd.json with L at lambda0 +- h for
h = 1e-5, 5e-4, 1e-3, 2e-3, the central differences D(h), and the
Richardson combinations R(h) = (4 D(h) - D(2h)) / 3 for h = 5e-4 and 1e-3,
which remove the h^2 term.  At this size L ~ 4.5e3 carries about 1e-14 L of
rounding that is not smooth in lambda, so D(1e-5) is noise-limited to about
1e-6; the Richardson values at the larger steps are the accurate ones, and
their difference bounds what is left.  With --bins also the 16 radial |p|
bins over [0, sqrt(12)] of dL/dn(p, 0) (each bin's occupations moved
together), as R(1e-3).  About 15 s per oracle run on 8 threads:
    python tests/golden/make_adjoint_fd_48.py [--bins] [--threads N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "qftdynamicsm.jl_amd"))

from oracle.oracle import Oracle  # noqa: E402
from qftkb import lattice  # noqa: E402

NX, SDIM, NMEM, DT, LAM, ON, M = 48, 3, 20, 0.05, 0.5, 1, 1.0
N0, Q = 4.0, 1.0
T_FINAL = 102
H = 1e-5
HS = (5e-4, 1e-3, 2e-3)  # steps of the Richardson combinations
NBINS = 16
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_48c_n20_fd.json")


def loss(lam, n0, threads):
    o = Oracle(NX, SDIM, NMEM, DT, lam, ON, M, M * M, nthreads=threads)
    o.init_particles(np.ascontiguousarray(n0), M)
    k2 = o.kL2()
    for t in range(3, T_FINAL + 1):
        o.evolve(t)
    L = float((np.exp(-k2) * o.measure(T_FINAL)["n"]).sum())
    o.close()
    return L


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bins", action="store_true")
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    a = ap.parse_args()
    k2 = lattice.kL2_table(NX, SDIM)
    n0 = lattice.n_box(k2, N0, Q)
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    res.update(dict(Nx=NX, sdim=SDIM, N_mem=NMEM, dt=DT, Lambda=LAM, ONgroup=ON, Mass=M, T_final=T_FINAL,
                    h=H, loss="sum_p exp(-kL2) n(p, T_final)", init=f"box n0={N0} Q={Q}"))
    if "dL_dlambda_richardson" not in res:
        t0 = time.time()
        D = {}
        for h in (H,) + HS:
            Lp, Lm = loss(LAM + h, n0, a.threads), loss(LAM - h, n0, a.threads)
            D[h] = (Lp - Lm) / (2 * h)
            res[f"L_plus_{h:g}"], res[f"L_minus_{h:g}"], res[f"D_{h:g}"] = Lp, Lm, D[h]
            print("h", h, D[h], flush=True)
        r1 = (4 * D[HS[0]] - D[HS[1]]) / 3
        r2 = (4 * D[HS[1]] - D[HS[2]]) / 3
        res.update({"dL_dlambda_fd_h": D[H], "richardson_5e-4": r1, "richardson_1e-3": r2,
                    "dL_dlambda_richardson": r1, "richardson_spread": abs(r1 - r2) / abs(r1),
                    "seconds_lambda": time.time() - t0})
        json.dump(res, open(OUT, "w"), indent=1)
        print("lambda richardson", r1, r2, flush=True)
    if a.bins:
        b = lattice.radial_bins(k2, SDIM, NBINS)
        fd = res.get("dL_dn0_bins_fd", [None] * NBINS)
        for i in range(NBINS):
            if fd[i] is not None or not np.any(b == i):
                continue
            d = {}
            for h in HS[1:]:
                e = np.where(b == i, h, 0.0)
                d[h] = (loss(LAM, n0 + e, a.threads) - loss(LAM, n0 - e, a.threads)) / (2 * h)
            fd[i] = (4 * d[HS[1]] - d[HS[2]]) / 3
            res["dL_dn0_bins_fd"] = fd
            json.dump(res, open(OUT, "w"), indent=1)
            print("bin", i, fd[i], flush=True)


if __name__ == "__main__":
    main()
