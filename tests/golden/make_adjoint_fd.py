"""Central finite differences of the oracle for the differentiation
workload: 24^3 lattice, N_mem = 200, 500 steps (t = 3 .. 502), dt = 0.1,
lambda = 1, O(1), m = 1, thermal T = 2 initial occupations, loss
L = sum_p exp(-|p|^2) n(p, 502).

Writes tests/golden/adjoint_24c_n200_fd.json with L(lambda0 +- h), h = 1e-5, and the
central difference; with --bins also the 16 radial |p| bins over
[0, sqrt(12)] of dL/dn(p, 0) (each bin's occupations moved together).
Slow (minutes per oracle run); run by hand:
    python tests/golden/make_adjoint_fd.py [--bins] [--threads N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402

NX, SDIM, NMEM, DT, LAM, ON, M = 24, 3, 200, 0.1, 1.0, 1, 1.0
T_FINAL = 502
H = 1e-5
NBINS = 16
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_24c_n200_fd.json")


def occupations(k2):
    return 1.0 / (np.exp(np.sqrt(k2 + M * M) / 2.0) - 1.0)


def bins_of(k2):
    edges = np.linspace(0.0, np.sqrt(4.0 * SDIM), NBINS + 1)
    return np.minimum(np.searchsorted(edges, np.sqrt(k2), side="right") - 1, NBINS - 1)


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
    o = Oracle(NX, SDIM, NMEM, DT, LAM, ON, M, M * M, nthreads=1)
    k2 = o.kL2()
    o.close()
    n0 = occupations(k2)
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    res.update(dict(Nx=NX, sdim=SDIM, N_mem=NMEM, dt=DT, Lambda=LAM, ONgroup=ON, Mass=M, T_final=T_FINAL,
                    h=H, loss="sum_p exp(-kL2) n(p, T_final)", init="thermal T=2"))
    if "dL_dlambda_fd" not in res:
        t0 = time.time()
        Lp = loss(LAM + H, n0, a.threads)
        Lm = loss(LAM - H, n0, a.threads)
        res.update(L_plus=Lp, L_minus=Lm, dL_dlambda_fd=(Lp - Lm) / (2 * H), seconds_lambda=time.time() - t0)
        json.dump(res, open(OUT, "w"), indent=1)
        print("lambda", res["dL_dlambda_fd"], flush=True)
    if a.bins:
        b = bins_of(k2)
        fd = res.get("dL_dn0_bins_fd", [None] * NBINS)
        for i in range(NBINS):
            if fd[i] is not None or not np.any(b == i):
                continue
            e = np.where(b == i, H, 0.0)
            fd[i] = (loss(LAM, n0 + e, a.threads) - loss(LAM, n0 - e, a.threads)) / (2 * H)
            res["dL_dn0_bins_fd"] = fd
            json.dump(res, open(OUT, "w"), indent=1)
            print("bin", i, fd[i], flush=True)


if __name__ == "__main__":
    main()
