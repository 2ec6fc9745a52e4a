"""Adjoint kernels over seeded random parameter sets, against the numpy
reverse-mode twin (tests/adjoint_twin.py, pinned against oracle finite
differences in tests/test_adjoint_twin.py).

The hand-picked adjoint cases fix one parameter family at a time; these draw
the lattice (1D-3D, ragged site blocks), window (filling and sliding), O(N)
group, coupling, time step, masses with Mass2 != Mass^2 (a negative Mass2,
the tachyonic model, in some), the expansion (LambdaLO, the quantum and the
classical sunset), the loss weights and the recording mode (a tape of every
dropped column, or window checkpoints every C steps) together.  Tolerance
1e-10 against the twin (both exact; FP64 rounding only).
"""
import numpy as np
import pytest

from oracle.kb_numpy import kL2_table
from tests.adjoint_twin import gradient

pytestmark = pytest.mark.gpu

TWIN_TOL = 1e-10
SEED = 20261204
NCASES = 30


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / s if s > 0 else np.max(np.abs(a - b))


def draw_cases():
    rng = np.random.default_rng(SEED)
    cases = []
    for i in range(NCASES):
        sdim = int(rng.choice([1, 2, 3, 3]))
        Nx = int(rng.integers(2, {1: 24, 2: 9, 3: 6}[sdim] + 1))
        Nmem = int(rng.integers(3, 13))
        T = Nmem + int(rng.integers(0, 11))
        Mass = float(rng.uniform(0.7, 1.3))
        Mass2 = -0.3 * Mass * Mass if rng.uniform() < 0.15 else Mass * Mass * float(rng.uniform(0.7, 1.3))
        pexp = int((0, 1, 4)[i % 3])
        args = (Nx, sdim, Nmem, float(rng.choice([0.05, 0.08, 0.1])), float(rng.uniform(0.3, 2.5)),
                int(rng.integers(1, 5)), Mass, Mass2, pexp)
        cases.append(dict(args=args, T=max(T, 4), C=int(rng.choice([0, 1, 3, 50])),
                          Tocc=float(rng.uniform(0.8, 3.0)), wscale=float(rng.uniform(0.3, 2.0))))
    return cases


CASES = draw_cases()


def _id(c):
    a = c["args"]
    return f"{a[0]}^{a[1]}-N{a[2]}-T{c['T']}-p{a[8]}-O{a[5]}-C{c['C']}"


@pytest.mark.parametrize("c", CASES, ids=[_id(c) for c in CASES])
def test_random_adjoint_matches_twin(c):
    from qftkb import KBContext
    args, T = c["args"], c["T"]
    print(f"\n[case] random adjoint {args} T={T} C={c['C']}", flush=True)
    k2 = kL2_table(args[0], args[1])
    m = args[6]
    n0 = 1.0 / (np.exp(np.sqrt(k2 + m * m) / (c["Tocc"] * m)) - 1.0)
    wgt = np.exp(-c["wscale"] * k2)
    g = KBContext(*args[:8], pexp=args[8])
    g.init_particles(n0, m)
    g.adj_record(c["C"], T)
    for t in range(3, T + 1):
        g.step(t)
    L = float((wgt * g.measure(T)["n"]).sum())
    gr = g.adj_gradient(wgt, n0=True, F0=True)
    Lt, lt, dnt, dFt = gradient(args, n0, m, T, wgt)
    assert L == pytest.approx(Lt, rel=1e-12)
    assert gr["lambda"] == pytest.approx(lt, rel=TWIN_TOL, abs=TWIN_TOL * abs(Lt))
    assert rel(gr["n0"], dnt) <= TWIN_TOL
    for k in range(3):
        assert rel(gr["F0"][k], dFt[k]) <= TWIN_TOL, k
    g.close()
