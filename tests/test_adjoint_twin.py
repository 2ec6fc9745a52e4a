"""The adjoint of the KB step, pinned on the CPU: the numpy reverse-mode twin
(tests/adjoint_twin.py) against central finite differences (h = 1e-5) of
the C oracle's forward run, for L = sum_p wgt(p) n(p, T)."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.adjoint_twin import gradient

H = 1e-5
TOL = 1e-6


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


def oracle_loss(args, n0, initmass, T, wgt):
    o = Oracle(*args)
    o.init_particles(np.ascontiguousarray(n0, dtype=np.float64), initmass)
    for t in range(3, T + 1):
        o.evolve(t)
    L = float((wgt * o.measure(T)["n"]).sum())
    o.close()
    return L


def with_lambda(args, lam):
    a = list(args)
    a[4] = lam
    return tuple(a)


CASES = [
    # (Nx, sdim, Nmem, dt, lambda, ON, Mass, Mass2, pexp), steps T
    pytest.param((4, 3, 5, 0.1, 1.3, 1, 1.0, 1.0, 1), 14, id="4^3-sliding-window"),
    pytest.param((4, 3, 16, 0.1, 1.0, 1, 1.0, 1.0, 1), 12, id="4^3-full-memory"),
    pytest.param((6, 2, 6, 0.08, 1.7, 4, 0.9, 0.81, 1), 15, id="6^2-O(4)-sliding"),
    pytest.param((5, 1, 4, 0.1, 2.0, 2, 1.0, 1.0, 1), 11, id="5-1d-minimal-window"),
    pytest.param((4, 3, 5, 0.1, 1.0, 1, 1.0, 1.0, 0), 14, id="4^3-LO"),
    pytest.param((4, 3, 6, 0.1, 1.6, 1, 1.0, 1.0, 4), 13, id="4^3-classical"),
]


@pytest.mark.parametrize("args,T", CASES)
def test_twin_gradient_matches_oracle_fd(args, T):
    from oracle.kb_numpy import kL2_table
    Nx, sdim = args[0], args[1]
    k2 = kL2_table(Nx, sdim)
    m = args[6]
    n0 = thermal(k2, m=m)
    wgt = np.exp(-k2)
    L, lb, dn0, _ = gradient(args, n0, m, T, wgt)
    assert L == pytest.approx(oracle_loss(args, n0, m, T, wgt), rel=1e-12)
    lam = args[4]
    fd = (oracle_loss(with_lambda(args, lam + H), n0, m, T, wgt)
          - oracle_loss(with_lambda(args, lam - H), n0, m, T, wgt)) / (2 * H)
    scale = max(abs(fd), 1e-3 * abs(L))
    assert abs(lb - fd) <= TOL * scale, (lb, fd)
    # per-site occupations: every site on these tiny lattices
    fds = np.empty_like(n0)
    for p in range(n0.size):
        e = np.zeros_like(n0)
        e[p] = H
        fds[p] = (oracle_loss(args, n0 + e, m, T, wgt) - oracle_loss(args, n0 - e, m, T, wgt)) / (2 * H)
    scale = np.max(np.abs(fds))
    assert np.max(np.abs(dn0 - fds)) <= TOL * scale, np.max(np.abs(dn0 - fds)) / scale
