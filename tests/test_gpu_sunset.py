"""kb_sunset: the sunset self-energies (calcSelfEnergyattp!,
TwoPIScalar_Evolution.jl:25-31) of lattices held in caller device memory,
through the C ABI, against the oracle on random non-symmetric rows and
against the self-energy stage of a KB step on the same rows.  Device memory
comes from tests/hipmem.py (plumbing only); tolerance 1e-10 against the oracle,
1e-14 against the step (the same kernels and twiddles)."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def device_sunset(Nx, sdim, lam, ON, F_rows, r_rows):
    """Run kb_sunset on host rows (nrows x V); returns (SigF, Sigr) on the host."""
    from hipmem import DeviceArray
    from qftkb._abi import sunset, sunset_workspace_bytes
    nrows, V = F_rows.shape
    F = DeviceArray.from_host(F_rows)
    R = DeviceArray.from_host(r_rows)
    SF, SR = DeviceArray(nrows * V), DeviceArray(nrows * V)
    nb = sunset_workspace_bytes(Nx, sdim, nrows)
    ws = DeviceArray(nb, np.uint8)
    assert ws.ptr % 256 == 0
    sunset(Nx, sdim, lam, ON, nrows, F.ptr, R.ptr, SF.ptr, SR.ptr, ws.ptr, nb)
    return SF.host().reshape(nrows, V), SR.host().reshape(nrows, V)


def oracle_rows(Nx, sdim, lam, ON, seed, N=6):
    """Random non-symmetric pairs in an oracle window [1, N]; returns the rows
    (F(N-1, tp), rho(N-1, tp)) for tp in [1, N-1] and the oracle's Sigma."""
    rng = np.random.default_rng(seed)
    o = Oracle(Nx, sdim, N, 0.1, lam, ON, 1.0, 1.0)
    o.init_particles(np.zeros(o.V), 1.0)
    for _ in range(N - 2):
        o.expand()
    for a in range(1, N + 1):
        for b in range(1, a + 1):
            o.set("F", a, b, rng.standard_normal(o.V))
            o.set("r", a, b, rng.standard_normal(o.V))
    tm = N - 1
    o.selfenergies(tm)
    tps = range(1, tm + 1)
    F = np.array([o.get("F", tm, tp) for tp in tps])
    r = np.array([o.get("r", tm, tp) for tp in tps])
    SF = np.array([o.sigma("F", tp) for tp in tps])
    SR = np.array([o.sigma("r", tp) for tp in tps])
    return F, r, SF, SR


@pytest.mark.parametrize("Nx,sdim", [(8, 3), (6, 3), (5, 3), (11, 3), (40, 3), (12, 2), (7, 2), (16, 1), (13, 1)])
def test_sunset_of_device_lattices_matches_oracle(Nx, sdim):
    print(f"\n[case] kb_sunset {Nx}^{sdim} random rows", flush=True)
    F, r, SF, SR = oracle_rows(Nx, sdim, 1.7, 3, seed=Nx * 7 + sdim)
    gF, gR = device_sunset(Nx, sdim, 1.7, 3, F, r)
    for q in range(F.shape[0]):
        assert rel(gF[q], SF[q]) <= 1e-10, ("SigF", q)
        assert rel(gR[q], SR[q]) <= 1e-10, ("Sigr", q)


@pytest.mark.parametrize("Nx,sdim,N", [(32, 3, 12), (9, 3, 7), (10, 2, 9)])
def test_sunset_equals_the_step_stage(Nx, sdim, N):
    """Rows F(t-1, tp), rho(t-1, tp) of a KB run, their Sigma from the step's
    own stage, and kb_sunset of the same rows."""
    from qftkb import KBContext
    print(f"\n[case] kb_sunset vs step stage {Nx}^{sdim} N_mem={N}", flush=True)
    g = KBContext(Nx, sdim, N, 0.1, 1.3, 1, 1.0, 1.0)
    k2 = g.kL2()
    g.init_particles(1.0 / (np.exp(np.sqrt(k2 + 1.0) / 2.0) - 1.0), 1.0)
    t_end = N + 4
    for t in range(3, t_end + 1):
        g.step(t)
    tmin, tmax = g.window()
    tps = range(tmin, tmax)
    F = np.array([g.get("F", tmax - 1, tp) for tp in tps])
    r = np.array([g.get("r", tmax - 1, tp) for tp in tps])
    SF = np.array([g.sigma("F", tp) for tp in tps])
    SR = np.array([g.sigma("r", tp) for tp in tps])
    gF, gR = device_sunset(Nx, sdim, 1.3, 1, F, r)
    for q in range(F.shape[0]):
        assert rel(gF[q], SF[q]) <= 1e-14, ("SigF", q)
        assert rel(gR[q], SR[q]) <= 1e-14, ("Sigr", q)
    g.close()


def test_sunset_rejects_bad_arguments():
    from hipmem import DeviceArray, synchronize
    from qftkb import KBError
    from qftkb._abi import sunset, sunset_workspace_bytes
    with pytest.raises(KBError):
        sunset_workspace_bytes(65, 3, 4)  # past the plane kernels
    with pytest.raises(KBError):
        sunset_workspace_bytes(8, 3, 0)
    V, n = 512, 3
    F, R, S, S2 = (DeviceArray(n * V) for _ in range(4))
    nb = sunset_workspace_bytes(8, 3, n)
    ws = DeviceArray(nb, np.uint8)
    with pytest.raises(KBError) as e:
        sunset(8, 3, 1.0, 1, n, F.ptr, R.ptr, S.ptr, S.ptr, ws.ptr, nb)
    assert "overlap" in str(e.value)
    with pytest.raises(KBError) as e:
        sunset(8, 3, 1.0, 1, n, F.ptr, R.ptr, S.ptr, S2.ptr, ws.ptr, nb - 8)
    assert "workspace" in str(e.value)
    synchronize()
