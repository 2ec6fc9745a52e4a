"""kb_sunset_noalloc (the sunset self-energies with the library's static
per-device scratch as workspace) against kb_sunset with a caller workspace,
through the C ABI: two calls in a row, the second with a smaller input, plus
the size limit.  Device memory comes from tests/hipmem.py (plumbing only)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def both(Nx, sdim, nrows, lam, ON, seed):
    """The same random rows through kb_sunset and kb_sunset_noalloc."""
    from hipmem import DeviceArray, synchronize
    from qftkb._abi import sunset, sunset_noalloc, sunset_workspace_bytes
    rng = np.random.default_rng(seed)
    V = Nx ** sdim
    Fh, Rh = rng.standard_normal(nrows * V), rng.standard_normal(nrows * V)
    F, R = DeviceArray.from_host(Fh), DeviceArray.from_host(Rh)
    out = [DeviceArray(nrows * V) for _ in range(4)]
    for o in out:
        o.fill(np.nan)
    nb = sunset_workspace_bytes(Nx, sdim, nrows)
    ws = DeviceArray(nb, np.uint8)
    sunset(Nx, sdim, lam, ON, nrows, F.ptr, R.ptr, out[0].ptr, out[1].ptr, ws.ptr, nb)
    sunset_noalloc(Nx, sdim, lam, ON, nrows, F.ptr, R.ptr, out[2].ptr, out[3].ptr)
    synchronize()
    h = [o.host().reshape(nrows, V) for o in out]
    return Fh.reshape(nrows, V), Rh.reshape(nrows, V), h


def test_noalloc_matches_workspace_entry_point_twice():
    print("\n[case] kb_sunset_noalloc 32^3 x 12 rows, then 16^3 x 5 rows", flush=True)
    for Nx, sdim, nrows, seed in ((32, 3, 12, 11), (16, 3, 5, 12)):
        F, R, (sf, sr, nf, nr) = both(Nx, sdim, nrows, 1.4, 2, seed)
        assert np.all(np.isfinite(nf)) and np.all(np.isfinite(nr))
        for q in range(nrows):
            assert rel(nf[q], sf[q]) <= 1e-10, ("SigF", Nx, q)
            assert rel(nr[q], sr[q]) <= 1e-10, ("Sigr", Nx, q)
        print(f"[case] {Nx}^{sdim}: bitwise equal = {np.array_equal(nf, sf) and np.array_equal(nr, sr)}",
              flush=True)
    # the second (smaller) call also against the oracle's sunset of the same rows
    from oracle.oracle import Oracle
    o = Oracle(16, 3, 7, 0.1, 1.4, 2, 1.0, 1.0)
    o.init_particles(np.zeros(o.V), 1.0)
    for _ in range(4):
        o.expand()  # window [1, 6]
    for q in range(5):
        o.set("F", 6, q + 1, F[q])
        o.set("r", 6, q + 1, R[q])
    o.set("F", 6, 6, F[0])
    o.set("r", 6, 6, R[0])
    o.expand()  # window [1, 7]: the oracle's Sigma(6, tp) for tp = 1 .. 6
    o.selfenergies(6)
    for q in range(5):
        assert rel(nf[q], o.sigma("F", q + 1)) <= 1e-10
        assert rel(nr[q], o.sigma("r", q + 1)) <= 1e-10


def test_noalloc_two_dimensional_and_limit():
    from hipmem import DeviceArray, synchronize
    from qftkb import KBError
    from qftkb._abi import sunset_noalloc
    F, R, (sf, sr, nf, nr) = both(24, 2, 9, 0.8, 1, 13)
    assert rel(nf, sf) <= 1e-10 and rel(nr, sr) <= 1e-10
    # 32 rows of 64^3 need 4 KiB + 32 * 64^2 * 33 * 16 B = 66 MiB of scratch: refused before any launch
    a, b, c, d = (DeviceArray(32 * 64 ** 3) for _ in range(4))
    with pytest.raises(KBError) as e:
        sunset_noalloc(64, 3, 1.0, 1, 32, a.ptr, b.ptr, c.ptr, d.ptr)
    assert "limit" in str(e.value)
    synchronize()
