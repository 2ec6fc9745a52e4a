"""Every tile shape of the memory-integral kernel, and the autotuner that
picks among them, against the oracle (tolerance 1e-10, norm-wise relative)."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10
ARGS = (6, 3, 140, 0.1, 1.3, 1, 1.0, 1.0)  # long window: several 32/64/128-row chunks per column block


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


@pytest.fixture(scope="module")
def reference():
    o = Oracle(*ARGS)
    n = thermal(o.kL2())
    o.init_particles(n, 1.0)
    rows = {}
    for t in range(3, 161):
        o.evolve(t)
        if t in (60, 145, 160):
            rows[t] = {w: [o.get(w, t, tp) for tp in range(o.window()[0], t + 1)] for w in "Fr"}
    return n, rows


def test_every_tile_shape_matches_oracle(reference):
    from qftkb import KBContext
    n, rows = reference
    shapes = KBContext(*ARGS, tiles=None).tile_shapes()
    assert len(shapes) >= 2
    for v, shape in enumerate(shapes):
        print(f"\n[case] tile shape {v} {shape}", flush=True)
        g = KBContext(*ARGS, tiles=v)
        assert g.tiles() == v
        g.init_particles(n, 1.0)
        for t in range(3, 161):
            g.step(t)
            if t in rows:
                tmin = g.window()[0]
                for w in "Fr":
                    for k, tp in enumerate(range(tmin, t + 1)):
                        assert rel(g.get(w, t, tp), rows[t][w][k]) <= RTOL, (v, w, t, tp)
        g.close()


def test_autotune_keeps_results(reference):
    from qftkb import KBContext
    n, rows = reference
    print("\n[case] autotune mid-run", flush=True)
    g = KBContext(*ARGS, tiles=None)
    g.init_particles(n, 1.0)
    for t in range(3, 146):
        g.step(t)
        if t == 100:
            r = g.autotune(reps=2)
            assert 0 <= r["variant"] < len(g.tile_shapes())
            assert g.tiles() == r["variant"]
            assert all(x > 0 for x in r["ms"])
    tmin = g.window()[0]
    for w in "Fr":
        for k, tp in enumerate(range(tmin, 146)):
            assert rel(g.get(w, 145, tp), rows[145][w][k]) <= RTOL
    g.close()


def test_every_fft_block_size_matches_oracle(reference):
    from qftkb import KBContext
    n, rows = reference
    probe = KBContext(*ARGS, tiles=None)
    shapes = probe.fft_shapes()
    probe.close()
    assert len(shapes) >= 2
    for v, threads in enumerate(shapes):
        print(f"\n[case] plane kernels with {threads} threads", flush=True)
        g = KBContext(*ARGS, tiles=None)
        g.set_fft_tiles(v)
        assert g.fft_tiles() == v
        g.init_particles(n, 1.0)
        for t in range(3, 61):
            g.step(t)
        tmin = g.window()[0]
        for w in "Fr":
            for k, tp in enumerate(range(tmin, 61)):
                assert rel(g.get(w, 60, tp), rows[60][w][k]) <= RTOL, (v, w, tp)
        r = g.autotune_fft(reps=2)
        assert 0 <= r["variant"] < len(shapes) and all(x > 0 for x in r["ms"])
        g.close()
