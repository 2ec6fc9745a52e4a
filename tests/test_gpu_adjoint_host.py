"""Host-snapshot checkpointing of the adjoint (kb_adj_record_host): the
gradient equals the taped one, and the number of recomputed forward steps
equals that of the binomial checkpointing schedule restated here.

Tolerance: 1e-11 against the taped gradient (the recomputed forward steps
repeat the recorded ones up to the order of the memory integral's LDS
sums)."""
import numpy as np
import pytest

TOL = 1e-11
ARGS = (6, 3, 8, 0.1, 1.4, 1, 1.0, 1.0)


def thermal(k2, T=2.0, m=1.0):
    return 1.0 / (np.exp(np.sqrt(k2 + m * m) / (T * m)) - 1.0)


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / s if s > 0 else np.max(np.abs(a - b))


# ---- the schedule, restated ---------------------------------------------------
def revolve_tables(nseg, slots):
    """F[n][f]: least segments advanced to reverse n segments whose start
    window is held, with f free slots; G[n][f]: the same when the recording
    run passes every boundary for free and tapes the last segment.  bF, bG:
    the first snapshot's distance in segments (0: keep none)."""
    F = [[0] * (slots + 1) for _ in range(nseg + 1)]
    G = [[0] * (slots + 1) for _ in range(nseg + 1)]
    bF = [[0] * (slots + 1) for _ in range(nseg + 1)]
    bG = [[0] * (slots + 1) for _ in range(nseg + 1)]
    for n in range(2, nseg + 1):
        for f in range(slots + 1):
            best, bm = n * (n - 1) // 2, 0
            for m in range(1, n if f > 0 else 1):
                c = m + F[n - m][f - 1] + F[m][f]
                if c < best:
                    best, bm = c, m
            F[n][f], bF[n][f] = best, bm
            gbest, gm = F[n - 1][f], 0
            for m in range(1, n if f > 0 else 1):
                c = G[n - m][f - 1] + F[m][f]
                if c < gbest:
                    gbest, gm = c, m
            G[n][f], bG[n][f] = gbest, gm
    return F, G, bF, bG


def binom(n, k):
    b = 1
    for i in range(1, k + 1):
        b = b * (n - k + i) // i
    return b


def griewank(n, s):
    """Least forward steps to reverse n steps with s checkpoints, the first
    at the start (binomial checkpointing): r n - beta(s + 1, r - 1) for
    beta(s, r - 1) < n <= beta(s, r), beta(s, r) = binomial(s + r, s)."""
    r = 0
    while binom(s + r, s) < n:
        r += 1
    return r * n - binom(s + r, s + 1)


def expected_recomputed(T, C, max_steps, slots):
    """Forward steps the reverse sweep recomputes: the recording run keeps
    the right spine of G; every segment but the last is restored from the
    nearest kept window below it and advanced to (keeping windows by F),
    then recomputed with its tape."""
    nseg = (max_steps - 2 + C - 1) // C
    F, G, bF, bG = revolve_tables(nseg, slots)
    plan, cur, n = [], 0, nseg
    for f in range(slots, 0, -1):
        if n < 2 or bG[n][f] == 0:
            break
        cur += bG[n][f]
        n -= bG[n][f]
        plan.append(cur)
    held = [b for b in plan if 2 + b * C <= T]
    seg_first = lambda t: 3 + C * ((t - 3) // C)
    have_from, total = seg_first(T), 0
    for t in range(T, 2, -1):
        if t >= have_from:
            continue
        first = seg_first(t)
        k = (first - 3) // C
        held = [b for b in held if b <= k]
        j = max(held, default=0)
        nfree = slots - len(held)
        cur = j
        while cur < k:
            n = k - cur + 1
            bm = bF[n][nfree] if n >= 2 else 0
            m = cur + bm if bm > 0 else k
            total += (m - cur) * C
            cur = m
            if cur < k:
                held.append(cur)
                nfree -= 1
        total += t - first + 1
        have_from = first
    return total


def test_schedule_is_binomial_checkpointing():
    """The table F equals Griewank's closed form for binomial checkpointing
    (f free slots plus the held start: s = f + 1 checkpoints); with no
    slot every segment is advanced to from the start; with a slot per
    segment each is recomputed once; more slots never cost more."""
    F, G, _, _ = revolve_tables(60, 5)
    for n in range(1, 61):
        for f in range(6):
            assert F[n][f] == griewank(n, f + 1), (n, f)
    T, C = 42, 4  # 10 segments, the last one (steps 39 .. 42) taped while recording
    assert expected_recomputed(T, C, T, 0) == sum(k * C + C for k in range(9))
    assert expected_recomputed(T, C, T, 20) == (T - 2) - C
    costs = [expected_recomputed(T, C, T, s) for s in range(0, 12)]
    assert all(a >= b for a, b in zip(costs, costs[1:])), costs


def gradient(C, T, host_windows=None, max_steps=None):
    from qftkb import KBContext
    c = KBContext(*ARGS)
    n0 = thermal(c.kL2())
    c.init_particles(n0, 1.0)
    c.adj_record(C, max_steps or T, host_windows=host_windows)
    for t in range(3, T + 1):
        c.step(t)
    wgt = np.exp(-c.kL2())
    g = c.adj_gradient(wgt, n0=True, F0=True)
    info = c.adj_info()
    assert c.window() == (1, 2)
    return c, g, info


@pytest.mark.gpu
@pytest.mark.parametrize("slots,max_steps", [(0, 42), (1, 42), (2, 42), (3, 60), (12, 42)])
def test_host_snapshots_match_tape(slots, max_steps):
    T, C = 42, 4
    print(f"\n[case] host snapshots: 6^3 N_mem=8 T={T} C={C} slots={slots} max_steps={max_steps}", flush=True)
    c0, g0, _ = gradient(0, T)
    c1, g1, info = gradient(C, T, host_windows=slots, max_steps=max_steps)
    assert abs(g1["lambda"] - g0["lambda"]) <= TOL * abs(g0["lambda"])
    assert rel(g1["n0"], g0["n0"]) <= TOL
    assert rel(g1["F0"], g0["F0"]) <= TOL
    assert info["recomputed_steps"] == expected_recomputed(T, C, max_steps, slots)
    c0.close()
    c1.close()


@pytest.mark.gpu
def test_host_snapshots_two_gradients_in_a_row():
    """Recording continues after a gradient: a second forward run and sweep
    with the same snapshot slots give the same gradient."""
    from qftkb import KBContext
    T, C = 30, 3
    c = KBContext(*ARGS)
    n0 = thermal(c.kL2())
    c.init_particles(n0, 1.0)
    c.adj_record(C, T, host_windows=2)
    wgt = np.exp(-c.kL2())
    out = []
    for _ in range(2):
        for t in range(3, T + 1):
            c.step(t)
        out.append(c.adj_gradient(wgt, n0=True))
    assert abs(out[1]["lambda"] - out[0]["lambda"]) <= TOL * abs(out[0]["lambda"])
    assert rel(out[1]["n0"], out[0]["n0"]) <= TOL
    c.close()


@pytest.mark.gpu
def test_host_snapshots_reject_bad_arguments():
    from qftkb import KBContext, KBError
    c = KBContext(*ARGS)
    c.init_particles(thermal(c.kL2()), 1.0)
    with pytest.raises(KBError):
        c.adj_record(0, 20, host_windows=2)  # host windows need C > 0
    with pytest.raises(KBError):
        c.adj_record(4, 20, host_windows=-1)
    c.close()
