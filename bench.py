"""KB steps/s of the FP64 2PI Kadanoff-Baym step at 32^3, N_mem = 400 on one
MI355X, with the memory-integral kernel's HBM roofline fraction and the
oracle (CPU restatement of the reference algorithm) timed on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Set-up follows the reference's benchmark harness (TwoPIScalar_Bechmark.jl:
62-104): the memory window is filled first (here with the free propagators
of the thermal T = 2 state, so every value is finite and realistic), then W
untimed steps, then K timed steps.  Every timed step is a complete evolve!:
self-energy FFT stage, Hartree reduction, memory integrals, leapfrog of the
new row and its diagonal.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "qftdynamicsm.jl_amd"))

import numpy as np  # noqa: E402

WORKLOAD = dict(Nx=32, sdim=3, NstepsinMemory=400, dt=0.1, Lambda=1.0, ONgroup=1, Mass=1.0, T=2.0)
# differentiation workload: gradients of L = sum_p exp(-|p|^2) n(p, t_f)
ADJ_WORKLOAD = dict(Nx=24, sdim=3, NstepsinMemory=200, dt=0.1, Lambda=1.0, ONgroup=1, Mass=1.0, T=2.0,
                    steps=500, checkpoint_every=50, nbins=16, h=1e-5)
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak


def thermal_n(kL2, Mass, T):
    return 1.0 / (np.exp(np.sqrt(kL2 + Mass ** 2) / (T * Mass)) - 1.0)


def bytes_memint(N, V):
    """Algorithmic bytes of one memory-integral + leapfrog launch: the stored
    window of N-1 live times read once ((N-1)N F and rho values per site),
    SigF/Sigr of N-1 times, the new row of N F and rho values written."""
    return ((N - 1) * N + 2 * (N - 1) + 2 * N) * V * 8


def bytes_resum(N, V):
    """Algorithmic bytes of one 1/N resummation launch over a full window of
    L = N-1 live times: sweep 1 reads the Ir triangle without row T
    ((L-1)(L-2)/2 values per site), sweep 2 the F2kr2 triangle with its
    diagonal (L(L+1)/2), sweep 3 the Ir triangle with row T (L(L-1)/2); plus
    Fr row T, F2kr2 row T and Fr column tmin read, Ir row T and the (IF, Ir)
    ring written (6L)."""
    L = N - 1
    return ((L - 1) * (L - 2) // 2 + L * (L + 1) // 2 + L * (L - 1) // 2 + 6 * L) * V * 8


def load_pmc(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(w, budget_s=20.0, max_steps=12):
    """Oracle (C + OpenMP, reference double-read order) on a bounded sample:
    the same 32^3 window filled the same way, then 3 to max_steps timed KB
    steps within about budget_s; the rate is the mean, the spread is in
    `sample`."""
    from oracle.oracle import Oracle

    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    N = w["NstepsinMemory"]
    V = w["Nx"] ** w["sdim"]
    need = 2 * N * (N + 1) // 2 * V * 8
    avail = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable"):
                    avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    note = ""
    if avail is not None and avail < 1.25 * need:
        N = 200
        note = f"; host RAM {avail / 2**30:.0f} GiB < 1.25x the N_mem=400 window, timed at N_mem=200"
    o = Oracle(w["Nx"], w["sdim"], N, w["dt"], w["Lambda"], w["ONgroup"], w["Mass"], w["Mass"] ** 2,
               nthreads=nthreads)
    n = thermal_n(o.kL2(), w["Mass"], w["T"])
    o.fill_free_window(N, n, w["Mass"])
    times = []
    t = N + 1
    t0 = time.perf_counter()
    while len(times) < max_steps:
        a = time.perf_counter()
        o.evolve(t)
        times.append(time.perf_counter() - a)
        t += 1
        if len(times) >= 3 and (time.perf_counter() - t0) + times[-1] > budget_s:
            break
    o.close()
    per = sum(times) / len(times)
    return {"value": 1.0 / per, "unit": "KB steps/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle (CPU restatement of the reference CPUfull algorithm, double-read order), "
                      f"32^3 N_mem={N} window, mean of {len(times)} timed steps after fill "
                      f"({min(times):.3f}-{max(times):.3f} s per step, {1 / max(times):.3f}-"
                      f"{1 / min(times):.3f} KB steps/s){note}"}


def ninv_transform_count(n):
    """Full-lattice transforms of one 1/N self-energy stage in the reference's
    CPUfull code (calcIrs!, calcIFs!, calcSelfEnergyattp!: Evolution.jl:
    170-206, 354-436, 683-698) over a window of n times, by its loop bounds;
    the oracle restates those loops and counts the same number."""
    irs = 3 * (n - 1) + 2 * (n - 1) * (n - 2)
    ifs = 0
    for k in range(n):  # t' = tmin + k
        first = max(0, k - 1)
        beg = 1 if k == 0 else k
        ifs += 2 + 8 * first + 4 * max(0, (n - 1) - beg) + 4 + (4 if k > 0 else 0) + 1
    return irs + ifs + 4 * n


def ninv_cpu_baseline(w, nsample=40):
    """The oracle's restatement of the reference's CPUfull 1/N self-energy
    stage (threaded over the time sums like the reference's chunks) timed
    once on the same 32^3 lattice with an N_mem = nsample window, scaled to
    N_mem by the exact transform count.  The memory integrals, a fraction of
    a second per step at N_mem = 400, are left out."""
    from oracle.oracle import Oracle

    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    N = w["NstepsinMemory"]
    o = Oracle(w["Nx"], w["sdim"], nsample, w["dt"], w["Lambda"], w["ONgroup"], w["Mass"], w["Mass"] ** 2,
               pexp=3, nthreads=nthreads)
    o.fill_free_window(nsample, thermal_n(o.kL2(), w["Mass"], w["T"]), w["Mass"])
    o.expand()
    c0 = o.ninv_transforms()
    t0 = time.perf_counter()
    o.ninv_selfenergies(nsample)
    ts = time.perf_counter() - t0
    done = o.ninv_transforms() - c0
    o.close()
    assert done == ninv_transform_count(nsample - 1), (done, ninv_transform_count(nsample - 1))
    per = ts * ninv_transform_count(N - 1) / done
    return {"value": 1.0 / per, "unit": "KB steps/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle 1/N self-energy stage (CPUfull calcIrs!/calcIFs!/calcSelfEnergyattp!), "
                      f"32^3 lattice, N_mem={nsample} window: {done} lattice transforms in {ts:.2f} s, "
                      f"scaled to the {ninv_transform_count(N - 1)} transforms of N_mem={N}; memory "
                      f"integrals not included"}


ENSEMBLE_SEED = 20240611
ENSEMBLE_NX = (16, 16, 24, 24, 32, 32, 40, 40)


def ensemble_runs(seed=ENSEMBLE_SEED, nmem=300):
    """The 8-run uneven ensemble: lattices {16,16,24,24,32,32,40,40}^3, lambda ~ U[0.5, 2],
    box occupations n0 ~ U[1, 4], Q ~ U[0.5, 1.5], drawn from one PCG64 seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = rng.uniform(0.5, 2.0, len(ENSEMBLE_NX))
    n0 = rng.uniform(1.0, 4.0, len(ENSEMBLE_NX))
    Q = rng.uniform(0.5, 1.5, len(ENSEMBLE_NX))
    runs = [dict(Nx=nx, sdim=3, NstepsinMemory=nmem, dt=0.1, Lambda=float(l), ONgroup=1, Mass=1.0, Mass2=1.0)
            for nx, l in zip(ENSEMBLE_NX, lam)]
    return runs, [dict(n0=float(a), Q=float(b)) for a, b in zip(n0, Q)]


def ensemble_bench(args):
    from qftkb import KBContext, KBEnsemble, lattice

    runs, occ = ensemble_runs(nmem=args.nmem if args.nmem != WORKLOAD["NstepsinMemory"] else 300)
    if args.sym:  # box occupations are cubic-symmetric: every member in orbit storage
        runs = [dict(r, symmetric=True) for r in runs]
    N = runs[0]["NstepsinMemory"]
    steps, warm = args.steps, args.warmup
    # all runs at once
    ens = KBEnsemble(runs)
    for m, o in zip(ens.members, occ):
        m.fill_free_window(N, lattice.n_box(m.kL2(), o["n0"], o["Q"]), 1.0)
    ens.use_graph(True)
    t = N + 1
    for _ in range(warm):
        ens.step(t)
        t += 1
    # the pooled memory integral's shape, timed on the last step (then one
    # step to capture the graph again)
    tuned = ens.autotune(3)
    ens.step(t)
    t += 1
    ens.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step(t)
        t += 1
    ens.synchronize()
    t_ens = time.perf_counter() - t0
    for m in ens.members:
        m.check_finite()
    ens.close()
    print(f"[bench] ensemble: {steps} steps of {len(runs)} runs in {t_ens:.4f} s", file=sys.stderr, flush=True)
    # the same runs one after another, each a plain context (graph on too)
    t_seq = 0.0
    per_run = []
    for r, o in zip(runs, occ) if not args.no_sequential else ():
        c = KBContext(**r)
        c.fill_free_window(N, lattice.n_box(c.kL2(), o["n0"], o["Q"]), 1.0)
        c.use_graph(True)
        t = N + 1
        for _ in range(warm):
            c.step(t)
            t += 1
        c.synchronize()
        a = time.perf_counter()
        for _ in range(steps):
            c.step(t)
            t += 1
        c.synchronize()
        dt_run = time.perf_counter() - a
        per_run.append(dt_run / steps * 1e3)
        t_seq += dt_run
        c.close()
    line = {
        "metric": "ensemble KB steps/s (8 uneven runs, 16^3..40^3, N_mem=300)" +
                  (" with cubic-orbit storage (separate metric)" if args.sym else ""),
        "value": len(runs) * steps / t_ens, "unit": "member KB steps/s",
        "ensemble_steps_per_s": steps / t_ens,
        "sequential_member_steps_per_s": len(runs) * steps / t_seq if t_seq else None,
        "speedup_vs_sequential": t_seq / t_ens if t_seq else None,
        "ms_per_ensemble_step": 1e3 * t_ens / steps,
        "sequential_ms_per_step_by_run": dict(zip([f"{r['Nx']}^3 #{i}" for i, r in enumerate(runs)], per_run)),
        "ensemble_tiles": {k: tuned[k] for k in ("rows", "batch", "threads")},
        "ensemble_tile_ms": [round(x, 4) for x in tuned["ms"]],
        "steps": steps, "warmup": warm, "dtype": "f64",
        "config": {"workload": "8-run uneven ensemble {16,16,24,24,32,32,40,40}^3", "Nx": list(ENSEMBLE_NX), "NstepsinMemory": N,
                   "seed": ENSEMBLE_SEED, "Lambda": [r["Lambda"] for r in runs], "box": occ,
                   "storage": "cubic orbits" if args.sym else "every site"},
    }
    print(json.dumps(line), flush=True)


def radial_bins(kL2, sdim, nbins):
    edges = np.linspace(0.0, np.sqrt(4.0 * sdim), nbins + 1)
    return np.minimum(np.searchsorted(edges, np.sqrt(kL2), side="right") - 1, nbins - 1)


def adjoint_bench(fd_bins=False, fd=True):
    """Cost of dL/dlambda and dL/dn(p,0) by the adjoint kernels relative to
    one forward run, at 24^3, N_mem = 200, 500 steps, checkpoints every 50
    (and with every dropped column taped), plus the check against central
    finite differences (h = 1e-5) of the forward solver."""
    from qftkb import KBContext

    a = ADJ_WORKLOAD
    Nx, N, dt, lam, ON, M = a["Nx"], a["NstepsinMemory"], a["dt"], a["Lambda"], a["ONgroup"], a["Mass"]
    T = 2 + a["steps"]

    def make(lmb):
        return KBContext(Nx, a["sdim"], N, dt, lmb, ON, M, M * M)

    def forward(c, graph):
        c.use_graph(graph)
        for t in range(3, T + 1):
            c.step(t)
        c.synchronize()

    # load every kernel once
    w0 = KBContext(6, 3, 5, dt, lam, ON, M, M * M)
    w0.init_particles(np.full(216, 0.5), M)
    w0.adj_record(4, 12)
    for t in range(3, 13):
        w0.step(t)
    w0.adj_gradient(np.ones(216))
    w0.close()

    c = make(lam)
    k2 = c.kL2()
    n0 = thermal_n(k2, M, a["T"])
    wgt = np.exp(-k2)
    c.init_particles(n0, M)
    t0 = time.perf_counter()
    forward(c, True)
    t_fwd = time.perf_counter() - t0
    L = float((wgt * c.measure(T)["n"]).sum())
    c.close()
    print(f"[bench] adjoint workload: forward {a['steps']} steps in {t_fwd:.3f} s", file=sys.stderr, flush=True)
    out = {"workload": f"{Nx}^3 lattice, N_mem={N}, {a['steps']} steps, thermal T=2, lambda=1, "
                       f"L = sum_p exp(-|p|^2) n(p, t_f)",
           "forward_s": t_fwd, "loss": L}
    grads = {}
    for C in (a["checkpoint_every"], 0, -1):
        g = make(lam)
        g.init_particles(n0, M)
        g.adj_record(C, T)
        t0 = time.perf_counter()
        forward(g, False)
        t1 = time.perf_counter()
        gr = g.adj_gradient(wgt)
        t2 = time.perf_counter()
        info = g.adj_info()
        g.close()
        key = f"checkpoint_every_{C}" if C > 0 else ("tape_every_column" if C == 0 else "packed_tape")
        out[key] = {"record_forward_s": t1 - t0, "reverse_s": t2 - t1, "gradient_s": t2 - t0,
                    "ratio": (t2 - t0) / t_fwd, "reverse_ratio": (t2 - t1) / t_fwd,
                    "recomputed_steps": info["recomputed_steps"], "adjoint_device_gb": info["device_bytes"] / 1e9}
        grads[C] = gr
        print(f"[bench] adjoint C={C}: record {t1 - t0:.3f} s, reverse {t2 - t1:.3f} s, "
              f"ratio {(t2 - t0) / t_fwd:.2f}", file=sys.stderr, flush=True)
    gr = grads[a["checkpoint_every"]]
    out["ratio"] = out[f"checkpoint_every_{a['checkpoint_every']}"]["ratio"]
    out["dL_dlambda"] = gr["lambda"]
    bins = radial_bins(k2, a["sdim"], a["nbins"])
    out["dL_dn0_bins"] = [float(gr["n0"][bins == b].sum()) for b in range(a["nbins"])]
    out["tape_vs_checkpoint_max_rel"] = float(np.max(np.abs(grads[0]["n0"] - gr["n0"])) / np.max(np.abs(gr["n0"])))

    if not fd:
        return out
    h = a["h"]

    def loss(lmb, n):
        c = make(lmb)
        c.init_particles(n, M)
        forward(c, True)
        v = float((wgt * c.measure(T)["n"]).sum())
        c.close()
        return v

    t0 = time.perf_counter()
    fd = (loss(lam + h, n0) - loss(lam - h, n0)) / (2 * h)
    out["fd_dL_dlambda"] = fd
    out["fd_rel_err_lambda"] = abs(gr["lambda"] - fd) / abs(fd)
    if fd_bins:
        fdb = []
        for b in range(a["nbins"]):
            e = np.where(bins == b, h, 0.0)
            fdb.append((loss(lam, n0 + e) - loss(lam, n0 - e)) / (2 * h) if np.any(bins == b) else 0.0)
        scale = max(abs(x) for x in fdb)
        out["fd_dL_dn0_bins"] = fdb
        out["fd_rel_err_bins"] = max(abs(x - y) for x, y in zip(out["dL_dn0_bins"], fdb)) / scale
    out["fd_s"] = time.perf_counter() - t0
    out["fd_cost_ratio"] = 2 * (1 + (a["nbins"] if fd_bins else 0)) * 1.0  # forward runs per gradient, FD
    gold = load_pmc(os.path.join(ROOT, "tests", "golden", "adjoint_24c_n200_fd.json"))
    if gold and "dL_dlambda_fd" in gold:
        out["oracle_fd_dL_dlambda"] = gold["dL_dlambda_fd"]
        out["oracle_fd_rel_err_lambda"] = abs(gr["lambda"] - gold["dL_dlambda_fd"]) / abs(gold["dL_dlambda_fd"])
    return out


# finite differences: central differences D(h) at h and 2h, Richardson R = (4 D(h) - D(2h)) / 3.  At
# 4000 steps L ~ 4.5e3 carries ~1e-14 L of rounding that is not smooth in lambda, so D(1e-5) is
# noise-limited to ~1e-6; at h = 5e-4 the noise is ~2e-8 and R removes the h^2 term.
LARGE_ADJ_WORKLOAD = dict(Nx=48, sdim=3, NstepsinMemory=300, dt=0.05, Lambda=0.5, ONgroup=1, Mass=1.0, n0=4.0,
                          Q=1.0, steps=4000, h=5e-4, nbins=16)


def large_adjoint_bench(fd=True):
    """The largest single-GPU differentiation: 48^3, N_mem=300 (80.7 GB
    window), dt=0.05, 4000 steps, box n0=4, Q=1, lambda=0.5; L = sum_p
    exp(-|p|^2) n(p, t_f).  Forward steps/s and the memory integral's HBM
    fraction; the gradient with the orbit-packed tape (the box state is
    cubic-symmetric: the 3702 dropped columns take 52 GB instead of 1.96 TB),
    its cost ratio, and GPU central differences of dL/dlambda."""
    from qftkb import KBContext, lattice

    a = LARGE_ADJ_WORKLOAD
    Nx, N, dt, lam, ON, M = a["Nx"], a["NstepsinMemory"], a["dt"], a["Lambda"], a["ONgroup"], a["Mass"]
    T = 2 + a["steps"]

    def make(lmb):
        return KBContext(Nx, a["sdim"], N, dt, lmb, ON, M, M * M)

    def forward(c, graph, label):
        c.use_graph(graph)
        t0 = time.perf_counter()
        for t in range(3, T + 1):
            c.step(t)
            if (t - 2) % 500 == 0:
                c.synchronize()
                print(f"[bench] {label}: step {t}/{T}, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                      flush=True)
        c.synchronize()
        return time.perf_counter() - t0

    c = make(lam)
    k2 = c.kL2()
    n0 = lattice.n_box(k2, a["n0"], a["Q"])
    wgt = np.exp(-k2)
    c.init_particles(n0, M)
    t_fwd = forward(c, True, "forward")
    L = float((wgt * c.measure(T)["n"]).sum())
    # memory-integral time on the full window (60 more steps, timed per kernel)
    c.enable_timing(True)
    c.reset_timing()
    for t in range(T + 1, T + 61):
        c.step(t)
    c.synchronize()
    ms_mi, n_mi = c.timing(0)
    c.enable_timing(False)
    c.close()
    avg = ms_mi / max(n_mi, 1) / 1e3
    B = bytes_memint(N, Nx ** 3)
    out = {"workload": f"{Nx}^3 lattice, N_mem={N}, dt={dt}, {a['steps']} steps, box n0={a['n0']} Q={a['Q']}, "
                       f"lambda={lam}, L = sum_p exp(-|p|^2) n(p, t_f)",
           "forward_s": t_fwd, "forward_steps_per_s": a["steps"] / t_fwd, "loss": L,
           "memint_ms": avg * 1e3, "memint_hbm_frac": B / avg / 1e9 / HBM_PEAK_GBS}
    g = make(lam)
    g.init_particles(n0, M)
    g.adj_record(-1, T)
    t_rec = forward(g, False, "recording forward")
    t0 = time.perf_counter()
    gr = g.adj_gradient(wgt)
    t_rev = time.perf_counter() - t0
    info = g.adj_info()
    g.close()
    print(f"[bench] packed-tape gradient: record {t_rec:.1f} s, reverse {t_rev:.1f} s", file=sys.stderr, flush=True)
    out["packed_tape"] = {"record_forward_s": t_rec, "reverse_s": t_rev, "gradient_s": t_rec + t_rev,
                          "ratio": (t_rec + t_rev) / t_fwd, "adjoint_device_gb": info["device_bytes"] / 1e9,
                          "recomputed_steps": info["recomputed_steps"]}
    out["ratio"] = out["packed_tape"]["ratio"]
    out["dL_dlambda"] = gr["lambda"]
    bins = radial_bins(k2, a["sdim"], a["nbins"])
    out["dL_dn0_bins"] = [float(gr["n0"][bins == b].sum()) for b in range(a["nbins"])]
    if fd:
        h = a["h"]
        D = {}
        for hh in (h, 2 * h):
            Ls = []
            for sgn in (1, -1):
                f = make(lam + sgn * hh)
                f.init_particles(n0, M)
                forward(f, True, f"finite difference {'+' if sgn > 0 else '-'}{hh:g}")
                Ls.append(float((wgt * f.measure(T)["n"]).sum()))
                f.close()
            D[hh] = (Ls[0] - Ls[1]) / (2 * hh)
        fdv = (4 * D[h] - D[2 * h]) / 3
        out["fd_central"] = {f"{hh:g}": v for hh, v in D.items()}
        out["fd_dL_dlambda"] = fdv
        out["fd_method"] = f"GPU forward runs, central differences at h={h:g} and {2 * h:g}, Richardson"
        out["fd_rel_err_lambda"] = abs(gr["lambda"] - fdv) / abs(fdv)
    return out


def large_adjoint_host_bench(C=150, host_windows=2, fd=True):
    """The same 48^3 differentiation for a state without the cubic symmetry
    (the box occupation modulated along x: n(p) = box(p) (1 + 0.2 sin(2 pi
    i_x / Nx + 0.3))), so the orbit-packed tape does not apply: the window
    (80.7 GB) is kept at segment boundaries in host memory (at most
    host_windows of them) under the binomial checkpointing schedule, and
    the device tape holds one C-step segment (kb_adj_record_host).  Forward
    steps/s, the gradient's cost ratio, the recomputed steps, and dL/dlambda
    against GPU central differences (h and 2h, Richardson)."""
    from qftkb import KBContext, lattice

    a = LARGE_ADJ_WORKLOAD
    Nx, N, dt, lam, ON, M = a["Nx"], a["NstepsinMemory"], a["dt"], a["Lambda"], a["ONgroup"], a["Mass"]
    T = 2 + a["steps"]

    def make(lmb):
        return KBContext(Nx, a["sdim"], N, dt, lmb, ON, M, M * M)

    def forward(c, graph, label):
        c.use_graph(graph)
        t0 = time.perf_counter()
        for t in range(3, T + 1):
            c.step(t)
            if (t - 2) % 500 == 0:
                c.synchronize()
                print(f"[bench] {label}: step {t}/{T}, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                      flush=True)
        c.synchronize()
        return time.perf_counter() - t0

    c = make(lam)
    k2 = c.kL2()
    ix = np.arange(Nx ** 3) % Nx
    n0 = lattice.n_box(k2, a["n0"], a["Q"]) * (1.0 + 0.2 * np.sin(2 * np.pi * ix / Nx + 0.3))
    wgt = np.exp(-k2)
    c.init_particles(n0, M)
    t_fwd = forward(c, True, "forward")
    L = float((wgt * c.measure(T)["n"]).sum())
    c.close()
    out = {"workload": f"{Nx}^3 lattice, N_mem={N}, dt={dt}, {a['steps']} steps, box n0={a['n0']} Q={a['Q']} "
                       f"modulated by 1 + 0.2 sin(2 pi i_x / {Nx} + 0.3) (no cubic symmetry), lambda={lam}, "
                       f"L = sum_p exp(-|p|^2) n(p, t_f)",
           "forward_s": t_fwd, "forward_steps_per_s": a["steps"] / t_fwd, "loss": L}
    g = make(lam)
    g.init_particles(n0, M)
    g.adj_record(C, T, host_windows=host_windows)
    t_rec = forward(g, False, "recording forward")
    t0 = time.perf_counter()
    gr = g.adj_gradient(wgt)
    t_rev = time.perf_counter() - t0
    info = g.adj_info()
    g.close()
    J = (N + 3) // 4  # column blocks of 4 slots; the history holds their slot pairs per 8-site block
    window_gb = (Nx ** 3 + 7) // 8 * (4 * (N * J - 4 * J * (J - 1) // 2)) * 8 * 16 / 1e9
    print(f"[bench] host-checkpointed gradient: record {t_rec:.1f} s, reverse with recomputation {t_rev:.1f} s, "
          f"{info['recomputed_steps']} steps recomputed", file=sys.stderr, flush=True)
    out["host_checkpoints"] = {"checkpoint_every": C, "host_windows": host_windows,
                               "host_gb": host_windows * window_gb, "record_forward_s": t_rec,
                               "reverse_s": t_rev, "gradient_s": t_rec + t_rev, "ratio": (t_rec + t_rev) / t_fwd,
                               "adjoint_device_gb": info["device_bytes"] / 1e9,
                               "recomputed_steps": info["recomputed_steps"]}
    out["ratio"] = out["host_checkpoints"]["ratio"]
    out["dL_dlambda"] = gr["lambda"]
    bins = radial_bins(k2, a["sdim"], a["nbins"])
    out["dL_dn0_bins"] = [float(gr["n0"][bins == b].sum()) for b in range(a["nbins"])]
    if fd:
        h = a["h"]
        D = {}
        for hh in (h, 2 * h):
            Ls = []
            for sgn in (1, -1):
                f = make(lam + sgn * hh)
                f.init_particles(n0, M)
                forward(f, True, f"finite difference {'+' if sgn > 0 else '-'}{hh:g}")
                Ls.append(float((wgt * f.measure(T)["n"]).sum()))
                f.close()
            D[hh] = (Ls[0] - Ls[1]) / (2 * hh)
        fdv = (4 * D[h] - D[2 * h]) / 3
        out["fd_central"] = {f"{hh:g}": v for hh, v in D.items()}
        out["fd_dL_dlambda"] = fdv
        out["fd_method"] = f"GPU forward runs, central differences at h={h:g} and {2 * h:g}, Richardson"
        out["fd_rel_err_lambda"] = abs(gr["lambda"] - fdv) / abs(fdv)
    return out


AUTOTUNE_WORKLOADS = [(32, 3, 400), (24, 3, 200), (48, 3, 300), (16, 3, 202), (40, 3, 300),
                      (16, 3, 300), (24, 3, 300), (32, 3, 300)]


def autotune_bench(out_path):
    """Pick the memory-integral tile shape per lattice and window: fill the
    window, take two steps, re-time the last step's integral in every shape."""
    from qftkb import KBContext
    from qftkb._abi import tiles_key

    w = WORKLOAD
    res = {"method": "kb_autotune: 3 timed launches per tile shape on a filled window (HIP events)",
           "choices": {}}
    for Nx, sd, N in AUTOTUNE_WORKLOADS:
        c = KBContext(Nx, sd, N, w["dt"], w["Lambda"], w["ONgroup"], w["Mass"], w["Mass"] ** 2, tiles=None)
        c.fill_free_window(N, thermal_n(c.kL2(), w["Mass"], w["T"]), w["Mass"])
        c.step(N + 1)
        c.step(N + 2)
        r = c.autotune(reps=3)
        r["shapes"] = c.tile_shapes()
        r["fft"] = c.autotune_fft(reps=3)
        r["fft"]["shapes"] = c.fft_shapes()
        res["choices"][tiles_key(Nx, sd, N)] = r
        print(f"[bench] autotune {tiles_key(Nx, sd, N)}: rows {r['rows']} batch {r['batch']} "
              f"threads {r['threads']} "
              f"ms {[round(x, 3) for x in r['ms']]}; fft threads {r['fft']['threads']} "
              f"ms {[round(x, 3) for x in r['fft']['ms']]}", file=sys.stderr, flush=True)
        c.close()
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--nmem", type=int, default=WORKLOAD["NstepsinMemory"])
    ap.add_argument("--nx", type=int, default=WORKLOAD["Nx"])
    ap.add_argument("--ensemble", action="store_true",
                    help="8 uneven runs (16^3..40^3) advanced as one ensemble vs the same runs one after another")
    ap.add_argument("--no-sequential", action="store_true",
                    help="with --ensemble: skip stepping the runs one after another")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost-ratio leg")
    ap.add_argument("--sym", action="store_true",
                    help="cubic-orbit storage (a separate metric: the stored window shrinks ~34x at 32^3)")
    ap.add_argument("--adjoint", action="store_true",
                    help="only the differentiation workload, with finite differences of all 16 radial bins")
    ap.add_argument("--no-fd", action="store_true", help="with --adjoint: skip the finite differences")
    ap.add_argument("--adjoint-large", action="store_true",
                    help="only the 48^3 N_mem=300 4000-step differentiation (packed tape; about 6 minutes)")
    ap.add_argument("--adjoint-large-host", action="store_true",
                    help="the 48^3 differentiation for a state without cubic symmetry: window snapshots in "
                         "pinned host memory under the binomial checkpointing schedule (about 10 minutes)")
    ap.add_argument("--host-windows", type=int, default=2,
                    help="with --adjoint-large-host: window snapshots kept in pinned host memory")
    ap.add_argument("--checkpoint-every", type=int, default=150,
                    help="with --adjoint-large-host: steps per segment (the device tape holds one)")
    ap.add_argument("--autotune", metavar="OUT_JSON",
                    help="pick the memory-integral tile shape for the standard workloads and write the choices")
    ap.add_argument("--ninverse", action="store_true",
                    help="the same workload in the 1/N expansion at NLO (a separate metric; the roofline "
                         "line is the resummation kernel's)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 2 GiB device copy (torch) to scale rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.full((1 << 28,), 1.0, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        # streaming-read rate of a library reduction (fp64 sum over 17 GB), the
        # practical ceiling the memory-integral kernel is compared with
        z = torch.full((1 << 31,), 1.0, dtype=torch.float64, device="cuda")
        for _ in range(2):
            z.sum()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(10):
            z.sum()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        from qftkb._abi import read_bandwidth
        print(json.dumps({"calibration": "copy", "bytes_read": x.numel() * 8, "bytes_written": y.numel() * 8,
                          "sum_read_GBs": z.numel() * 8 / ms / 1e6,
                          "streaming_read_GBs": read_bandwidth(0, 16 << 30)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus N > 1 is not supported "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)
    if args.ensemble:
        ensemble_bench(args)
        return
    if args.adjoint:
        print(json.dumps({"adjoint": adjoint_bench(fd_bins=True, fd=not args.no_fd)}), flush=True)
        return
    if args.adjoint_large:
        print(json.dumps({"adjoint_large": large_adjoint_bench(fd=not args.no_fd)}), flush=True)
        return
    if args.adjoint_large_host:
        r = large_adjoint_host_bench(C=args.checkpoint_every, host_windows=args.host_windows, fd=not args.no_fd)
        print(json.dumps({"adjoint_large_host": r}), flush=True)
        return
    if args.autotune:
        print(json.dumps({"autotune": autotune_bench(args.autotune)}), flush=True)
        return

    from qftkb import KBContext

    w = dict(WORKLOAD)
    w["NstepsinMemory"] = args.nmem
    w["Nx"] = args.nx
    N = w["NstepsinMemory"]
    V = w["Nx"] ** w["sdim"]
    pexp = 3 if args.ninverse else 1
    ctx = KBContext(w["Nx"], w["sdim"], N, w["dt"], w["Lambda"], w["ONgroup"], w["Mass"], w["Mass"] ** 2,
                    symmetric=args.sym, pexp=pexp)
    Vs = ctx.storage_info()[0]  # stored sites: V, or the cubic orbits
    ft = ctx.fft_tiles()
    ctx_orbit_product = ft < 0
    tile_shape = list(ctx.tile_shapes()[ctx.tiles()]) + ["orbit product" if ft < 0 else ctx.fft_shapes()[ft]]
    n = thermal_n(ctx.kL2(), w["Mass"], w["T"])
    ctx.fill_free_window(N, n, w["Mass"])
    print(f"[bench] window filled: {w['Nx']}^3 N_mem={N}, device bytes {ctx.device_bytes() / 1e9:.2f} GB",
          file=sys.stderr, flush=True)
    t = N + 1
    for _ in range(args.warmup):
        ctx.step(t)
        t += 1
    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.step(t)
        t += 1
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ms_mi, n_mi = ctx.timing(0)
    ms_fft, n_fft = ctx.timing(1)
    ms_h, _ = ctx.timing(2)
    ms_rs, n_rs = ctx.timing(5)
    ctx.enable_timing(False)
    ctx.check_finite()
    print(f"[bench] {args.steps} steps in {elapsed:.4f} s; memint {ms_mi / max(n_mi, 1):.3f} ms, "
          f"sigma {ms_fft / max(n_fft, 1):.3f} ms per step", file=sys.stderr, flush=True)
    ctx.close()

    steps_per_s = args.steps / elapsed
    avg_mi_s = ms_mi / max(n_mi, 1) / 1e3
    B = bytes_memint(N, Vs)
    achieved = B / avg_mi_s / 1e9
    pmc = load_pmc(os.path.join(ROOT, "profiles", "memint_pmc_32c_n400.json"))
    traffic = None
    if pmc and pmc.get("Nx") == w["Nx"] and pmc.get("N") == N and not args.sym:
        traffic = pmc.get("traffic_bytes")
    no_cpu = args.no_cpu
    if args.sym or args.ninverse:
        args.no_adjoint = args.no_cpu = True
    if args.ninverse:
        traffic = None
        pmc_n = load_pmc(os.path.join(ROOT, "profiles", "ninv_pmc_32c_n400.json"))
        if pmc_n and pmc_n.get("Nx") == w["Nx"] and pmc_n.get("N") == N and not args.sym:
            traffic = pmc_n.get("traffic_bytes")
    size = f"{w['Nx']}^3 N_mem={N}"
    headline = (w["Nx"], N) == (WORKLOAD["Nx"], WORKLOAD["NstepsinMemory"])
    line = {
        "metric": ("KB steps/s at 32^3 N_mem=400 FP64 (% HBM peak); adjoint/forward cost ratio"
                   if headline and not (args.sym or args.ninverse) else
                   f"KB steps/s at {size} FP64, 1/N expansion at NLO (separate metric)" if args.ninverse else
                   f"KB steps/s at {size} FP64 with cubic-orbit storage (separate metric)" if args.sym else
                   f"KB steps/s at {size} FP64 (separate workload)"),
        "value": steps_per_s,
        "unit": "KB steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (window filled with free thermal propagators, as the reference benchmark fills "
                "its window without physics)",
        "config": {"workload": f"2PI KB step, {'NinverseNLO bubble chain' if args.ninverse else 'LambdaNLO sunset'}, "
                               f"phi^4 O(1), {w['Nx']}^3 lattice, N_mem={N}, thermal T=2, lambda=1, m=1, dt=0.1",
                   "Nx": w["Nx"], "sdim": 3, "NstepsinMemory": N, "modes": V, "stored_sites": Vs,
                   "storage": "cubic orbits" if args.sym else "every site"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "memint_step", "algorithmic_bytes": B,
                     "avg_kernel_ms": avg_mi_s * 1e3, "sigma_stage_ms": ms_fft / max(n_fft, 1),
                     "tiles_rows_batch_threads_fftthreads": tile_shape,
                     "hartree_ms": ms_h / max(n_mi, 1),
                     # the whole step: memory-integral bytes plus a fused self-energy's
                     # 4 (N-1) V doubles, over the wall time per step
                     "step_achieved": (B + 4 * (N - 1) * Vs * 8) / (elapsed / args.steps) / 1e9},
        "cpu_baseline": None,
        "adjoint": None,
    }
    if args.sym and not args.ninverse and ctx_orbit_product:
        # the self-energy stage as orbit products: 2 products of 2 V^2 (2L) useful flops each,
        # over the stage's HIP-event time (which includes the 5 us gather)
        stage_s = ms_fft / max(n_fft, 1) / 1e3
        flops = 2 * 2 * Vs * Vs * 2 * (N - 1)
        line["roofline_self_energy"] = {"bound": "mfma", "achieved": flops / stage_s / 1e12,
                                        "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                                        "frac": flops / stage_s / 1e12 / FP64_MFMA_PEAK_TFLOPS,
                                        "flops": flops, "stage_ms": stage_s * 1e3,
                                        "kernels": "orbit_gather + orbit_gemm<0> + orbit_gemm<3>"}
    if args.ninverse:
        avg_rs_s = ms_rs / max(n_rs, 1) / 1e3
        Br = bytes_resum(N, Vs)
        ach = Br / avg_rs_s / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                            # orbit storage splits the sweeps (3 partial + 3 finish kernels)
                            "kernel": "ninv_split_sweep<1..3> + ninv_split_finish<1..3>" if args.sym
                            else "ninv_resum_kernel",
                            "algorithmic_bytes": Br, "avg_kernel_ms": avg_rs_s * 1e3,
                            "memint_ms": avg_mi_s * 1e3, "memint_frac": achieved / HBM_PEAK_GBS,
                            "self_energy_stage_ms": ms_fft / max(n_fft, 1),
                            "fft_passes_ms": ms_fft / max(n_fft, 1) - avg_rs_s * 1e3,
                            "hartree_ms": ms_h / max(n_mi, 1)}
        if not no_cpu:
            line["cpu_baseline"] = ninv_cpu_baseline(w)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
