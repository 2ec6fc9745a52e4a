"""Output records of a run: the measurement array, the problem, and the
energy-drift curve.

The reference ends a run by serialising `thesolution.measurearray` and
`thesolution.problem` with Julia's `Serialization` into
`<datapath>/Measurearray.jld` and `<datapath>/Problem.jld`
(G_SimulationArgs.jl:116-119), where datapath is the run directory named by
getParameterstring under the data root of `etc/paths.txt`
(DefAbstractSimSetup.jl:64-80, TwoPIScalar_Modelfile.jl:15-44).  Julia's
serialisation format is not readable outside Julia, so the same content is
written here in portable containers:

  Measurearray.npz  one entry per measurement k (measurearray[k]): the
                    scalar fields as length-K arrays (time, E, E_pot, E_kin,
                    E_mem, M2, n_unstable) and the per-site / history
                    fields as K x ... arrays (F, ddF, r, n, omega, c, Fonet,
                    ronet, SigF, Sigr; history fields padded with NaN where a
                    window was shorter);
  Problem.json      the parameter Dict, the parameter string and the
                    SimSetup fields;
  EnergyDrift.csv   k (measurement index), time, E, E - E(first),
                    (E - E(first)) / |E(first)|, E_kin, E_pot, E_mem, M2 per
                    measurement: the energy-drift curve a run is compared on.

Nothing here touches the device.
"""
from __future__ import annotations

import csv
import json
import os
from dataclasses import fields

import numpy as np

from .problem import MeasurementTwoPIScalar

SCALARS = ("time", "E", "E_pot", "E_kin", "E_mem", "M2", "phi", "n_unstable")
SITE_FIELDS = ("F", "ddF", "r", "n", "omega", "c")
HISTORY_FIELDS = ("Fonet", "ronet", "SigF", "Sigr")


def read_paths(path: str = "etc/paths.txt") -> dict:
    """Data, post and plot roots from a paths file in the reference's format
    (line 1 a comment, then `data#<path>`, `post#<path>`, `plot#<path>`;
    DefAbstractSimSetup.jl:68-71 drops the first five characters)."""
    with open(path) as f:
        lines = f.read().splitlines()
    return {"data": lines[1][5:], "post": lines[2][5:], "plot": lines[3][5:]}


def run_paths(simsetup, roots: dict) -> dict:
    """Run directories <root>/<parameterstring> (DefAbstractSimSetup.jl:73-76)."""
    ps = simsetup.parameterstring
    return {k: os.path.join(v, ps) for k, v in roots.items()}


def _stack(records, name):
    vals = [getattr(m, name) for m in records]
    if all(v is None for v in vals):
        return None
    n = max(0 if v is None else len(v) for v in vals)
    out = np.full((len(vals), n), np.nan)
    for i, v in enumerate(vals):
        if v is not None:
            out[i, :len(v)] = v
    return out


def save_measurearray(path: str, measurearray) -> None:
    """Measurearray.npz; entries that were never measured are skipped and
    their indices recorded in `k`."""
    ks = [k for k, m in enumerate(measurearray) if m is not None]
    recs = [measurearray[k] for k in ks]
    arrays = {"k": np.asarray(ks, dtype=np.int64), "length": np.int64(len(measurearray))}
    for name in SCALARS:
        arrays[name] = np.asarray([getattr(m, name) for m in recs], dtype=float)
    for name in SITE_FIELDS + HISTORY_FIELDS:
        a = _stack(recs, name)
        if a is not None:
            arrays[name] = a
    np.savez(path, **arrays)


def load_measurearray(path: str) -> list:
    """The list saved by save_measurearray (None where nothing was measured)."""
    with np.load(path, allow_pickle=False) as z:
        out = [None] * int(z["length"])
        names = set(z.files)
        for i, k in enumerate(z["k"]):
            kw = {}
            for name in SCALARS:
                v = z[name][i]
                kw[name] = int(v) if name == "n_unstable" else float(v)
            for name in SITE_FIELDS + HISTORY_FIELDS:
                if name in names:
                    row = z[name][i]
                    kw[name] = row[~np.isnan(row)] if name in HISTORY_FIELDS else row.copy()
            out[int(k)] = MeasurementTwoPIScalar(**kw)
    return out


def energy_drift(measurearray):
    """Rows (k, time, E, E - E0, (E - E0)/|E0|, E_kin, E_pot, E_mem, M2) with
    E0 the first measured energy."""
    recs = [(k, m) for k, m in enumerate(measurearray) if m is not None]
    if not recs:
        return []
    E0 = recs[0][1].E
    return [(k, m.time, m.E, m.E - E0, (m.E - E0) / abs(E0) if E0 != 0 else float("nan"), m.E_kin, m.E_pot,
             m.E_mem, m.M2) for k, m in recs]


def save_energy_drift(path: str, measurearray) -> None:
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["k", "time", "E", "dE", "dE_rel", "E_kin", "E_pot", "E_mem", "M2"])
        for row in energy_drift(measurearray):
            w.writerow([row[0]] + [repr(float(x)) for x in row[1:]])


def save_problem(path: str, problem, parameters: dict) -> None:
    ss = problem.simsetup
    rec = {"parameters": parameters, "parameterstring": ss.parameterstring,
           "simsetup": {f.name: getattr(ss, f.name) for f in fields(ss)}}
    with open(path, "w") as f:
        json.dump(rec, f, indent=1, default=lambda o: o.item() if hasattr(o, "item") else str(o))


def store_solution(solution, parameters: dict, datapath: str | None = None) -> dict:
    """Write Measurearray.npz, Problem.json and EnergyDrift.csv into the run
    directory (default: solution.problem.simsetup.datapath), as the end of
    G_SimulationArgs.jl does with its two .jld files.  Returns the paths."""
    d = datapath or solution.problem.simsetup.datapath
    if not d:
        raise ValueError("no run directory: pass datapath or set simsetup.datapath")
    os.makedirs(d, exist_ok=True)
    paths = {"measurearray": os.path.join(d, "Measurearray.npz"), "problem": os.path.join(d, "Problem.json"),
             "energy_drift": os.path.join(d, "EnergyDrift.csv")}
    save_measurearray(paths["measurearray"], solution.measurearray)
    save_problem(paths["problem"], solution.problem, parameters)
    save_energy_drift(paths["energy_drift"], solution.measurearray)
    return paths

