"""Output records (measurement array, problem, energy-drift curve): round
trips on the CPU, and a short GPU run whose stored energy-drift curve
matches the oracle's energies."""
import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

from qftkb import MeasurementTwoPIScalar, SimSetup, output


def fake_records(rng):
    recs = [None] * 6
    for k in (0, 2, 3, 5):
        L = 3 + k
        recs[k] = MeasurementTwoPIScalar(
            time=0.1 * k, E=10.0 + 1e-3 * k, E_pot=4.0, E_kin=5.0, E_mem=1e-3 * k, M2=1.2, n_unstable=k % 2,
            F=rng.standard_normal(8), ddF=rng.standard_normal(8), r=np.zeros(8), n=rng.random(8),
            omega=rng.random(8), c=rng.random(8), Fonet=rng.standard_normal(L), ronet=rng.standard_normal(L),
            SigF=rng.standard_normal(L), Sigr=rng.standard_normal(L))
    return recs


def test_measurearray_round_trip(tmp_path):
    recs = fake_records(np.random.default_rng(5))
    p = tmp_path / "Measurearray.npz"
    output.save_measurearray(str(p), recs)
    back = output.load_measurearray(str(p))
    assert len(back) == len(recs)
    for a, b in zip(recs, back):
        if a is None:
            assert b is None
            continue
        for name in output.SCALARS:
            assert getattr(a, name) == getattr(b, name)
        for name in output.SITE_FIELDS + output.HISTORY_FIELDS:
            np.testing.assert_array_equal(getattr(a, name), getattr(b, name))


def test_energy_drift_rows_and_file(tmp_path):
    recs = fake_records(np.random.default_rng(6))
    rows = output.energy_drift(recs)
    assert [r[0] for r in rows] == [0, 2, 3, 5]
    assert rows[0][3] == 0.0 and rows[2][3] == pytest.approx(3e-3)
    assert rows[3][4] == pytest.approx(5e-3 / 10.0)
    p = tmp_path / "EnergyDrift.csv"
    output.save_energy_drift(str(p), recs)
    lines = p.read_text().splitlines()
    assert lines[0].split(",")[:4] == ["k", "time", "E", "dE"]
    assert float(lines[2].split(",")[2]) == recs[2].E


def test_paths_file_and_run_dirs(tmp_path):
    # the reference's etc/paths.txt format: a comment, then data#/post#/plot#
    f = tmp_path / "paths.txt"
    f.write_text("# directories\ndata#/scratch/data\npost#/scratch/post\nplot#/scratch/plot\n")
    roots = output.read_paths(str(f))
    assert roots == {"data": "/scratch/data", "post": "/scratch/post", "plot": "/scratch/plot"}
    ss = SimSetup(parameterstring="phi4LambdaNLO_L100_ON1_RNone_T20_Nx8_sdim3")
    assert output.run_paths(ss, roots)["data"] == "/scratch/data/phi4LambdaNLO_L100_ON1_RNone_T20_Nx8_sdim3"


def test_store_solution_writes_three_files(tmp_path):
    recs = fake_records(np.random.default_rng(7))
    ss = SimSetup(parameterstring="x", datapath=str(tmp_path / "run"), Nsteps=10, Nmeas=5, NstepsbetweenNmeas=2)
    sol = SimpleNamespace(problem=SimpleNamespace(simsetup=ss), measurearray=recs)
    paths = output.store_solution(sol, {"Nx": 8, "Lambda": 1.0})
    for p in paths.values():
        assert os.path.exists(p)
    rec = json.load(open(paths["problem"]))
    assert rec["parameters"]["Nx"] == 8 and rec["simsetup"]["NstepsbetweenNmeas"] == 2


@pytest.mark.gpu
def test_gpu_run_records_match_oracle(tmp_path):
    """A short run through the reference-shaped driver loop: the stored
    energy-drift curve agrees with the oracle's energies (1e-10)."""
    from oracle.oracle import Oracle
    from qftkb import TwoPIScalarfile, run
    print("\n[case] stored run 6^3, 24 steps, measured every 4", flush=True)
    parameters = {"Mass": 1.0, "Nx": 6, "sdim": 3, "dt": 0.1, "Nsteps": 24, "NstepsinMemory": 9, "Nmeas": 6,
                  "Mod": "phi4", "Lambda": 1.0, "ONgroup": 1, "init": "Thermal", "T": 2.0, "Reno": "RNone",
                  "Pexp": "LambdaNLO", "Num": "GPUfull", "Nchunks": 0, "tag": "", "ov": True}
    sol = run(TwoPIScalarfile(), dict(parameters))
    paths = output.store_solution(sol, parameters, datapath=str(tmp_path / "run"))
    back = output.load_measurearray(paths["measurearray"])
    o = Oracle(6, 3, 9, 0.1, 1.0, 1, 1.0, 1.0)
    k2 = o.kL2()
    o.init_particles(1.0 / (np.exp(np.sqrt(k2 + 1.0) / 2.0) - 1.0), 1.0)
    E = {2: o.measure(2)["E"]}
    for t in range(3, 27):
        o.evolve(t)
        if (t - 2) % 4 == 0:
            E[t] = o.measure(t)["E"]
    for k, m in enumerate(back):
        t = 2 + 4 * k
        assert m is not None
        assert abs(m.E - E[t]) <= 1e-10 * abs(E[t]), (t, m.E, E[t])
    rows = output.energy_drift(back)
    assert abs(rows[-1][3] - (E[26] - E[2])) <= 1e-10 * abs(E[2])
