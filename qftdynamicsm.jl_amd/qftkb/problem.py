"""Host-side mirror of the reference's TwoPIScalar problem/solution API,
driving the MI355X back-end through libqftkb.

The reference builds a QFTdynamicsProblem from a parameter Dict
(TwoPIScalar_Modelfile.jl:46-124), allocates a solution and scratch
(:133-143), then runs initialize!, evolve! and measure!
(TwoPIScalar_Initialization.jl:289-325, TwoPIScalar_Evolution.jl:924-981,
TwoPIScalar_Measurement.jl:223-228).  The same names and parameter keys are
kept here; the numerics strings for this back-end are "GPUfull" (every
site stored, the drop-in for CPUfull) and "GPUorbits" (one value per orbit
of the lattice's cubic group, exact for the cubic-symmetric states that every
initialisation here produces; the counterpart of the reference's reduced
CPUred storage).

Left out on purpose: the run-directory side effects of
createfilestructandSimSetup (reading etc/paths.txt, purging and creating
directories, DefAbstractSimSetup.jl:64-108); the run-directory name itself
is reproduced by getParameterstring.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import lattice
from ._abi import PEXP_CODES, KBContext


class TwoPIScalarfile:
    """Model-file marker type (TwoPIScalar_Modelfile.jl:4)."""


# ---------------------------------------------------------------------------
# parameters / set-up (DefAbstractSimSetup.jl)
# ---------------------------------------------------------------------------
def splitter(idxmin: int, n: int, k: int):
    """Chunk ranges of the reference's thread split (DefAbstractSimSetup.jl:7-10)."""
    xz = [int(math.ceil(v)) for v in np.linspace(0, n, k + 1)]
    return [range(xz[i] + idxmin, xz[i + 1] + idxmin) for i in range(k)]


def check_parameters(parameters: dict) -> None:
    """DefAbstractSimSetup.jl:21-31."""
    if parameters["NstepsinMemory"] == 0:
        parameters["NstepsinMemory"] = parameters["Nsteps"] + 2
    if parameters["Nmeas"] == 0:
        parameters["Nmeas"] = parameters["Nsteps"]


@dataclass
class SimSetup:
    """DefAbstractSimSetup.jl:33-48."""
    parameterstring: str = ""
    datapath: str = ""
    postpath: str = ""
    plotpath: str = ""
    Nsteps: int = 0
    Nmeas: int = 0
    NstepsbetweenNmeas: int = 0
    lastEvolStep: int = 0
    NstepsinMemory: int = 0


def isnewsimulation(simsetup: SimSetup) -> bool:
    return simsetup.lastEvolStep == 0


def getParameterstring(modelfile, parameters) -> str:
    """Run-directory name, TwoPIScalar_Modelfile.jl:15-44."""
    s = parameters["Mod"] + parameters["Pexp"]
    s += "_L" + str(int(math.floor(parameters["Lambda"] * 100)))
    s += "_ON" + str(parameters["ONgroup"])
    s += "_" + parameters["Reno"]
    if parameters["init"] == "Thermal":
        s += "_T" + str(_julia_int(parameters["T"] * 10))
    elif parameters["init"] == "Pnr":
        s += "_n" + str(_julia_int(parameters["n"] * 10))
    else:
        s += "_" + parameters["init"]
    s += "_"
    s += "Nx" + str(parameters["Nx"])
    s += "_sdim" + str(parameters["sdim"])
    s += "_Nsteps" + str(parameters["Nsteps"])
    s += "_dt" + str(_julia_int(parameters["dt"] * 1000))
    s += "_M" + str(_julia_int(parameters["Mass"] * 100))
    s += "_NstepsinMemory" + str(parameters["NstepsinMemory"])
    s += "_Num" + parameters["Num"]
    if parameters.get("tag", "") != "":
        s += "_" + parameters["tag"]
    return s


def _julia_int(x):
    """Julia's Int64(x): exact conversion or InexactError."""
    if float(x) != int(round(float(x))):
        raise ValueError(f"InexactError: Int64({x})")
    return int(round(float(x)))


def createSimSetup(modelfile, parameters) -> SimSetup:
    """The bookkeeping half of createfilestructandSimSetup
    (DefAbstractSimSetup.jl:64-108)."""
    ss = SimSetup()
    ss.parameterstring = getParameterstring(modelfile, parameters)
    ss.NstepsinMemory = parameters["NstepsinMemory"]
    ss.lastEvolStep = 0
    ss.NstepsbetweenNmeas = int(math.floor(parameters["Nsteps"] / parameters["Nmeas"]))
    ss.Nsteps = ss.NstepsbetweenNmeas * parameters["Nmeas"]
    ss.Nmeas = parameters["Nmeas"]
    return ss


# ---------------------------------------------------------------------------
# problem types (TwoPIScalar_Problemdefinition.jl, _Discretization.jl,
# _ModelExpansionRenormalization.jl)
# ---------------------------------------------------------------------------
@dataclass
class Discretization:
    Mass: float
    Nx: int
    sdim: int
    Nsteps: int
    dt: float
    vol: int
    ivol: float
    NstepsinMemory: int
    Nmeas: int
    kL2: np.ndarray = field(repr=False, default=None)


@dataclass
class Phi4:
    Lambda: float
    ONgroup: int
    Mass2: float
    tachyonic: bool = False


@dataclass
class Init:
    kind: str
    params: dict


@dataclass
class TwoPIScalarProblem:
    """Fields of QFTdynamicsProblem (AbstractTypesSimulation.jl:19-27)."""
    model: Phi4
    pexp: str
    disc: Discretization
    init: Init
    reno: str
    num: str
    simsetup: SimSetup


# the device back-ends: every site stored, or one value per cubic orbit
NUMERICS = ("GPUfull", "GPUorbits")

_INIT_KEYS = {"Thermal": ("T",), "Pnr": ("n",), "Gauss": ("xi", "eta", "sig"), "Box": ("n0", "Q"),
              "TopHatT1": (), "TopHatT2": (), "TopHatT3": ()}


def _make_problem(modelfile, parameters: dict) -> TwoPIScalarProblem:
    """QFTdynamicsProblem(::TwoPIScalarfile, parameters), Modelfile.jl:46-124."""
    check_parameters(parameters)
    if parameters["Num"] not in NUMERICS:
        raise ValueError(f"Num={parameters['Num']!r}: this package implements the numerics {NUMERICS}")
    simsetup = createSimSetup(modelfile, parameters)
    Nx, sdim = parameters["Nx"], parameters["sdim"]
    V = Nx ** sdim
    disc = Discretization(parameters["Mass"], Nx, sdim, parameters["Nsteps"], parameters["dt"], V, 1 / V,
                          parameters["NstepsinMemory"], parameters["Nmeas"], lattice.kL2_table(Nx, sdim))
    if parameters["Mod"] == "phi4":
        model = Phi4(parameters["Lambda"], parameters["ONgroup"], parameters["Mass"] ** 2)
    elif parameters["Mod"] == "phi4tachyonic":
        model = Phi4(parameters["Lambda"], parameters["ONgroup"], parameters["Mass"] ** 2, tachyonic=True)
    else:
        raise ValueError(f"unknown Mod {parameters['Mod']!r}")
    kind = parameters["init"]
    if kind in ("Quench", "Tsunami"):
        # the reference has no getparticlenr method for these (Initialization.jl:133-203)
        raise NotImplementedError(f"init={kind!r} has no particle-number definition in the reference")
    if kind not in _INIT_KEYS:
        raise ValueError(f"unknown init {kind!r}")
    init = Init(kind, {k: parameters[k] for k in _INIT_KEYS[kind]})
    pexp = parameters["Pexp"]
    if pexp not in PEXP_CODES:
        raise ValueError(f"unknown Pexp {pexp!r}")
    reno = parameters["Reno"]
    if reno not in ("RNone", "RMass"):
        raise ValueError(f"unknown Reno {reno!r}")
    model = Renormalize(model, pexp, reno, disc)
    return TwoPIScalarProblem(model, pexp, disc, init, reno, parameters["Num"], simsetup)


def Renormalize(model: Phi4, pexp, reno, disc) -> Phi4:
    """ModelExpansionRenormalization.jl:18-36."""
    if reno == "RNone":
        return model
    H = lattice.hartreeprefactor(model.Lambda, model.ONgroup, pexp)
    ct = lattice.counterterm(disc.kL2, disc.Mass, H)
    if model.tachyonic:
        return Phi4(model.Lambda, model.ONgroup, -model.Mass2 + ct, tachyonic=True)
    return Phi4(model.Lambda, model.ONgroup, model.Mass2 + ct)


def QFTdynamicsProblem(modelfile, parameters: dict) -> TwoPIScalarProblem:
    """The reference's constructor spelling (Modelfile.jl:46)."""
    return _make_problem(modelfile, parameters)


# ---------------------------------------------------------------------------
# solution
# ---------------------------------------------------------------------------
@dataclass
class MeasurementTwoPIScalar:
    """TwoPIScalar_Measurement.jl:2-23.  Vector fields are per site (every
    site its own level); see lattice.levels for a per-shell view."""
    time: float = 0.0
    E: float = 0.0
    E_pot: float = 0.0
    E_kin: float = 0.0
    E_mem: float = 0.0
    M2: float = 0.0
    phi: float = 0.0
    F: np.ndarray = None
    ddF: np.ndarray = None
    r: np.ndarray = None
    n: np.ndarray = None
    omega: np.ndarray = None
    c: np.ndarray = None
    Fonet: np.ndarray = None
    ronet: np.ndarray = None
    SigF: np.ndarray = None
    Sigr: np.ndarray = None
    n_unstable: int = 0


class TwoPIScalarSimDataGPU:
    """Device-resident SimData: wraps a kb_ctx.  Mirrors the fields of
    TwoPIScalarSimDataCPUfull (SimDataStruct.jl:63-89) that callers read."""

    def __init__(self, problem: TwoPIScalarProblem, device: int = 0):
        d, m = problem.disc, problem.model
        pexp = PEXP_CODES[problem.pexp]
        self.ctx = KBContext(d.Nx, d.sdim, d.NstepsinMemory, d.dt, m.Lambda, m.ONgroup, d.Mass, m.Mass2,
                             pexp=pexp, device=device, symmetric=problem.num == "GPUorbits")
        self.ctx.set_kL2(d.kL2)
        self.NstepsinMemory = d.NstepsinMemory
        self.Nx, self.sdim = d.Nx, d.sdim

    @property
    def indices(self):
        return list(self.ctx.window())

    @property
    def hartreemass2(self):
        return [self.ctx.hartreemass2()]

    @property
    def kL2values(self):
        return self.ctx.kL2()


@dataclass
class QFTdynamicsSolutionTwoPIScalar:
    problem: TwoPIScalarProblem
    simdata: TwoPIScalarSimDataGPU
    measurearray: list


def QFTdynamicsSolution(modelfile, problem: TwoPIScalarProblem, device: int = 0):
    """Modelfile.jl:133-143.  Returns (solution, tmpdata); the device back-end
    keeps its scratch inside the context, so tmpdata is None."""
    simdata = TwoPIScalarSimDataGPU(problem, device)
    measurearray = [None] * (problem.simsetup.Nmeas + 1)
    return QFTdynamicsSolutionTwoPIScalar(problem, simdata, measurearray), None


def getparticlenr(init: Init, disc: Discretization) -> np.ndarray:
    k = init.kind
    if k == "Pnr":
        return lattice.n_particle(disc.kL2, init.params["n"])
    if k == "Thermal":
        return lattice.n_thermal(disc.kL2, disc.Mass, init.params["T"])
    if k.startswith("TopHatT"):
        return lattice.n_tophat(disc.kL2, int(k[-1]), disc.sdim, disc.Mass)
    if k == "Box":
        return lattice.n_box(disc.kL2, init.params["n0"], init.params["Q"])
    raise ValueError(f"no particle number for init {k!r}")


def getGapMass(problem: TwoPIScalarProblem) -> float:
    n = getparticlenr(problem.init, problem.disc)
    H = lattice.hartreeprefactor(problem.model.Lambda, problem.model.ONgroup, problem.pexp)
    return lattice.gap_mass(problem.disc.kL2, n, problem.model.Mass2, H)


def initialize(solution: QFTdynamicsSolutionTwoPIScalar, tmpdata=None) -> None:
    """initialize! (Initialization.jl:289-325)."""
    problem = solution.problem
    disc, model, init = problem.disc, problem.model, problem.init
    ctx = solution.simdata.ctx
    if isnewsimulation(problem.simsetup):
        if init.kind == "Gauss":
            p = init.params
            ctx.init_rows(*lattice.gauss_rows(disc.vol, disc.dt, p["xi"], p["eta"], p["sig"]))
        else:
            if model.tachyonic or init.kind in ("TopHatT1", "Thermal", "Box"):
                m0 = disc.Mass
            else:  # Pnr, TopHatT2, TopHatT3: LO fixed point
                m0 = getGapMass(problem)
            ctx.init_particles(getparticlenr(init, disc), m0)
    problem.simsetup.lastEvolStep = 2
    measure(solution, 2)


def evolve(solution: QFTdynamicsSolutionTwoPIScalar, tmpdata, t: int) -> None:
    """evolve! (Evolution.jl:924-981) on the device."""
    solution.simdata.ctx.step(t)
    solution.problem.simsetup.lastEvolStep = t


def createMeasurement(t: int, solution: QFTdynamicsSolutionTwoPIScalar) -> MeasurementTwoPIScalar:
    ctx = solution.simdata.ctx
    m = ctx.measure(t)
    h = ctx.measure_history(t)
    return MeasurementTwoPIScalar(time=m["time"], E=m["E"], E_pot=m["E_pot"], E_kin=m["E_kin"],
                                  E_mem=m["E_mem"], M2=m["M2"], F=m["F"], ddF=m["ddF"], r=m["r"], n=m["n"],
                                  omega=m["omega"], c=m["c"], Fonet=h["Fonet"], ronet=h["ronet"],
                                  SigF=h["SigF"], Sigr=h["Sigr"], n_unstable=m["n_unstable"])


def measure(solution: QFTdynamicsSolutionTwoPIScalar, t: int) -> None:
    """measure! (Measurement.jl:223-228)."""
    k = (t - 2) / solution.problem.simsetup.NstepsbetweenNmeas
    if k != int(k):
        raise ValueError(f"InexactError: measurement at t={t} is not on the measurement grid")
    solution.measurearray[int(k)] = createMeasurement(t, solution)


def run(modelfile, parameters: dict, device: int = 0, progress=None):
    """The driver loop of G_Simulation.jl:12-31."""
    problem = _make_problem(modelfile, parameters)
    solution, tmpdata = QFTdynamicsSolution(modelfile, problem, device)
    initialize(solution, tmpdata)
    ss = problem.simsetup
    for t in range(ss.lastEvolStep + 1, ss.Nsteps + 3):
        evolve(solution, tmpdata, t)
        if (t - 2) % ss.NstepsbetweenNmeas == 0:
            measure(solution, t)
        if progress:
            progress(t)
    return solution
