"""qftkb: host-side mirror of the reference's TwoPIScalar (2PI Kadanoff-Baym)
problem / solution API, running the KB step on MI355X through libqftkb.so.

    from qftkb import TwoPIScalarfile, QFTdynamicsProblem, QFTdynamicsSolution
    from qftkb import initialize, evolve, measure
    problem = QFTdynamicsProblem(TwoPIScalarfile(), parameters)
    solution, tmpdata = QFTdynamicsSolution(TwoPIScalarfile(), problem)
    initialize(solution, tmpdata)
    for t in range(3, problem.simsetup.Nsteps + 3):
        evolve(solution, tmpdata, t)
        ...
"""
from ._abi import KBContext, KBEnsemble, KBError, load_library, LIB_PATH, EXPORTED, FFT_ORBIT_PRODUCT  # noqa: F401
from . import lattice, output  # noqa: F401
from .problem import (  # noqa: F401
    TwoPIScalarfile, QFTdynamicsProblem, QFTdynamicsSolution, TwoPIScalarProblem, SimSetup,
    MeasurementTwoPIScalar, check_parameters, getParameterstring, splitter, initialize, evolve,
    measure, createMeasurement, getparticlenr, getGapMass, Renormalize, run,
)
