"""MI355X-native field time-stepping for semba-fdtd (Yee + CPML + pole-residue
ADE), hand-written HIP for gfx950 behind a C ABI (include/sembafdtd.h).

The compute path is ``libsembafdtd.so`` in this directory; there is no CPU
fallback.
"""
from .problem import (Problem, NodalSource, EPS0, MU0, EX, EY, EZ, HX, HY, HZ,
                      KIND_PEC, KIND_PMC, KIND_EDISP, KIND_MDISP, cavity_problem, scatterer_problem)
from ._lib import FdtdError, LIB_PATH, load as load_library
from .solver import Solver, build_tables

__all__ = ["Problem", "NodalSource", "Solver", "build_tables", "FdtdError", "LIB_PATH",
           "load_library", "cavity_problem", "scatterer_problem", "EPS0", "MU0",
           "EX", "EY", "EZ", "HX", "HY", "HZ", "KIND_PEC", "KIND_PMC", "KIND_EDISP", "KIND_MDISP"]
