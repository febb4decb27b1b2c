"""crysscft_sample_amd: the crysSCFT propagator hot path on MI355X (gfx950).

The compute lives in ``libcrys_solver.so`` (HIP, C ABI in include/crys_solver.h);
this package is its Python mirror of the reference's solver plugin API and
SCFT driver.
"""
from ._lib import CrysError, LibraryNotBuilt, lib  # noqa: F401
from .solver import (  # noqa: F401
    CrysPSSolver3m,
    CrysPSSolverBase,
    CrysPSSolverP1,
    CrysPSSolverPmmm,
    SOLVERS,
    Solver3m,
    SolverP1,
    SolverPmmmDCT,
    axis_operator,
)
from .scft import SCFT, SCFTResult, diblock_scft_simple_iterate  # noqa: F401

__version__ = "0.1.0"
