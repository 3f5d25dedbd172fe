"""MI355X (gfx950) elas3dxtal_pcg solver: host-side mirror of the reference routines.

The compute path is libe3x.so (HIP kernels behind the C ABI of include/e3x.h);
this package only binds it.
"""
from . import _lib
from .solver import (AML_DEFAULT, LDEMB_DEFAULT, TOL_DEFAULT, Elas3D, SolveResult, effective_stiffness, solve,
                     solve_batch, unit_strain)

__all__ = ["Elas3D", "SolveResult", "solve", "solve_batch", "effective_stiffness", "unit_strain", "TOL_DEFAULT", "AML_DEFAULT",
           "LDEMB_DEFAULT", "_lib"]
