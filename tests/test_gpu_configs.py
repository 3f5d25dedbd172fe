"""The larger workloads at their full sizes on the GPU (too large for the CPU oracle:
the FP64 GPU path they compare against is pinned to the oracle on smaller grids).

- 300^3, 1500 grains, A = 2, 3 % porosity: the mixed-precision form (FP32 stencil,
  vectors and block-Jacobi; FP64 u, dots and norms; true-residual replacement every 50
  iterations) against the FP64 solve; the error of every output is printed.
- 400^3, 4000 grains, A = 3, 2 % spheres + lack-of-fusion ellipsoids (20x5x3 voxels):
  the six unit load cases as six right-hand sides, FP64 and mixed; the mixed effective
  stiffness within 1e-6 of the FP64 one.
"""
import numpy as np
import pytest

from elas3dxtal import _lib as L
from elas3dxtal import configs
from elas3dxtal.solver import effective_stiffness, solve

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_cube300_mixed_against_fp64():
    w = configs.build("cube300")
    c11, c12, c44 = configs.materials("cube300")
    e6 = configs.exx_strain()
    args = (w["pix"], w["rod"], w["shape"], w["nphase"], c11, c12, c44, e6)
    r64 = solve(*args, precision=L.PREC_FP64)
    r32 = solve(*args, precision=L.PREC_MIXED, replace_every=50)
    assert r64.exit[0] == r32.exit[0] == L.PCG_CONVERGED
    errs = dict(utot=abs(r32.utot[0] - r64.utot[0]) / abs(r64.utot[0]),
                stress6=_rel(r32.stress6[0], r64.stress6[0]),
                strain6=_rel(r32.strain6[0], r64.strain6[0]),
                lstep=(int(r64.lstep[0]), int(r32.lstep[0])))
    print("300^3 mixed vs FP64:", errs)
    assert abs(int(r32.lstep[0]) - int(r64.lstep[0])) <= max(2, int(0.05 * r64.lstep[0]))
    assert errs["utot"] < 1e-6
    assert errs["stress6"] < 1e-5
    assert errs["strain6"] < 1e-12


def test_cube400_effective_stiffness_mixed_against_fp64():
    w = configs.build("cube400")
    c11, c12, c44 = configs.materials("cube400")
    args = (w["pix"], w["rod"], w["shape"], w["nphase"], c11, c12, c44)
    ce64, l64, x64 = effective_stiffness(*args, precision=L.PREC_FP64)
    ce32, l32, x32 = effective_stiffness(*args, precision=L.PREC_MIXED, replace_every=50)
    assert np.all(x64 == L.PCG_CONVERGED) and np.all(x32 == L.PCG_CONVERGED)
    err = _rel(ce32, ce64)
    print("400^3 six load cases: lstep FP64", l64.tolist(), "mixed", l32.tolist(),
          "C_eff mixed vs FP64", err)
    print("C_eff FP64 (GPa):\n", np.round(ce64 / 1e9, 3))
    assert err < 1e-6
