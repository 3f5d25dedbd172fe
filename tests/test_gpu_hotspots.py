"""GPU hotspot pass (e3x_hotspots, the reference's MATLAB post-processing steps for
top_peaks.txt, applications/sec3_2/A_2_0/elas3dxtal_postprocessing.m:86-243) against
the oracle's restatement.  Peak positions and counts must be identical; values agree
to 2e-6 relative, the rounding of single-precision sums in two orders (the device sums
the 3-D kernel directly, the oracle in separable passes as imgaussfilt3 does)."""
import numpy as np
import pytest

import _oracle as oc
from _cases import hot_fixture, voronoi_case
from elas3dxtal import _lib as L
from elas3dxtal.solver import Elas3D, solve

pytestmark = pytest.mark.gpu

C11, C12, C44 = 246.5e9, 147.3e9, 124.7e9


def test_plateaus_and_grid_corner_match_oracle():
    shape, vm, at = hot_fixture()
    pix = np.ones(vm.size, np.int32)
    with Elas3D(shape, 2) as s:
        s.ppixel(pix, np.zeros(3))
        idx, val, n = s.hotspots(vm=vm, nsolid=1, sigma=0.05, threshold_mpa=250.0, npeaks=5)
        # a constant field: every voxel is a regional maximum (all in one plateau)
        idx_c, _, n_c = s.hotspots(vm=np.full(vm.size, 300e6), nsolid=1, sigma=0.05, npeaks=3)
    idx_o, val_o, n_o = oc.hotspots(shape, pix, vm, 1, sigma=0.05, thr_mpa=250.0, npeaks=5)
    assert n == n_o == 4
    assert list(idx) == list(idx_o[:4]) == [at(2, 2, 2), at(3, 2, 2), at(7, 6, 5), at(0, 0, 0)]
    assert np.array_equal(val.astype(np.float32), val_o[:4])
    assert n_c == vm.size and list(idx_c) == [0, 1, 2]


@pytest.mark.parametrize("sigma,nsolid", [(0.4, 3), (0.4, 4), (1.2, 3)])
def test_random_field_with_pores_matches_oracle(sigma, nsolid):
    shape = (23, 19, 17)
    rng = np.random.default_rng(70)
    vm = rng.uniform(50e6, 400e6, int(np.prod(shape)))
    pix = rng.integers(1, 4, vm.size).astype(np.int32)
    pix[rng.random(vm.size) < 0.08] = 4
    thr = 240.0 if sigma < 1 else 226.0
    with Elas3D(shape, 4) as s:
        s.ppixel(pix, np.zeros(9))
        idx, val, n = s.hotspots(vm=vm, nsolid=nsolid, sigma=sigma, threshold_mpa=thr, npeaks=64)
    idx_o, val_o, n_o = oc.hotspots(shape, pix, vm, nsolid, sigma=sigma, thr_mpa=thr, npeaks=64)
    assert n == n_o and n > 10
    k = min(n, 64)
    assert list(idx) == list(idx_o[:k])
    assert np.max(np.abs(val - val_o[:k]) / val_o[:k]) < 2e-6


def test_solved_field_matches_oracle_and_writes_top_peaks(tmp_path):
    """The handle's own von Mises field (vm = NULL: computed on the device) of a solved
    polycrystal with pores, with the scripts' nsolid = max(pix) and with pores excluded."""
    from elas3dxtal.io import format_top_peaks
    shape = (20, 18, 16)
    pix, rod = voronoi_case(shape, 8, seed=71, pore_frac=0.03)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    res, s = solve(pix, rod, shape, 9, C11, C12, C44, e6, keep=True)
    try:
        assert res.exit[0] == L.PCG_CONVERGED
        _, vm = s.stress_fullfield(0)
        thr = float(np.percentile(vm / 1e6, 97))
        for nsolid in (9, 8):
            idx, val, n = s.hotspots(nsolid=nsolid, threshold_mpa=thr, npeaks=10)
            idx_o, val_o, n_o = oc.hotspots(shape, pix, vm, nsolid, sigma=0.4, thr_mpa=thr, npeaks=10)
            assert n == n_o and n >= 3
            assert list(idx) == list(idx_o[:min(n, 10)])
            assert np.max(np.abs(val - val_o[:len(val)]) / val_o[:len(val)]) < 2e-6
        txt = format_top_peaks(shape, idx, val)
        lines = txt.splitlines()
        assert lines[0] == "Top 3 Local Maxima (Von Mises Stress):" and len(lines) == 4
        m = int(idx[0])
        assert lines[1].endswith(f"at ({m % 20:.2f}, {m // 20 % 18:.2f}, {m // 360:.2f})")
    finally:
        s.close()


def test_hotspot_argument_errors():
    shape = (4, 4, 4)
    with Elas3D(shape, 2) as s:
        with pytest.raises(L.E3xError) as e:
            s.hotspots(vm=np.zeros(64))
        assert e.value.code == L.E3X_ESTATE     # no microstructure yet
        s.ppixel(np.ones(64, np.int32), np.zeros(3))
        for kw in (dict(sigma=0.0), dict(sigma=1.6), dict(npeaks=65)):
            with pytest.raises(L.E3xError) as e:
                s.hotspots(vm=np.zeros(64), nsolid=1, **kw)
            assert e.value.code == L.E3X_EINVAL
        with pytest.raises(L.E3xError) as e:
            s.hotspots(nsolid=1)                 # own field before femat
        assert e.value.code == L.E3X_ESTATE
