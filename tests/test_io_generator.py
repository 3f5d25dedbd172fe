"""Generator, orientation conversion and HDF5 schema (CPU only)."""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

import _oracle as oc
from elas3dxtal import _lib as L
from elas3dxtal import configs, io


def test_generator_is_deterministic_and_seed_sensitive():
    a = io.generate((20, 18, 16), 12, 7, pore_mode=io.PORES_SPHERES, pore_fraction=0.02)
    b = io.generate((20, 18, 16), 12, 7, pore_mode=io.PORES_SPHERES, pore_fraction=0.02)
    c = io.generate((20, 18, 16), 12, 8, pore_mode=io.PORES_SPHERES, pore_fraction=0.02)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


@pytest.mark.parametrize("shape,ngrain", [((16, 16, 16), 10), ((24, 17, 9), 40), ((8, 8, 8), 1)])
def test_voronoi_equals_periodic_kdtree(shape, ngrain):
    """Without pores every voxel carries the grain of its nearest seed under the
    minimum-image metric; rebuild the seeds from the same RNG stream and compare with
    scipy's periodic k-d tree."""
    pix, rod, eul, frac = io.generate(shape, ngrain, 11)
    assert frac == 0.0 and pix.min() >= 1 and pix.max() <= ngrain
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) + 0.5
    seeds = _mt_seeds(11, ngrain, shape)
    tree = cKDTree(seeds, boxsize=np.array(shape, float))
    d, lab = tree.query(pts, k=min(2, ngrain))
    if ngrain > 1:
        first = lab[:, 0]
        tie = np.isclose(d[:, 0], d[:, 1], rtol=0, atol=1e-12)
        ok = (pix - 1 == first) | tie
        assert np.all(ok)
    else:
        assert np.all(pix == 1)


def _mt_seeds(seed, ngrain, shape):
    """The generator's seed draws: std::mt19937_64(seed), u = (x >> 11) * 2^-53."""
    mt = _MT64(seed)
    out = np.empty((ngrain, 3))
    for g in range(ngrain):
        for c in range(3):
            out[g, c] = (mt.next() >> 11) * 2.0 ** -53 * shape[c]
    return out


class _MT64:
    """Reference MT19937-64 (Matsumoto & Nishimura 2004), for reproducing draws."""
    def __init__(self, seed):
        self.mt = [0] * 312
        self.idx = 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & 0xFFFFFFFFFFFFFFFF

    def next(self):
        if self.idx >= 312:
            um, lm = 0xFFFFFFFF80000000, 0x7FFFFFFF
            for i in range(312):
                x = (self.mt[i] & um) | (self.mt[(i + 1) % 312] & lm)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[i] = self.mt[(i + 156) % 312] ^ xa
            self.idx = 0
        x = self.mt[self.idx]
        self.idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & 0xFFFFFFFFFFFFFFFF


def test_mt64_reference_value():
    # the 10000th output of std::mt19937_64 with the default seed 5489 (C++11 [rand.predef])
    mt = _MT64(5489)
    for _ in range(9999):
        mt.next()
    assert mt.next() == 9981545732273789042


def test_pore_fractions_and_phase_ids():
    pix, _, _, frac = io.generate((40, 40, 40), 30, 3, pore_mode=io.PORES_VOXELS, pore_fraction=0.02)
    assert frac == np.mean(pix == 31) and abs(frac - 0.02) < 1e-12
    pix, _, _, frac = io.generate((40, 40, 40), 30, 3, pore_mode=io.PORES_SPHERES, pore_fraction=0.03,
                                  pore_median=3.0, pore_sigma=0.4)
    assert frac == np.mean(pix == 31) and 0.03 <= frac < 0.05
    pix, _, _, frac = io.generate((60, 60, 30), 30, 3, pore_mode=io.PORES_SPHERES, pore_fraction=0.02,
                                  lof_fraction=0.01, lof_axes=(20, 5, 3))
    assert frac >= 0.03 and frac == np.mean(pix == 31)


def test_orientations_uniform_and_rodrigues_convention():
    _, rod, eul, _ = io.generate((8, 8, 8), 4000, 5)
    # Q(r) of the reference equals the passive Bunge matrix of the returned Euler angles
    for g in range(0, 4000, 97):
        q = oc.rodrigues_matrix(rod[g])
        gmat = Rotation.from_euler("ZXZ", eul[g]).as_matrix().T      # passive g = R^T
        assert np.allclose(q, gmat, atol=1e-10)
    # uniform on SO(3): the rotation angle has density (1 - cos w)/pi on [0, pi]
    w = 2 * np.arctan(np.linalg.norm(rod, axis=1))
    hist, edges = np.histogram(w, bins=6, range=(0, np.pi))
    cdf = lambda x: (x - np.sin(x)) / np.pi
    want = 4000 * np.diff(cdf(edges))
    assert np.all(np.abs(hist - want) < 5 * np.sqrt(want) + 5)


def test_bunge_round_trip():
    rng = np.random.default_rng(0)
    e = np.stack([rng.random(200) * 2 * np.pi, np.arccos(1 - 2 * rng.random(200)), rng.random(200) * 2 * np.pi], 1)
    r = io.rodrigues_from_bunge(e)
    for g in range(200):
        assert np.allclose(oc.rodrigues_matrix(r[g]), Rotation.from_euler("ZXZ", e[g]).as_matrix().T, atol=1e-10)
    e2 = io.bunge_from_rodrigues(r)
    r2 = io.rodrigues_from_bunge(e2)
    for g in range(200):
        assert np.allclose(oc.rodrigues_matrix(r2[g]), oc.rodrigues_matrix(r[g]), atol=1e-10)


def test_hdf5_input_round_trip_and_errors(tmp_path):
    shape = (10, 9, 8)
    pix, rod, _, _ = io.generate(shape, 6, 2, pore_mode=io.PORES_VOXELS, pore_fraction=0.05)
    f = str(tmp_path / "input_structure_poly.h5")
    io.write_input(f, shape, pix, rod)
    npix, nori, dims = io.probe_input(f)
    assert (npix, nori, dims) == (720, 18, shape)
    p2, r2 = io.read_input(f, 720, 7)
    assert np.array_equal(p2, pix) and np.array_equal(r2, rod.ravel())
    with pytest.raises(Exception) as ei:
        io.read_input(f, 721, 7)                      # /pix size != ns
    assert ei.value.code == 1
    with pytest.raises(Exception) as ei:
        io.read_input(f, 720, 8)                      # /orientation size != 3*(nphase-1)
    assert ei.value.code == 1
    bad = pix.copy()
    bad[0] = 9
    io.write_input(f, shape, bad, rod)
    with pytest.raises(Exception) as ei:
        io.read_input(f, 720, 7)                      # label out of [1, nphase]
    assert ei.value.code == 4
    with pytest.raises(Exception) as ei:
        io.read_input(str(tmp_path / "missing.h5"), 720, 7)
    assert ei.value.code == 6


def test_hdf5_fullfield_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    ns = 1000
    pix = rng.integers(1, 5, ns).astype(np.int32)
    vm, st = rng.random(ns), rng.standard_normal((6, ns))
    f = str(tmp_path / "fullfield_poly.h5")
    io.write_fullfield(f, pix, vm, st)
    p2, v2, s2 = io.read_fullfield(f, ns)
    assert np.array_equal(p2, pix) and np.array_equal(v2, vm) and np.array_equal(s2, st)


def test_config_materials():
    c11, c12, c44 = configs.materials("cube200")
    assert abs(2 * c44 / (c11 - c12) - 2.0) < 1e-12
    assert configs.materials("cube50") == (246.5e9, 147.3e9, 124.7e9)
    for a, want in [(0.5, 646.1), (1.0, 396.7), (5.0, 197.18)]:
        assert abs(configs.c11_for_zener(a) - want) < 1e-9
    e = configs.six_unit_strains()
    assert e[0, 0] == np.float32(1e-3) and e[5, 5] == 0.5 * np.float32(1e-3)


def test_request_from_files_reads_schema_and_controls(tmp_path):
    """A serving request from the reference's input file and JSON controls: the grid from
    the /pix dimensions attribute, nphase from /orientation, the reference's defaults for
    what the controls leave out, and errors for unknown or conflicting keys."""
    from _cases import voronoi_case
    pix, rod = voronoi_case((8, 7, 6), 4, seed=1)
    p = str(tmp_path / "in.h5")
    io.write_input(p, (8, 7, 6), pix, rod)
    rq, pp, oo, ff = io.request_from_files(p, "")
    try:
        assert (rq.nx, rq.ny, rq.nz, rq.nphase) == (8, 7, 6, 5)
        assert (rq.c11, rq.c12, rq.c44) == (246.5e9, 147.3e9, 124.7e9)
        assert list(rq.strain6) == [float(np.float32(1e-3)), 0, 0, 0, 0, 0]
        assert rq.tol == float(np.float32(1e-8)) and rq.ldemb == 10000 and rq.precision == 0
        assert rq.flag_m == 1 and ff is None
        assert np.array_equal(np.ctypeslib.as_array(pp, shape=(pix.size,)), pix)
        assert np.array_equal(np.ctypeslib.as_array(oo, shape=(rod.size,)), rod.ravel())
    finally:
        io.request_free(pp, oo)
    ctl = {"zener": 2.0, "precision": "mixed", "replace_every": 25, "strain6": [0, 0, 1e-3, 0, 0, 0],
           "fullfield": str(tmp_path / "ff.h5")}
    rq, pp, oo, ff = io.request_from_files(p, ctl)
    io.request_free(pp, oo)
    assert rq.c11 == 147.3e9 + 2 * 124.7e9 / 2.0 and rq.precision == 1 and rq.replace_every == 25
    assert list(rq.strain6) == [0, 0, 1e-3, 0, 0, 0] and ff == ctl["fullfield"]
    for bad in ('{"bogus": 1}', '{"c11": 1e9, "zener": 2}', '{"strain6": [1, 2]}', '{"precision": "fp16"}',
                '{"c11": 1e9', '[1, 2]'):
        with pytest.raises(L.E3xError) as e:
            io.request_from_files(p, bad)
        assert e.value.code == L.E3X_EINVAL, bad
    with pytest.raises(L.E3xError) as e:        # the grid is neither in the file nor given
        q = str(tmp_path / "noattr.h5")
        io.write_input(q, (8, 7, 6), pix, rod)
        rq, pp, oo, _ = io.request_from_files(q, '{"grid": [8, 7, 5]}')   # wrong size: f90:2939-2942
    assert e.value.code == L.E3X_EINVAL
