"""The benchmark workloads of the elas3dxtal_pcg solve, as generator parameters and
material constants.  Moduli in Pa; strains are tensor components {exx,eyy,ezz,exz,eyz,exy}."""
import numpy as np

from . import io
from .solver import AML_DEFAULT

C12_GPA, C44_GPA = 147.3, 124.7


def c11_for_zener(a, c12=C12_GPA, c44=C44_GPA):
    """Zener ratio A = 2 C44 / (C11 - C12)  ->  C11 (GPa)."""
    return c12 + 2.0 * c44 / a


CONFIGS = {
    # Plain FP64 PCG, 50^3, 20 grains, 2% random single-voxel pores, exx only.
    "cube50": dict(n=50, ngrain=20, c=(246.5, 147.3, 124.7), pores=dict(pore_mode=io.PORES_VOXELS, pore_fraction=0.02),
                 seed=1001),
    # Plain FP64 PCG, 200^3, 500 grains, Zener A in {0.5, 1, 2, 5} (A=2 is the headline),
    # 1% lognormal spheres (median 3 voxels, sigma 0.4), exx.
    "cube200": dict(n=200, ngrain=500, zener=2.0,
                 pores=dict(pore_mode=io.PORES_SPHERES, pore_fraction=0.01, pore_median=3.0, pore_sigma=0.4),
                 seed=2001),
    # Mixed precision, 300^3, 1500 grains, A=2, 3% spheres, reliable update every 50.
    "cube300": dict(n=300, ngrain=1500, zener=2.0,
                 pores=dict(pore_mode=io.PORES_SPHERES, pore_fraction=0.03, pore_median=3.0, pore_sigma=0.4),
                 seed=3001),
    # Six unit load cases, 400^3, 4000 grains, A=3, 2% spheres + lack-of-fusion ellipsoids
    # (semi-axes 20 x 5 x 3 voxels, random yaw about z).
    "cube400": dict(n=400, ngrain=4000, zener=3.0,
                 pores=dict(pore_mode=io.PORES_SPHERES, pore_fraction=0.02, pore_median=3.0, pore_sigma=0.4,
                            lof_fraction=0.005, lof_axes=(20.0, 5.0, 3.0)),
                 seed=4001),
}


def materials(cfg, zener=None):
    c = CONFIGS[cfg]
    if "c" in c and zener is None:
        c11, c12, c44 = c["c"]
    else:
        a = zener if zener is not None else c["zener"]
        c11, c12, c44 = c11_for_zener(a), C12_GPA, C44_GPA
    return c11 * 1e9, c12 * 1e9, c44 * 1e9


def build(cfg, n=None, ngrain=None, seed=None):
    """Generate a workload: returns dict(shape, nphase, pix, rod, euler, pore_fraction)."""
    c = CONFIGS[cfg]
    n = n or c["n"]
    ng = ngrain or c["ngrain"]
    pix, rod, eul, frac = io.generate((n, n, n), ng, seed if seed is not None else c["seed"], **c["pores"])
    return dict(shape=(n, n, n), nphase=ng + 1, pix=pix, rod=rod, euler=eul, pore_fraction=frac)


def exx_strain():
    return np.array([AML_DEFAULT, 0, 0, 0, 0, 0])


def six_unit_strains(magnitude=AML_DEFAULT):
    """Three normal and three engineering-shear unit load cases (tensor shear = half)."""
    e = np.zeros((6, 6))
    for k in range(3):
        e[k, k] = magnitude
    for k in range(3, 6):
        e[k, k] = 0.5 * magnitude
    return e


def serving_requests(n_requests, seed=5000):
    """The serving stream's requests, each with its own seeded microstructure: grid edge
    uniform over {64, 96, 128, 160, 200}, 10-800 Voronoi grains with uniform SO(3)
    orientations, Zener ratio A uniform on [0.5, 5], porosity uniform on [0, 5] % as
    lognormal spheres (median radius 3 voxels, sigma 0.4), one random unit macro strain
    of magnitude 1e-3 (tensor shear 0.5e-3).  `seed` draws the parameters and request k's
    microstructure is generated with seed + k.  Returns dicts
    (shape, nphase, pix, rod, c11, c12, c44 in Pa, e6)."""
    rng = np.random.default_rng(seed)
    edges = [64, 96, 128, 160, 200]
    reqs = []
    for k in range(n_requests):
        e = int(edges[int(rng.integers(0, len(edges)))])
        ng = int(rng.integers(10, 801))
        frac = float(rng.uniform(0.0, 0.05))
        pix, rod, _, _ = io.generate((e, e, e), ng, seed + k, pore_mode=io.PORES_SPHERES,
                                     pore_fraction=frac, pore_median=3.0, pore_sigma=0.4)
        a = float(rng.uniform(0.5, 5.0))
        e6 = np.zeros(6)
        j = int(rng.integers(0, 6))
        e6[j] = 1e-3 if j < 3 else 0.5e-3
        reqs.append(dict(shape=(e, e, e), nphase=ng + 1, pix=pix, rod=rod, zener=a,
                         c11=c11_for_zener(a) * 1e9, c12=C12_GPA * 1e9, c44=C44_GPA * 1e9, e6=e6))
    return reqs
