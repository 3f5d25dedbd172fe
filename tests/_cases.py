"""Seeded test microstructures shared by the oracle pins and the GPU parity tests."""
import numpy as np

import _indep as ind


def voronoi_case(shape, ngrain, seed, pore_frac=0.02):
    """Periodic Voronoi grains (nearest seed, minimum image) + random single-voxel pores
    (phase ngrain+1).  Returns pix (1-based labels, node order m = i + nx(j + ny k)) and
    Rodrigues vectors of the grains."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    dims = np.array([nx, ny, nz], float)
    seeds = rng.random((ngrain, 3)) * dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) + 0.5
    best = np.full(pts.shape[0], np.inf)
    lab = np.zeros(pts.shape[0], np.int32)
    for g in range(ngrain):
        d = np.abs(pts - seeds[g])
        d = np.minimum(d, dims - d)
        d2 = np.sum(d * d, axis=1)
        upd = d2 < best
        best[upd] = d2[upd]
        lab[upd] = g + 1
    npore = int(round(pore_frac * pts.shape[0]))
    if npore:
        lab[rng.choice(pts.shape[0], npore, replace=False)] = ngrain + 1
    rod = ind.random_rodrigues(rng, ngrain)
    return lab, rod


# A case whose PCG resets its direction (f90:1961-1981): cubic grains with negative
# pore moduli (K, G) = (-160, 10) GPa make the stiffness indefinite, and the
# preconditioned residual jumps by more than 10x several times before it converges.
RESET_CASE = dict(shape=(6, 6, 6), ngrain=2, seed=2, pore_frac=0.1, phasemod=(-160e9, 10e9))


def unit_strains(magnitude):
    """Six unit load cases {exx, eyy, ezz, exz, eyz, exy} (tensor components; the shears
    are engineering strains of `magnitude`, tensor component half of it)."""
    e = np.zeros((6, 6))
    for k in range(6):
        e[k, k] = magnitude if k < 3 else 0.5 * magnitude
    return e


def ceff_from_cases(stress6, strain6):
    """C_eff with <sigma>_k = C_eff <eps>_k for the six cases (rows k of the inputs)."""
    S = np.asarray(stress6, float).T
    E = np.asarray(strain6, float).T
    return S @ np.linalg.inv(E)


def hot_fixture():
    """A 9x8x7 field at 100 MPa with: a two-voxel plateau of 300 MPa (both are regional
    maxima), a three-voxel plateau of 280 MPa touching a 290 MPa voxel (not maxima; the
    290 is), and a lone 260 MPa voxel at the grid's corner (a maximum: no wrap)."""
    shape = (9, 8, 7)
    nx, ny, nz = shape
    vm = np.full(int(np.prod(shape)), 100e6)
    at = lambda i, j, k: i + nx * (j + ny * k)
    vm[at(2, 2, 2)] = vm[at(3, 2, 2)] = 300e6
    vm[at(6, 5, 4)] = vm[at(6, 6, 4)] = vm[at(6, 6, 5)] = 280e6
    vm[at(7, 6, 5)] = 290e6
    vm[at(0, 0, 0)] = 260e6
    return shape, vm, at
