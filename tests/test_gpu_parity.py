"""GPU parity of the HIP library (through the C ABI) against the CPU oracle.

Tolerances: the FP64 path must match within 1e-10 relative on fields, energies and
averages (every table and single-pass quantity within 1e-12); the iteration count
of the reference's stop rule must match within 1 percent.
"""
import numpy as np
import pytest

import _indep as ind
import _oracle as oc
from _cases import RESET_CASE, ceff_from_cases, unit_strains, voronoi_case
from elas3dxtal import _lib as L
from elas3dxtal.solver import Elas3D, effective_stiffness, solve

pytestmark = pytest.mark.gpu

C11, C12, C44 = 246.5e9, 147.3e9, 124.7e9


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def oracle_tables(nphase, rod, pm=(0.0, 0.0)):
    rot = oc.rotate_stiffness(rod, C11, C12, C44)
    cm = oc.cmod(nphase, rot, list(pm))
    return cm, oc.dk(cm)


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert L.device_count() >= 1, "no HIP device visible"


# ---------------------------------------------------------------- femat tables, b, C
@pytest.mark.parametrize("shape,ngrain", [((4, 4, 4), 3), ((9, 7, 5), 6), ((16, 16, 16), 12)])
def test_femat_tables_rhs_constant(shape, ngrain):
    pix, rod = voronoi_case(shape, ngrain, seed=100 + ngrain)
    nphase = ngrain + 1
    e6 = [1.1e-3, -0.4e-3, 0.7e-3, 0.3e-3, -0.2e-3, 0.5e-3]
    with Elas3D(shape, nphase) as s:
        s.ppixel(pix, rod)
        cval = s.femat(C11, C12, C44, e6)
        cm, dk = s.tables()
        b = s.get(L.FIELD_B)
        prob = s.assig()
    cm_o, dk_o = oracle_tables(nphase, rod)
    # the same operations in the same order, each rounded on its own: the same bits
    assert np.array_equal(cm, cm_o)
    assert np.array_equal(dk, dk_o)
    b_o, c_o = oc.femat_rhs(shape, pix, dk_o, e6)
    assert _rel(b, b_o) < 1e-12
    assert abs(cval[0] - c_o) < 1e-12 * abs(c_o)
    assert np.array_equal(prob, oc.assig(pix, nphase))          # integer counts: exact


# ---------------------------------------------------------------- stencil + energy
@pytest.mark.parametrize("shape,ngrain", [((4, 4, 4), 2), ((11, 6, 9), 5), ((24, 20, 16), 8)])
def test_energy_and_stencil(shape, ngrain):
    pix, rod = voronoi_case(shape, ngrain, seed=200 + ngrain, pore_frac=0.05)
    nphase = ngrain + 1
    ns = int(np.prod(shape))
    e6 = [0.0, 0.0, oc.AML_F32, 0.2e-3, 0.0, -0.1e-3]
    rng = np.random.default_rng(7)
    u = oc.affine(shape, e6) + 1e-4 * rng.standard_normal(3 * ns)
    cm_o, dk_o = oracle_tables(nphase, rod)
    b_o, c_o = oc.femat_rhs(shape, pix, dk_o, e6)
    gb_o, ut_o = oc.energy(shape, pix, dk_o, b_o, c_o, u)
    with Elas3D(shape, nphase) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.set(L.FIELD_U, u)
        ut, gg = s.energy()
        gb = s.get(L.FIELD_GB)
        # bare stencil: b = 0 gives gb = K x
        x = rng.standard_normal(3 * ns)
        s.set(L.FIELD_B, np.zeros(3 * ns))
        s.set(L.FIELD_U, x)
        s.energy()
        kx = s.get(L.FIELD_GB)
    assert _rel(gb, gb_o) < 1e-11
    assert abs(ut[0] - ut_o) < 1e-12 * abs(ut_o)
    assert abs(gg[0] - oc.sumsq(gb_o)) < 1e-10 * oc.sumsq(gb_o)
    assert _rel(kx, oc.apply(shape, pix, dk_o, x)) < 1e-12


def test_affine_guess_and_homogeneous_equilibrium():
    shape = (10, 9, 8)
    ns = int(np.prod(shape))
    e6 = [0.9e-3, -0.3e-3, 0.2e-3, 0.4e-3, -0.1e-3, 0.25e-3]
    with Elas3D(shape, 2) as s:
        s.ppixel(np.ones(ns, np.int32), np.array([0.3, -0.2, 0.5]))
        s.femat(C11, C12, C44, e6)
        s.affine_guess()
        u = s.get(L.FIELD_U)
        ut, gg = s.energy()
        cm, _ = s.tables()
        gb = s.get(L.FIELD_GB)
        b = s.get(L.FIELD_B)
    assert np.array_equal(u, oc.affine(shape, e6))
    assert np.max(np.abs(gb)) < 1e-10 * np.max(np.abs(b))
    eps = np.array([e6[0], e6[1], e6[2], 2 * e6[3], 2 * e6[4], 2 * e6[5]])
    assert abs(ut[0] - 0.5 * ns * eps @ cm[0] @ eps) < 1e-12 * abs(ut[0])


# ---------------------------------------------------------------- PCG
def _oracle_pcg(shape, pix, rod, e6, tol=oc.TOL_F32, ldemb=10000, variant=0):
    nphase = int(pix.max()) if rod is None else len(rod) + 1
    cm_o, dk_o = oracle_tables(nphase, rod)
    b_o, c_o = oc.femat_rhs(shape, pix, dk_o, e6)
    u0 = oc.affine(shape, e6)
    g0, ut0 = oc.energy(shape, pix, dk_o, b_o, c_o, u0)
    gg0 = oc.sumsq(g0)
    r = oc.dembx(shape, pix, dk_o, u0, g0, gg0, tol=tol, ldemb=ldemb, variant=variant)
    return r, gg0


def _oracle_at(shape, pix, rod, e6, K, variant=0):
    """The oracle's PCG run for exactly K iterations (tol = 0: the iteration limit ends
    it, f90:1574): the iterates a run that stopped at K holds, for comparing at equal
    iteration counts when the stop test fell one iteration apart."""
    r, _ = _oracle_pcg(shape, pix, rod, e6, tol=0.0, ldemb=K, variant=variant)
    assert r["status"] == oc.OC_MAXITER and r["lstep"] == K
    return r


@pytest.mark.parametrize("shape,ngrain", [((8, 8, 8), 4), ((13, 10, 7), 6), ((20, 20, 20), 10)])
def test_dembx_matches_oracle(shape, ngrain):
    pix, rod = voronoi_case(shape, ngrain, seed=300 + ngrain)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    r_o, gg0_o = _oracle_pcg(shape, pix, rod, e6)
    with Elas3D(shape, ngrain + 1) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.affine_guess()
        ut0, gg0 = s.energy()
        assert abs(gg0[0] - gg0_o) < 1e-11 * gg0_o
        d = s.dembx(gg0)
        u = s.get(L.FIELD_U)
        h = s.get(L.FIELD_H)
    assert d["exit"][0] == L.PCG_CONVERGED == r_o["status"]
    K = int(d["lstep"][0])
    print(f"lstep GPU {K}, oracle {r_o['lstep']} (difference {K - r_o['lstep']})")
    assert abs(K - r_o["lstep"]) <= max(1, int(0.01 * r_o["lstep"]))
    # fields at equal iteration counts: the oracle's own stop, or its run of exactly K
    # iterations when the stop test fell elsewhere (the iterate and the direction after
    # iteration K are those of a stop at K, f90:1996-2007)
    ref = r_o if K == r_o["lstep"] else _oracle_at(shape, pix, rod, e6, K)
    assert _rel(u, ref["u"]) < 1e-10
    assert _rel(d["history"][0], ref["history"]) < 1e-8
    assert abs(d["gg"][0] - ref["gg"]) < 1e-8 * ref["gg"]
    # the direction after the stop test (f90:1961-1996, 2007): h is of the size of the
    # final residual, so its relative agreement is looser than u's
    assert _rel(h, ref["h"]) < 1e-5


@pytest.mark.parametrize("shape", [(2, 2, 2), (2, 3, 2), (3, 2, 5), (2, 9, 2), (5, 2, 3)])
@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_thin_grids_match_oracle(shape, precision):
    """Grids thinner than a stencil tile in every direction, down to two nodes: a node's
    -1 and +1 neighbours in a 2-wide direction are the same node (f90:350-372), and one
    tile covers the grid several times over through the periodic wrap."""
    pix, rod = voronoi_case(shape, 2, seed=7)
    e6 = [oc.AML_F32, 0, 0, 0, 0.3e-3, 0]
    r_o, gg0_o = _oracle_pcg(shape, pix, rod, e6)
    assert r_o["status"] == oc.OC_CONVERGED
    with Elas3D(shape, 3) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.affine_guess()
        u0 = s.get(L.FIELD_U)
        _, gg0 = s.energy()
        d = s.dembx(gg0, precision=precision)
        u = s.get(L.FIELD_U)
    assert abs(gg0[0] - gg0_o) < 1e-11 * gg0_o
    assert d["exit"][0] == L.PCG_CONVERGED
    if precision == L.PREC_FP64:
        assert d["lstep"][0] == r_o["lstep"]
        assert _rel(u, r_o["u"]) < 1e-10
    else:
        assert abs(int(d["lstep"][0]) - r_o["lstep"]) <= 2
        assert _rel(u - u0, r_o["u"] - u0) < 1e-5


def test_full_solve_matches_oracle_cube50():
    """Plain FP64 PCG on a 50^3 periodic cube: 20 Voronoi grains, 2% single-voxel pores,
    C11=246.5, C12=147.3, C44=124.7 GPa, exx = float32(1e-3)."""
    shape = (50, 50, 50)
    pix, rod = voronoi_case(shape, 20, seed=1000, pore_frac=0.02)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    ro = oc.solve(shape, 21, pix, rod, C11, C12, C44, e6)
    res, s = solve(pix, rod, shape, 21, C11, C12, C44, e6, keep=True)
    try:
        u = s.get(L.FIELD_U)
        sf, vm = s.stress_fullfield()
    finally:
        s.close()
    assert res.exit[0] == L.PCG_CONVERGED
    K = int(res.lstep[0])
    print(f"lstep GPU {K}, oracle {ro['lstep']}")
    assert abs(K - ro["lstep"]) <= max(1, int(0.01 * ro["lstep"]))
    if K != ro["lstep"]:
        # equal iteration counts: the oracle's main-program sequence with exactly K PCG
        # iterations (tol = 0, ldemb = K)
        ro = oc.solve(shape, 21, pix, rod, C11, C12, C44, e6, tol=0.0, ldemb=K)
    assert abs(res.utot[0] - ro["utot"]) < 1e-10 * abs(ro["utot"])
    assert abs(res.utot0[0] - ro["utot0"]) < 1e-10 * abs(ro["utot0"])
    assert _rel(res.stress6[0], ro["stress6"]) < 1e-10
    assert _rel(res.strain6[0], ro["strain6"]) < 1e-10
    assert _rel(u, ro["u"]) < 1e-10
    cm_o, _ = oracle_tables(21, rod)
    sf_o, vm_o = oc.fullfield(shape, pix, cm_o, e6, ro["u"])
    assert _rel(sf, sf_o) < 1e-9
    assert _rel(vm, vm_o) < 1e-9


def test_divergence_reset_matches_oracle():
    """An indefinite case (negative pore moduli) on which the PCG resets its direction
    several times (f90:1961-1981): same resets, iterations and history as the oracle."""
    rc = RESET_CASE
    shape = rc["shape"]
    pix, rod = voronoi_case(shape, rc["ngrain"], rc["seed"], rc["pore_frac"])
    nphase = rc["ngrain"] + 1
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    ro = oc.solve(shape, nphase, pix, rod, C11, C12, C44, e6, phasemod=rc["phasemod"])
    res, s = solve(pix, rod, shape, nphase, C11, C12, C44, e6, phasemod=rc["phasemod"], keep=True)
    try:
        u = s.get(L.FIELD_U)
    finally:
        s.close()
    assert ro["reset_count"] >= 2
    assert res.reset_count[0] == ro["reset_count"]
    assert res.exit[0] == L.PCG_CONVERGED
    assert res.lstep[0] == ro["lstep"]
    # indefinite operator: summation-order rounding grows to ~1e-6 (see the oracle pin)
    assert _rel(res.history[0], ro["history"]) < 1e-5
    assert abs(res.utot[0] - ro["utot"]) < 1e-8 * abs(ro["utot"])
    assert _rel(u, ro["u"]) < 1e-6


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_hah_break_on_all_pore_grid(precision):
    shape = (4, 4, 4)
    with Elas3D(shape, 2) as s:
        s.ppixel(np.full(64, 2, np.int32), np.zeros(3))
        s.femat(C11, C12, C44, [1e-3, 0, 0, 0, 0, 0])
        s.affine_guess()
        u0 = s.get(L.FIELD_U)
        ut, gg0 = s.energy()
        assert gg0[0] == 0.0
        d = s.dembx(gg0, precision=precision)
        assert d["exit"][0] == L.PCG_HAH_BREAK and d["lstep"][0] == 1
        assert np.array_equal(s.get(L.FIELD_U), u0)


@pytest.mark.parametrize("ldemb", [5, 6])
def test_maxiter_and_history_prefix(ldemb):
    """Stop by the iteration limit after an odd and an even number of iterations (the
    last direction is formed in either of the two direction buffers)."""
    shape = (12, 12, 12)
    pix, rod = voronoi_case(shape, 5, seed=400)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    r_o, _ = _oracle_pcg(shape, pix, rod, e6, ldemb=ldemb)
    assert r_o["status"] == oc.OC_MAXITER
    with Elas3D(shape, 6) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.affine_guess()
        _, gg0 = s.energy()
        d = s.dembx(gg0, ldemb=ldemb)
        u = s.get(L.FIELD_U)
        h = s.get(L.FIELD_H)
    assert d["exit"][0] == L.PCG_MAXITER and d["lstep"][0] == ldemb
    assert _rel(d["history"][0], r_o["history"]) < 1e-10
    assert _rel(u, r_o["u"]) < 1e-12
    assert _rel(h, r_o["h"]) < 1e-10


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_nist_absolute_stop_rule(precision):
    """E3X_STOP_NIST_ABS ends the PCG at the first iteration K whose unpreconditioned
    gradient has sum(g^2) < 1e-8 * ns (NIST ELAS3D's gtest).  The oracle, run with tol = 0
    for K and K - 1 iterations, must show the crossing there, and the iterates up to K are
    those of the reference's rule.  Moduli in GPa, the units that test is scaled for."""
    c11, c12, c44 = C11 * 1e-9, C12 * 1e-9, C44 * 1e-9
    shape = (12, 12, 12)
    ns = int(np.prod(shape))
    pix, rod = voronoi_case(shape, 5, seed=410)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    with Elas3D(shape, 6) as s:
        s.ppixel(pix, rod)
        s.femat(c11, c12, c44, e6)
        with pytest.raises(L.E3xError) as e:
            s.set_stop_rule(2)
        assert e.value.code == L.E3X_EINVAL
        s.set_stop_rule(L.STOP_NIST_ABS)
        s.affine_guess()
        _, gg0 = s.energy()
        d = s.dembx(gg0, precision=precision)
        u = s.get(L.FIELD_U)
        gb = s.get(L.FIELD_GB)
        # back to the default rule on the same handle: the reference's stop iteration
        s.set_stop_rule(L.STOP_REFERENCE)
        s.affine_guess()
        s.energy()
        d_ref = s.dembx(gg0, precision=precision)
    K = int(d["lstep"][0])
    assert d["exit"][0] == L.PCG_CONVERGED and 1 < K < 10000
    rot = oc.rotate_stiffness(rod, c11, c12, c44)
    dk_o = oc.dk(oc.cmod(6, rot, [0.0, 0.0]))
    b_o, c_o = oc.femat_rhs(shape, pix, dk_o, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk_o, b_o, c_o, u0)
    ks = range(K - 1, K + 1) if precision == L.PREC_FP64 else range(max(K - 2, 1), K + 3)
    runs = {k: oc.dembx(shape, pix, dk_o, u0, g0, oc.sumsq(g0), tol=0.0, ldemb=k) for k in ks}
    gsq = {k: oc.sumsq(r["gb"]) for k, r in runs.items()}
    gtest = 1e-8 * ns
    # the reference's rule on the same handle; in these units its absolute hAh floor
    # (|hAh| < 1e-14, f90:1899) may end the loop before rel_res < tol does
    r_ref = oc.dembx(shape, pix, dk_o, u0, g0, oc.sumsq(g0), tol=oc.TOL_F32)
    assert d_ref["exit"][0] == r_ref["status"]
    if precision == L.PREC_FP64:
        assert abs(int(d_ref["lstep"][0]) - r_ref["lstep"]) <= max(1, int(0.01 * r_ref["lstep"]))
        assert gsq[K] < gtest <= gsq[K - 1]
        assert _rel(u, runs[K]["u"]) < 1e-10
        assert abs(oc.sumsq(gb) - gsq[K]) < 1e-8 * gsq[K]
    else:
        # FP32 gradient: the crossing may move by up to two iterations either way
        assert min(gsq.values()) < gtest
        assert _rel(u, runs[K]["u"]) < 1e-5


def test_nist_absolute_stop_rule_per_rhs():
    """Three right-hand sides under E3X_STOP_NIST_ABS stop where their single solves do."""
    c11, c12, c44 = C11 * 1e-9, C12 * 1e-9, C44 * 1e-9
    shape = (14, 12, 10)
    pix, rod = voronoi_case(shape, 4, seed=420)
    e3 = np.array([[oc.AML_F32, 0, 0, 0, 0, 0], [0, 0, 0, 0, 2e-3, 0], [0, 5e-4, -5e-4, 0, 0, 0]])
    many = solve(pix, rod, shape, 5, c11, c12, c44, e3, stop_rule=L.STOP_NIST_ABS)
    for r in range(3):
        one = solve(pix, rod, shape, 5, c11, c12, c44, e3[r], stop_rule=L.STOP_NIST_ABS)
        assert many.exit[r] == one.exit[0] == L.PCG_CONVERGED
        assert many.lstep[r] == one.lstep[0]
        assert abs(many.utot[r] - one.utot[0]) < 1e-12 * abs(one.utot[0])


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_isotropic_matrix_void_inclusion_matches_oracle(precision):
    """flag_m = 0, the Eshelby programs' mode (applications/sec3_1/pcg/n_N/
    elas3dxtal_pcg.f90:59-79, 190-198): one isotropic matrix phase with the Voigt bulk
    and shear moduli of the cubic constants, and a centred spherical void (E1 = nu1 = 0)
    under ezz."""
    shape = (24, 22, 20)
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r2 = (i - nx / 2 + 0.5) ** 2 + (j - ny / 2 + 0.5) ** 2 + (k - nz / 2 + 0.5) ** 2
    pix = np.where(r2.ravel() <= 5.0 ** 2, 2, 1).astype(np.int32)
    rod = np.zeros((1, 3))
    k0, g0 = (C11 + 2 * C12) / 3.0, (C11 - C12 + 3 * C44) / 5.0
    pm = (k0, g0, 0.0, 0.0)
    e6 = [0, 0, oc.AML_F32, 0, 0, 0]
    ro = oc.solve(shape, 2, pix, rod, C11, C12, C44, e6, phasemod=pm, flag_m=0)
    res = solve(pix, rod, shape, 2, C11, C12, C44, e6, phasemod=pm, flag_m=0, precision=precision)
    assert res.exit[0] == L.PCG_CONVERGED == ro["status"]
    if precision == L.PREC_FP64:
        assert abs(int(res.lstep[0]) - ro["lstep"]) <= max(1, int(0.01 * ro["lstep"]))
        assert abs(res.utot[0] - ro["utot"]) < 1e-10 * abs(ro["utot"])
        assert _rel(res.stress6[0], ro["stress6"]) < 1e-10
    else:
        assert abs(int(res.lstep[0]) - ro["lstep"]) <= max(2, int(0.05 * ro["lstep"]))
        assert abs(res.utot[0] - ro["utot"]) < 1e-6 * abs(ro["utot"])
        assert _rel(res.stress6[0], ro["stress6"]) < 1e-5
    # the void carries no load: the mean szz falls below the matrix's c33 * ezz
    assert 0 < res.stress6[0][2] < (k0 + 4 * g0 / 3) * oc.AML_F32


@pytest.mark.parametrize("variant,stop_rule,precision", [
    (L.PCG_SWEEP, L.STOP_REFERENCE, L.PREC_FP64), (L.PCG_PLAIN_CG, L.STOP_REFERENCE, L.PREC_MIXED),
    (L.PCG_PLAIN_CG, L.STOP_NIST_ABS, L.PREC_FP64), (L.PCG_REFERENCE, L.STOP_NIST_ABS, L.PREC_MIXED)])
def test_loop_options_per_rhs_equal_single_solves(variant, stop_rule, precision):
    """The loop variants and the NIST stop rule with three right-hand sides in one PCG:
    each stops where its single solve does, with the same energy.  Moduli in GPa for the
    NIST rule (its units); in Pa for the reference's (in GPa its absolute hAh floor,
    f90:1899, would end the loop first)."""
    unit = 1e-9 if stop_rule == L.STOP_NIST_ABS else 1.0
    c11, c12, c44 = C11 * unit, C12 * unit, C44 * unit
    shape = (12, 10, 9)
    pix, rod = voronoi_case(shape, 5, seed=430)
    e3 = np.array([[oc.AML_F32, 0, 0, 0, 0, 0], [0, 0, 0, 0.8e-3, 0, 0], [0, 0.4e-3, 0, 0, 0, -0.3e-3]])
    kw = dict(precision=precision, replace_every=7, stop_rule=stop_rule, variant=variant)
    many = solve(pix, rod, shape, 6, c11, c12, c44, e3, **kw)
    for r in range(3):
        one = solve(pix, rod, shape, 6, c11, c12, c44, e3[r], **kw)
        assert many.exit[r] == one.exit[0] == L.PCG_CONVERGED
        assert many.lstep[r] == one.lstep[0]
        assert abs(many.utot[r] - one.utot[0]) <= 1e-13 * abs(one.utot[0])


@pytest.mark.parametrize("variant", [L.PCG_SWEEP, L.PCG_PLAIN_CG])
@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_pcg_variants_match_oracle(variant, precision):
    """The anisotropy-sweep and plain-CG loops against the oracle's, for stop tolerances
    that end them after an odd and an even number of iterations (the final direction is
    then h_K, held in either direction buffer); then the same handle back on the
    reference loop (the block-Jacobi inverse is rebuilt after the plain CG's identity)."""
    shape = (13, 10, 7)
    pix, rod = voronoi_case(shape, 6, seed=310)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    _, dk_o = oracle_tables(7, rod)
    b_o, c_o = oc.femat_rhs(shape, pix, dk_o, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk_o, b_o, c_o, u0)
    parities = set()
    with Elas3D(shape, 7) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.set_pcg_variant(variant)
        for tol in (1e-3, 1e-4, 1e-5, 1e-6, 1e-7, oc.TOL_F32):
            r_o = oc.dembx(shape, pix, dk_o, u0, g0, oc.sumsq(g0), tol=tol, variant=variant)
            s.affine_guess()
            _, gg0 = s.energy()
            d = s.dembx(gg0, tol=tol, precision=precision)
            u, h = s.get(L.FIELD_U), s.get(L.FIELD_H)
            K = int(d["lstep"][0])
            assert d["exit"][0] == r_o["status"] == L.PCG_CONVERGED
            assert d["reset_count"][0] == 0
            if precision == L.PREC_FP64:
                assert abs(K - r_o["lstep"]) <= max(1, int(0.01 * r_o["lstep"]))
                parities.add(K & 1)
                if K == r_o["lstep"]:
                    assert _rel(u, r_o["u"]) < 1e-10
                    assert _rel(d["history"][0], r_o["history"]) < 1e-8
                    assert _rel(h, r_o["h"]) < 1e-5
                else:
                    # equal iteration counts: the oracle's loop run for exactly K
                    # iterations.  Its direction is then the one after the update these
                    # loops skip on convergence (A_2_0:1518, cg:1050), so only u and the
                    # history compare.
                    r_k = oc.dembx(shape, pix, dk_o, u0, g0, oc.sumsq(g0), tol=0.0, ldemb=K, variant=variant)
                    assert r_k["lstep"] == K
                    assert _rel(u, r_k["u"]) < 1e-10
                    assert _rel(d["history"][0], r_k["history"]) < 1e-8
            else:
                parities.add(K & 1)
                assert abs(K - r_o["lstep"]) <= max(2, int(0.05 * r_o["lstep"]))
                assert _rel(u, r_o["u"]) < 1e-5
        s.set_pcg_variant(L.PCG_REFERENCE)
        s.affine_guess()
        _, gg0 = s.energy()
        d = s.dembx(gg0)
        u = s.get(L.FIELD_U)
    assert parities == {0, 1}
    r_ref = oc.dembx(shape, pix, dk_o, u0, g0, oc.sumsq(g0))
    assert d["exit"][0] == L.PCG_CONVERGED
    K = int(d["lstep"][0])
    assert abs(K - r_ref["lstep"]) <= max(1, int(0.01 * r_ref["lstep"]))
    if K != r_ref["lstep"]:
        r_ref = oc.dembx(shape, pix, dk_o, u0, g0, oc.sumsq(g0), tol=0.0, ldemb=K)
    assert _rel(u, r_ref["u"]) < 1e-10


@pytest.mark.parametrize("ldemb,replace_every", [(15, 0), (16, 0), (17, 0), (17, 8), (18, 7), (16, 3), (14, 1),
                                                  (15, 2)])
def test_mixed_steps_reach_u_once_in_order(ldemb, replace_every):
    """The mixed form stopped by the iteration limit after an odd or an even number of
    iterations, without and with replacements every 1, 2, 3, 7 or 8 iterations.  The u
    steps land two at a time in the applies, or in the update of a replacement
    iteration, or after the loop.  The PCG correction u - u0 must match the FP64
    oracle's after the same number of iterations, to FP32 rounding.  One step lost or
    applied twice is off by more than ten times the tolerance: consecutive iterates here
    differ by ~1e-4 of it."""
    shape = (12, 12, 12)
    pix, rod = voronoi_case(shape, 5, seed=400)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    r_o, _ = _oracle_pcg(shape, pix, rod, e6, ldemb=ldemb)
    assert r_o["status"] == oc.OC_MAXITER
    u0 = oc.affine(shape, e6)
    with Elas3D(shape, 6) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.affine_guess()
        _, gg0 = s.energy()
        d = s.dembx(gg0, ldemb=ldemb, precision=L.PREC_MIXED, replace_every=replace_every)
        u = s.get(L.FIELD_U)
        h = s.get(L.FIELD_H)
    assert d["exit"][0] == L.PCG_MAXITER and d["lstep"][0] == ldemb
    err_u, err_h = _rel(u - u0, r_o["u"] - u0), _rel(h, r_o["h"])
    print(f"ldemb {ldemb} replace {replace_every}: correction {err_u:.2e}, direction {err_h:.2e}")
    assert err_u < 1e-5
    assert err_h < 1e-3
    assert _rel(d["history"][0], r_o["history"]) < 1e-4


@pytest.mark.parametrize("replace_every", [0, 7, 1000])
def test_mixed_precision_against_fp64_oracle(replace_every):
    """FP32 stencil/vectors/preconditioner with FP64 u and dots, true-residual
    replacement every `replace_every` iterations (0: never): the converged solution
    agrees with the FP64 oracle to the level of the stop tolerance."""
    shape = (16, 14, 12)
    pix, rod = voronoi_case(shape, 7, seed=600)
    e6 = [oc.AML_F32, 0, 0, 0, 0.3e-3, 0]
    ro = oc.solve(shape, 8, pix, rod, C11, C12, C44, e6)
    res = solve(pix, rod, shape, 8, C11, C12, C44, e6, precision=L.PREC_MIXED, replace_every=replace_every)
    assert res.exit[0] == L.PCG_CONVERGED
    assert abs(int(res.lstep[0]) - ro["lstep"]) <= max(2, int(0.05 * ro["lstep"]))
    assert abs(res.utot[0] - ro["utot"]) < 1e-6 * abs(ro["utot"])
    assert _rel(res.stress6[0], ro["stress6"]) < 1e-5
    assert _rel(res.strain6[0], ro["strain6"]) < 1e-12


@pytest.mark.parametrize("precision,replace_every", [(L.PREC_FP64, 0), (L.PREC_MIXED, 50), (L.PREC_MIXED, 7),
                                                    (L.PREC_MIXED, 6)])
def test_multi_rhs_equals_single_solves(precision, replace_every):
    """Three right-hand sides in one PCG give each one's single solve (they stop at
    different iterations), also with true-gradient replacements every 7 or 6 iterations
    (the u steps pending at a replacement: two after an even count of iterations since
    the last one, one after an odd count)."""
    shape = (12, 10, 8)
    pix, rod = voronoi_case(shape, 6, seed=500)
    strains = np.array([[oc.AML_F32, 0, 0, 0, 0, 0], [0, 0, 0, 0.5e-3, 0, 0], [0, 0.7e-3, 0, 0, 0, 0.2e-3]])
    batch = solve(pix, rod, shape, 7, C11, C12, C44, strains, precision=precision, replace_every=replace_every)
    for r in range(3):
        one = solve(pix, rod, shape, 7, C11, C12, C44, strains[r], precision=precision, replace_every=replace_every)
        assert batch.lstep[r] == one.lstep[0]
        assert abs(batch.utot[r] - one.utot[0]) <= 1e-13 * abs(one.utot[0])
        assert _rel(batch.stress6[r], one.stress6[0]) < 1e-13


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_multi_rhs_stops_in_later_chunks(precision):
    """Three right-hand sides whose stops fall in three different iteration chunks past
    the first two (the host enqueues 32-iteration chunks on grids below 2^17 nodes and
    examines a chunk's state only once the next one is enqueued): after a stop the host
    has seen, the apply's grid leaves that right-hand side out, and past a predicted
    stop the chunks end one iteration after it.  Each right-hand side must still give its
    single solve (bitwise, same precision) and, in FP64, the oracle's iterate at its stop
    iteration.  tol = 1e-10 so that the stops come late (oracle: 89, 111, 147)."""
    shape = (40, 40, 40)
    pix, rod = voronoi_case(shape, 30, seed=521, pore_frac=0.03)
    c11 = C12 + 2 * C44 / 5.0
    strains = np.array([[1e-3, 1e-3, 1e-3, 0, 0, 0], [1e-3, 1e-3, 1e-3, 1e-4, 0, 0], [oc.AML_F32, 0, 0, 0, 0, 0]])
    kw = dict(precision=precision, replace_every=50, tol=1e-10)
    batch, s = solve(pix, rod, shape, 31, c11, C12, C44, strains, keep=True, **kw)
    try:
        u_b = [s.get(L.FIELD_U, r) for r in range(3)]
    finally:
        s.close()
    stops = [int(v) for v in batch.lstep]
    print("stops", stops)
    assert len({(k - 1) // 32 for k in stops}) == 3 and min(stops) > 64
    for r in range(3):
        one, s1 = solve(pix, rod, shape, 31, c11, C12, C44, strains[r], keep=True, **kw)
        try:
            u1 = s1.get(L.FIELD_U)
        finally:
            s1.close()
        assert batch.exit[r] == one.exit[0] == L.PCG_CONVERGED
        assert batch.lstep[r] == one.lstep[0]
        assert batch.utot[r] == one.utot[0]
        assert np.array_equal(u_b[r], u1)
        assert np.array_equal(batch.history[r], one.history[0])
        if precision == L.PREC_FP64:
            _check_deep_fp64_against_oracle(shape, pix, rod, c11, strains[r], int(batch.lstep[r]), u_b[r],
                                            batch.history[r], batch.utot[r])


def _check_deep_fp64_against_oracle(shape, pix, rod, c11, e6, K, u, history, utot):
    """FP64 iterate at stop iteration K against the oracle's at K, for solves driven far
    below the reference's tolerance.  There the PCG amplifies rounding: started from its
    own initial gradient perturbed by one ulp, the oracle's iterate after ~120 iterations
    of the A = 5 case moves by ~2e-8 and its residual history by ~2e-2.  So the bar is ten
    times the oracle's own one-ulp spread at K, and never looser than the usual 1e-10 (u)
    and 1e-8 (history)."""
    nphase = len(rod) + 1
    dk_o = oc.dk(oc.cmod(nphase, oc.rotate_stiffness(rod, c11, C12, C44), [0.0, 0.0]))
    b_o, c_o = oc.femat_rhs(shape, pix, dk_o, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk_o, b_o, c_o, u0)
    gg0 = oc.sumsq(g0)
    ro = oc.dembx(shape, pix, dk_o, u0, g0, gg0, tol=1e-10)
    assert abs(K - ro["lstep"]) <= max(1, int(0.01 * ro["lstep"]))
    if K != ro["lstep"]:
        ro = oc.dembx(shape, pix, dk_o, u0, g0, gg0, tol=0.0, ldemb=K)
    rng = np.random.default_rng(K)
    g1 = g0 * (1.0 + 2.0 ** -53 * rng.choice([-1.0, 1.0], g0.size))
    rp = oc.dembx(shape, pix, dk_o, u0, g1, gg0, tol=0.0, ldemb=K)
    spread_u, spread_h = _rel(rp["u"], ro["u"]), _rel(rp["history"], ro["history"])
    _, ut_o = oc.energy(shape, pix, dk_o, b_o, c_o, ro["u"])
    _, ut_p = oc.energy(shape, pix, dk_o, b_o, c_o, rp["u"])
    spread_e = abs(ut_p - ut_o) / abs(ut_o)
    err_u, err_h, err_e = _rel(u, ro["u"]), _rel(history, ro["history"]), abs(utot - ut_o) / abs(ut_o)
    print(f"K {K}: u {err_u:.2e} (oracle one-ulp spread {spread_u:.2e}), history {err_h:.2e} ({spread_h:.2e}), "
          f"energy {err_e:.2e} ({spread_e:.2e})")
    assert err_u < max(1e-10, 10 * spread_u)
    assert err_h < max(1e-8, 10 * spread_h)
    assert err_e < max(1e-10, 10 * spread_e)


@pytest.mark.parametrize("precision,tol", [(L.PREC_FP64, 1e-10), (L.PREC_MIXED, 1e-6)])
def test_effective_stiffness_matches_oracle(precision, tol):
    """Six unit load cases as six right-hand sides of one PCG: the 6x6 effective
    stiffness against six oracle solves (FP64 within 1e-10, the mixed form within 1e-6)."""
    shape = (10, 9, 8)
    ng = 5
    pix, rod = voronoi_case(shape, ng, seed=800, pore_frac=0.02)
    S, E = [], []
    for e6 in unit_strains(oc.AML_F32):
        r = oc.solve(shape, ng + 1, pix, rod, C11, C12, C44, list(e6), want_fields=False)
        S.append(r["stress6"])
        E.append(r["strain6"])
    ce_o = ceff_from_cases(S, E)
    ce, lstep, ex = effective_stiffness(pix, rod, shape, ng + 1, C11, C12, C44, magnitude=oc.AML_F32,
                                        precision=precision)
    assert np.all(ex == L.PCG_CONVERGED)
    assert _rel(ce, ce_o) < tol


def test_effective_stiffness_single_crystal():
    shape = (6, 5, 4)
    ce, lstep, ex = effective_stiffness(np.ones(120, np.int32), np.zeros(3), shape, 2, C11, C12, C44)
    c = np.zeros((6, 6))
    c[:3, :3] = C12
    c[[0, 1, 2], [0, 1, 2]] = C11
    c[[3, 4, 5], [3, 4, 5]] = C44
    assert np.all(ex == L.PCG_HAH_BREAK)      # the affine guess is the solution
    assert _rel(ce, c) < 1e-12


def test_more_than_65535_phases_matches_oracle():
    """A 48^3 grid with 70,000 grains (the reference's anisotropy-sweep inputs have 81,488,
    applications/sec3_2/A_2_0/elas3dxtal_pcg.f90:59) plus the pore phase: the handle holds
    32-bit phase ids.  Labels above 65,535 occur; the solve matches the oracle as the
    16-bit path does."""
    from elas3dxtal import io
    shape, ng = (48, 48, 48), 70000
    pix, rod, _, _ = io.generate(shape, ng, 4801, pore_mode=io.PORES_VOXELS, pore_fraction=0.02)
    assert pix.max() == ng + 1 and np.count_nonzero(pix > 65535) > 1000
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    ro = oc.solve(shape, ng + 1, pix, rod, C11, C12, C44, e6)
    res, s = solve(pix, rod, shape, ng + 1, C11, C12, C44, e6, keep=True)
    try:
        u = s.get(L.FIELD_U)
        prob = s.assig()
        _, vm = s.stress_fullfield()
    finally:
        s.close()
    K = int(res.lstep[0])
    print(f"70,001 phases: lstep GPU {K}, oracle {ro['lstep']}")
    assert res.exit[0] == L.PCG_CONVERGED
    assert abs(K - ro["lstep"]) <= max(1, int(0.01 * ro["lstep"]))
    if K != ro["lstep"]:
        ro = oc.solve(shape, ng + 1, pix, rod, C11, C12, C44, e6, tol=0.0, ldemb=K)
    assert np.array_equal(prob, oc.assig(pix, ng + 1))
    assert abs(res.utot0[0] - ro["utot0"]) < 1e-10 * abs(ro["utot0"])
    assert abs(res.utot[0] - ro["utot"]) < 1e-10 * abs(ro["utot"])
    assert _rel(u, ro["u"]) < 1e-10
    assert _rel(res.stress6[0], ro["stress6"]) < 1e-10
    assert _rel(res.history[0], ro["history"]) < 1e-8
    cm_o = oc.cmod(ng + 1, oc.rotate_stiffness(rod, C11, C12, C44), [0.0, 0.0])
    _, vm_o = oc.fullfield(shape, pix, cm_o, e6, ro["u"])
    assert _rel(vm, vm_o) < 1e-9


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_wide_phase_ids_give_the_same_bits(precision):
    """The same microstructure on a handle with 16-bit phase ids and on one with 32-bit
    ids (nphase above 65,535, the extra grains unused): the lane order of the stencil
    tiles ranks phases within each tile, so both take the same order and every sum the
    same order of operations: identical results."""
    shape = (20, 18, 16)
    pix, rod = voronoi_case(shape, 9, seed=4802, pore_frac=0.03)
    pix = np.where(pix == 10, 70001, pix).astype(np.int32)   # the pore phase as the last of 70,001
    rng = np.random.default_rng(4803)
    rod_wide = np.concatenate([rod, ind.random_rodrigues(rng, 70000 - 9)])
    pix_narrow = np.where(pix == 70001, 10, pix).astype(np.int32)
    e6 = [oc.AML_F32, 0, 0, 0, 0.3e-3, 0]
    kw = dict(precision=precision, replace_every=7)
    narrow, s1 = solve(pix_narrow, rod, shape, 10, C11, C12, C44, e6, keep=True, **kw)
    wide, s2 = solve(pix, rod_wide, shape, 70001, C11, C12, C44, e6, keep=True, **kw)
    try:
        u1, u2 = s1.get(L.FIELD_U), s2.get(L.FIELD_U)
    finally:
        s1.close()
        s2.close()
    assert narrow.lstep[0] == wide.lstep[0] and narrow.exit[0] == wide.exit[0] == L.PCG_CONVERGED
    assert np.array_equal(u1, u2)
    assert narrow.utot[0] == wide.utot[0]
    assert np.array_equal(narrow.stress6[0], wide.stress6[0])


@pytest.mark.parametrize("precision,variants", [(L.PREC_FP64, (5, 6, 8)), (L.PREC_MIXED, (6, 8))])
def test_apply_launch_variants_give_the_same_bits(precision, variants):
    """The tuner's launch variants of the PCG apply (waves per SIMD of its register budget)
    compute the same bits; with tuning left to the library it picks one of them at the
    first PCG and reports it."""
    shape = (24, 20, 16)
    pix, rod = voronoi_case(shape, 8, seed=4901, pore_frac=0.02)
    e6 = [oc.AML_F32, 0, 0, 0, 0.3e-3, 0]
    out = {}
    for w in variants + (0,):
        with Elas3D(shape, 9) as s:
            s.ppixel(pix, rod)
            s.set_tuning(precision, w)
            s.femat(C11, C12, C44, e6)
            s.affine_guess()
            _, gg0 = s.energy()
            d = s.dembx(gg0, precision=precision, replace_every=7)
            ut, _ = s.energy()
            out[w] = (int(d["lstep"][0]), float(ut[0]), s.get(L.FIELD_U), s.tuning(precision))
    ref = out[variants[0]]
    for w, (lstep, ut, u, tun) in out.items():
        assert lstep == ref[0] and ut == ref[1] and np.array_equal(u, ref[2]), w
        assert tun[0] == (w if w else tun[0]) and tun[0] in variants
    assert out[0][3][1] > 0     # measured


def test_bad_phase_label_and_call_order():
    with Elas3D((4, 4, 4), 3) as s:
        with pytest.raises(L.E3xError) as ei:
            s.energy()
        assert ei.value.code == L.E3X_ESTATE
        bad = np.ones(64, np.int32)
        bad[5] = 4
        with pytest.raises(L.E3xError) as ei:
            s.ppixel(bad, np.zeros(6))
        assert ei.value.code == L.E3X_EPHASE
        bad[5] = 0
        with pytest.raises(L.E3xError) as ei:
            s.ppixel(bad, np.zeros(6))
        assert ei.value.code == L.E3X_EPHASE
