"""Pins of the CPU oracle against mathematics (no GPU).

The reference ships no tests, fixtures or golden outputs, and it cannot be built
in this image (its HDF5 Fortran module is gfortran-only).  Each oracle function is
therefore checked here against an independent formulation (tests/_indep.py),
closed forms, or brute force on tiny grids.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import _indep as ind
import _oracle as oc
from _cases import hot_fixture

C11, C12, C44 = 246.5e9, 147.3e9, 124.7e9


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ---------------------------------------------------------------- ib table (f90:326-369)
@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 4, 4), (5, 5, 5), (3, 4, 5), (6, 3, 4)])
def test_ib_matches_rolled_index_grid(shape):
    nx, ny, nz = shape
    ib = oc.build_ib(nx, ny, nz)
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    im = [0, 1, 1, 1, 0, -1, -1, -1] * 3 + [0, 0, 0]
    jm = [1, 1, 0, -1, -1, -1, 0, 1] * 3 + [0, 0, 0]
    km = [0] * 8 + [-1] * 8 + [1] * 8 + [-1, 1, 0]
    for n in range(27):
        rolled = np.roll(idx, shift=(-km[n], -jm[n], -im[n]), axis=(0, 1, 2)).ravel()
        assert np.array_equal(ib[n], rolled), n
    # the 27 neighbours are distinct when every extent is >= 3
    assert np.all(np.sort(ib, axis=0)[1:] != np.sort(ib, axis=0)[:-1])


# ---------------------------------------------------------------- rotation (f90:2470-2637)
def test_rodrigues_matrix_is_active_rotation_by_2atan():
    rng = np.random.default_rng(1)
    for r in ind.random_rodrigues(rng, 50):
        ang = 2.0 * np.arctan(np.linalg.norm(r))
        ref = Rotation.from_rotvec(ang * r / np.linalg.norm(r)).as_matrix()
        assert np.allclose(oc.rodrigues_matrix(r), ref, atol=1e-14)
    assert np.array_equal(oc.rodrigues_matrix(np.zeros(3)), np.eye(3))


def test_rotated_stiffness_equals_tensor_rotation():
    rng = np.random.default_rng(2)
    rod = ind.random_rodrigues(rng, 40)
    rot = oc.rotate_stiffness(rod, C11, C12, C44)
    c4 = ind.cubic_tensor(C11, C12, C44)
    for g in range(len(rod)):
        q = oc.rodrigues_matrix(rod[g])
        ref = ind.voigt(ind.rotate_tensor(c4, q))
        assert _rel(rot[g], ref) < 1e-13, g
    # a plausible slip (transposed Q) must fail this check
    q = oc.rodrigues_matrix(rod[0])
    assert _rel(rot[0], ind.voigt(ind.rotate_tensor(c4, q.T))) > 1e-3


def test_rotation_closed_forms():
    rng = np.random.default_rng(3)
    rod = ind.random_rodrigues(rng, 20)
    c_cubic = ind.voigt(ind.cubic_tensor(C11, C12, C44))
    # r = 0 gives the crystal stiffness
    assert _rel(oc.rotate_stiffness(np.zeros(3), C11, C12, C44)[0], c_cubic) < 1e-15
    # Zener A = 1 (2 C44 = C11 - C12) is isotropic: invariant under every rotation
    c11_iso = C12 + 2 * C44
    rot = oc.rotate_stiffness(rod, c11_iso, C12, C44)
    c_iso = ind.voigt(ind.cubic_tensor(c11_iso, C12, C44))
    assert _rel(rot, np.broadcast_to(c_iso, rot.shape)) < 1e-13
    # Mandel-scaled eigenvalues are rotation invariant: {C11+2C12, C11-C12 x2, 2C44 x3}
    mand = np.diag([1, 1, 1, np.sqrt(2.0), np.sqrt(2.0), np.sqrt(2.0)])
    want = np.sort([C11 + 2 * C12, C11 - C12, C11 - C12, 2 * C44, 2 * C44, 2 * C44])
    for r6 in oc.rotate_stiffness(rod, C11, C12, C44):
        ev = np.sort(np.linalg.eigvalsh(mand @ r6 @ mand))
        assert _rel(ev, want) < 1e-12


# ---------------------------------------------------------------- cmod + dk (f90:655-805)
def test_dk_equals_gauss_quadrature_with_swapped_shear_order():
    """dk integrates B^T cmod B with B rows (xx,yy,zz,xz,yz,xy) (f90:769-780) while cmod
    is in (xx,yy,zz,yz,xz,xy) order, so the physical stiffness is P45 cmod P45."""
    rng = np.random.default_rng(4)
    rod = ind.random_rodrigues(rng, 6)
    rot = oc.rotate_stiffness(rod, C11, C12, C44)
    cm = oc.cmod(7, rot, [0.0, 0.0])
    assert np.all(cm[6] == 0.0)
    dk = oc.dk(cm)
    for p in range(6):
        ref = ind.element_stiffness_gauss(ind.P45 @ cm[p] @ ind.P45)
        assert _rel(dk[p].reshape(24, 24), ref) < 1e-13, p
        # without the swap the match fails for a general orientation
        bad = ind.element_stiffness_gauss(cm[p])
        assert _rel(dk[p].reshape(24, 24), bad) > 1e-3


def test_dk_symmetry_and_rigid_body_null_space():
    rng = np.random.default_rng(5)
    rot = oc.rotate_stiffness(ind.random_rodrigues(rng, 4), C11, C12, C44)
    dk = oc.dk(oc.cmod(5, rot, [0.0, 0.0]))
    scale = np.max(np.abs(dk))
    for p in range(4):
        k = dk[p].reshape(24, 24)
        assert np.max(np.abs(k - k.T)) < 1e-13 * scale
        for c in range(3):                      # rigid translations
            t = np.zeros(24); t[c::3] = 1.0
            assert np.max(np.abs(k @ t)) < 1e-13 * scale
        for w in np.eye(3):                     # infinitesimal rigid rotations u = w x X
            v = np.cross(w, ind.POS.astype(float)).reshape(24)
            assert np.max(np.abs(k @ v)) < 1e-13 * scale
        assert np.min(np.linalg.eigvalsh(k)) > -1e-12 * scale


def test_cmod_isotropic_mode():
    k1, g1, k2, g2 = 160e9, 80e9, 10e9, 5e9
    cm = oc.cmod(2, None, [k1, g1, k2, g2], flag_m=0)
    for p, (kk, gg) in enumerate([(k1, g1), (k2, g2)]):
        lam = kk - 2.0 * gg / 3.0
        want = np.zeros((6, 6))
        want[:3, :3] = lam
        want[np.arange(3), np.arange(3)] = lam + 2 * gg
        want[np.arange(3, 6), np.arange(3, 6)] = gg
        assert _rel(cm[p], want) < 1e-15


# ---------------------------------------------------------------- grids / inputs
def _random_case(shape, ngrain, seed, pore_frac=0.1):
    rng = np.random.default_rng(seed)
    nphase = ngrain + 1
    ns = int(np.prod(shape))
    pix = rng.integers(1, ngrain + 1, ns).astype(np.int32)
    pix[rng.random(ns) < pore_frac] = nphase
    rod = ind.random_rodrigues(rng, ngrain)
    rot = oc.rotate_stiffness(rod, C11, C12, C44)
    cm = oc.cmod(nphase, rot, [0.0, 0.0])
    return pix, rod, cm, oc.dk(cm)


# ---------------------------------------------------------------- stencil (f90:1176-1321, 1585-1884)
@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 3, 5), (4, 4, 4)])
def test_stencil_equals_dense_assembly(shape):
    pix, _, _, dk = _random_case(shape, 3, seed=10 + shape[0])
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    ns = int(np.prod(shape))
    rng = np.random.default_rng(11)
    for _ in range(3):
        x = rng.standard_normal(3 * ns)
        assert _rel(oc.apply(shape, pix, dk, x), kg @ x) < 1e-13
    # symmetry and rigid translations
    x, y = rng.standard_normal(3 * ns), rng.standard_normal(3 * ns)
    assert abs(x @ oc.apply(shape, pix, dk, y) - y @ oc.apply(shape, pix, dk, x)) < 1e-12 * np.abs(kg).max() * ns
    t = np.zeros(3 * ns); t[:ns] = 1.0
    assert np.max(np.abs(oc.apply(shape, pix, dk, t))) < 1e-12 * np.abs(kg).max()


# ---------------------------------------------------------------- b, C, energy (f90:815-1100, 1107-1353)
@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 3, 5)])
def test_rhs_and_constant_equal_jump_assembly(shape):
    pix, _, _, dk = _random_case(shape, 3, seed=20)
    e6 = [1.1e-3, -0.4e-3, 0.7e-3, 0.3e-3, -0.2e-3, 0.5e-3]
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    b_ref, c_ref = ind.rhs_from_jumps(shape, pix, ind.dk_to_kmats(dk), e6)
    assert _rel(b, b_ref) < 1e-13
    assert abs(cval - c_ref) < 1e-13 * abs(c_ref)


def test_jump_constant_to_rounding_on_a_large_grid():
    """The constant C summed over the ~5e4 boundary voxels of a 128^3 grid agrees with the
    exact sum of the voxel terms (computed independently, grouped by wrap pattern) to a
    few ulps: summed as one running total, as the reference's reduction does, it drifts by
    ~1e-12 here, which the energy amplifies ~200-fold through the cancellation of C
    against b.u."""
    shape = (128, 128, 128)
    pix, _, _, dk = _random_case(shape, 40, seed=21)
    e6 = [1e-3, 0.0, 0.0, 0.0, 0.0, 0.0]
    _, cval = oc.femat_rhs(shape, pix, dk, e6)
    exact = ind.jump_constant_exact(shape, pix, ind.dk_to_kmats(dk), e6)
    print(f"C oracle {cval!r}, exact {exact!r}, rel {abs(cval - exact) / abs(exact):.2e}")
    assert abs(cval - exact) < 1e-15 * abs(exact)


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 5, 3)])
def test_energy_equals_total_field_element_energy(shape):
    pix, _, _, dk = _random_case(shape, 4, seed=30)
    e6 = [0.0, 0.0, 1e-3, 0.4e-3, 0.0, -0.3e-3]
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    rng = np.random.default_rng(31)
    ns = int(np.prod(shape))
    u = 1e-3 * rng.standard_normal(3 * ns)
    gb, utot = oc.energy(shape, pix, dk, b, cval, u)
    ref = ind.element_energy_total(shape, pix, ind.dk_to_kmats(dk), u, e6)
    assert abs(utot - ref) < 1e-12 * abs(ref)
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    assert _rel(gb, kg @ u + b) < 1e-12


def test_homogeneous_affine_field_is_exact_equilibrium():
    """Single grain, no pores: u = E x solves the periodic problem, so g = Ku + b = 0 and
    utot = ns/2 * eps^T cmod eps with eps in the B-row order (xx,yy,zz,2xz,2yz,2xy)."""
    shape = (5, 4, 6)
    ns = int(np.prod(shape))
    rod = np.array([[0.3, -0.2, 0.5]])
    cm = oc.cmod(2, oc.rotate_stiffness(rod, C11, C12, C44), [0.0, 0.0])
    dk = oc.dk(cm)
    pix = np.ones(ns, np.int32)
    e6 = np.array([0.9e-3, -0.3e-3, 0.2e-3, 0.4e-3, -0.1e-3, 0.25e-3])
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u = oc.affine(shape, e6)
    gb, utot = oc.energy(shape, pix, dk, b, cval, u)
    assert np.max(np.abs(gb)) < 1e-10 * np.max(np.abs(b))
    eps = np.array([e6[0], e6[1], e6[2], 2 * e6[3], 2 * e6[4], 2 * e6[5]])
    want = 0.5 * ns * eps @ cm[0] @ eps
    assert abs(utot - want) < 1e-12 * abs(want)


# ---------------------------------------------------------------- preconditioner (f90:1498-1520, 2046-2147)
def test_invert3x3_and_singular_block():
    rng = np.random.default_rng(40)
    for _ in range(50):
        a = rng.standard_normal((3, 3)) * 1e11
        assert _rel(oc.invert3x3(a), np.linalg.inv(a)) < 1e-12
    assert np.all(oc.invert3x3(np.zeros((3, 3))) == 0.0)
    assert np.all(oc.invert3x3(np.full((3, 3), 1e-6)) == 0.0)   # |det| < 1e-14


def test_precond_blocks_are_inverse_dense_diagonal_blocks():
    shape = (4, 3, 3)
    pix, _, _, dk = _random_case(shape, 2, seed=41, pore_frac=0.0)
    ns = int(np.prod(shape))
    pix[:] = np.where(np.arange(ns) < 8, 3, pix)                 # some pores
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    mi = oc.precond(shape, pix, dk)
    for m in range(ns):
        blk = kg[np.ix_([m, ns + m, 2 * ns + m], [m, ns + m, 2 * ns + m])]
        assert _rel(mi[m], np.linalg.inv(blk)) < 1e-12
    # a node whose 8 voxels are all pore has a zero block
    pix_all = np.full(ns, 3, np.int32)
    assert np.all(oc.precond(shape, pix_all, dk) == 0.0)


# ---------------------------------------------------------------- PCG (f90:1356-2043)
def test_pcg_converges_to_dense_direct_solution():
    shape = (4, 4, 3)
    ns = int(np.prod(shape))
    pix, rod, cm, dk = _random_case(shape, 3, seed=50, pore_frac=0.05)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk, b, cval, u0)
    r = oc.dembx(shape, pix, dk, u0, g0, oc.sumsq(g0), tol=1e-13, ldemb=2000)
    assert r["status"] == oc.OC_CONVERGED
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    u_direct = np.linalg.lstsq(kg, -b, rcond=1e-13)[0]
    # u is unique modulo the null space: compare strain fields and energies
    sf1, _ = oc.fullfield(shape, pix, cm, e6, r["u"])
    sf2, _ = oc.fullfield(shape, pix, cm, e6, u_direct)
    assert _rel(sf1, sf2) < 1e-7
    _, ut1 = oc.energy(shape, pix, dk, b, cval, r["u"])
    _, ut2 = oc.energy(shape, pix, dk, b, cval, u_direct)
    assert abs(ut1 - ut2) < 1e-10 * abs(ut2)


def test_pcg_reference_stop_rule_and_history():
    shape = (6, 5, 4)
    pix, rod, cm, dk = _random_case(shape, 5, seed=51, pore_frac=0.02)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u0 = oc.affine(shape, e6)
    g0, ut0 = oc.energy(shape, pix, dk, b, cval, u0)
    gg0 = oc.sumsq(g0)
    r = oc.dembx(shape, pix, dk, u0, g0, gg0)
    assert r["status"] == oc.OC_CONVERGED
    h = r["history"]
    assert len(h) == r["lstep"]
    assert h[-1] < oc.TOL_F32 and np.all(h[:-1] >= oc.TOL_F32)     # first crossing exits
    assert abs(h[-1] - np.sqrt(r["gg"] / gg0)) <= 1e-15 * h[-1]
    _, ut1 = oc.energy(shape, pix, dk, b, cval, r["u"])
    assert ut1 < ut0                                              # energy decreased


def test_pcg_divergence_reset_matches_dense_restatement():
    """The 10x divergence reset (f90:1961-1981) on an indefinite case: the oracle's
    resets, iteration count, history and solution equal those of a dense-matrix PCG
    written independently from the reference's text."""
    from _cases import RESET_CASE as rc, voronoi_case
    shape = rc["shape"]
    ns = int(np.prod(shape))
    pix, rod = voronoi_case(shape, rc["ngrain"], rc["seed"], rc["pore_frac"])
    nphase = rc["ngrain"] + 1
    cm = oc.cmod(nphase, oc.rotate_stiffness(rod, C11, C12, C44), list(rc["phasemod"]))
    dk = oc.dk(cm)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk, b, cval, u0)
    gg0 = oc.sumsq(g0)
    r = oc.dembx(shape, pix, dk, u0, g0, gg0)
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    d = ind.pcg_reference_dense(kg, ind.block_jacobi_dense(kg, ns), u0, g0, gg0, oc.TOL_F32, 10000)
    assert d["resets"] >= 2
    assert r["reset_count"] == d["resets"]
    assert r["status"] == oc.OC_CONVERGED and d["status"] == 1
    assert r["lstep"] == d["lstep"]
    # the indefinite operator amplifies summation-order rounding (stencil vs dense
    # product) to ~1e-6 in the residual history; the reset decisions are unaffected
    assert _rel(r["history"], d["history"]) < 1e-5
    assert _rel(r["u"], d["u"]) < 1e-6
    assert _rel(r["h"], d["h"]) < 1e-2       # h is of the size of the final residual


def _pcg_case(shape, ngrain, seed):
    pix, rod, cm, dk = _random_case(shape, ngrain, seed=seed, pore_frac=0.03)
    e6 = [oc.AML_F32, 0, 0, 0, 0.4e-3, 0]
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk, b, cval, u0)
    return pix, dk, u0, g0, oc.sumsq(g0)


@pytest.mark.parametrize("variant", [oc.PCG_SWEEP, oc.PCG_PLAIN_CG])
def test_pcg_variants_match_dense_restatement(variant):
    """The sweep and plain-CG loops of the oracle against the dense-matrix loop written
    from those programs' text (no reset, stop test before the direction update; plain
    CG without preconditioner)."""
    shape = (5, 4, 4)
    ns = int(np.prod(shape))
    pix, dk, u0, g0, gg0 = _pcg_case(shape, 3, seed=52)
    r = oc.dembx(shape, pix, dk, u0, g0, gg0, variant=variant)
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    minv = None if variant == oc.PCG_PLAIN_CG else ind.block_jacobi_dense(kg, ns)
    d = ind.pcg_reference_dense(kg, minv, u0, g0, gg0, oc.TOL_F32, 10000, reset=False, test_first=True)
    assert r["status"] == oc.OC_CONVERGED and d["status"] == 1
    assert r["lstep"] == d["lstep"] and r["reset_count"] == 0
    assert _rel(r["history"], d["history"]) < 1e-8
    assert _rel(r["u"], d["u"]) < 1e-10
    assert _rel(r["h"], d["h"]) < 1e-4       # h is of the size of the final residual


def test_pcg_directions_are_conjugate():
    """Preconditioned CG with the Fletcher-Reeves gamma: successive directions are
    K-conjugate (h_i' K h_j = 0, i != j) and the preconditioned gradients M^-1-orthogonal
    (g_i' M^-1 g_j = 0).  A wrong gamma, a transposed block or a dropped term breaks it."""
    shape = (4, 4, 3)
    ns = int(np.prod(shape))
    pix, dk, u0, g0, gg0 = _pcg_case(shape, 3, seed=55)
    kg = ind.assemble_dense(shape, pix, ind.dk_to_kmats(dk))
    minv = ind.block_jacobi_dense(kg, ns)
    runs = [oc.dembx(shape, pix, dk, u0, g0, gg0, tol=0.0, ldemb=k) for k in range(1, 6)]
    hs = [r["h"] for r in runs]                        # directions 2..6
    for i in range(len(hs)):
        for j in range(i):
            c = abs(hs[i] @ kg @ hs[j]) / np.sqrt((hs[i] @ kg @ hs[i]) * (hs[j] @ kg @ hs[j]))
            assert c < 1e-8, (i, j, c)
    gs = [g0] + [r["gb"] for r in runs]

    def prec(v):
        return np.einsum("mij,mj->mi", minv, v.reshape(3, ns).T).T.ravel()

    for i in range(len(gs)):
        for j in range(i):
            c = abs(gs[i] @ prec(gs[j])) / np.sqrt((gs[i] @ prec(gs[i])) * (gs[j] @ prec(gs[j])))
            assert c < 1e-8, (i, j, c)


def test_plain_cg_residuals_are_orthogonal():
    """Textbook CG (M = I, Fletcher-Reeves gamma): successive gradients are mutually
    orthogonal.  A preconditioner left in, or a wrong gamma, breaks it."""
    shape = (4, 4, 3)
    pix, dk, u0, g0, gg0 = _pcg_case(shape, 3, seed=53)
    gs = [g0] + [oc.dembx(shape, pix, dk, u0, g0, gg0, tol=0.0, ldemb=k, variant=oc.PCG_PLAIN_CG)["gb"]
                 for k in range(1, 7)]
    for i in range(len(gs)):
        for j in range(i):
            c = abs(gs[i] @ gs[j]) / (np.linalg.norm(gs[i]) * np.linalg.norm(gs[j]))
            assert c < 1e-8, (i, j, c)


def test_sweep_loop_is_the_reference_loop_without_its_last_direction():
    """Without a reset the sweep loop takes the reference loop's steps; it stops before
    the direction update, so its final h is the direction the reference loop holds after
    K - 1 iterations."""
    shape = (6, 5, 4)
    pix, dk, u0, g0, gg0 = _pcg_case(shape, 4, seed=54)
    src = oc.dembx(shape, pix, dk, u0, g0, gg0)
    sw = oc.dembx(shape, pix, dk, u0, g0, gg0, variant=oc.PCG_SWEEP)
    assert src["reset_count"] == 0 and src["lstep"] == sw["lstep"]
    assert np.array_equal(src["u"], sw["u"]) and np.array_equal(src["history"], sw["history"])
    km1 = oc.dembx(shape, pix, dk, u0, g0, gg0, tol=0.0, ldemb=sw["lstep"] - 1)
    assert np.array_equal(sw["h"], km1["h"])
    assert not np.array_equal(src["h"], sw["h"])


def test_pcg_hah_break_on_all_pore_grid():
    shape = (3, 3, 3)
    ns = 27
    pix = np.full(ns, 2, np.int32)
    cm = oc.cmod(2, oc.rotate_stiffness(np.zeros(3), C11, C12, C44), [0.0, 0.0])
    dk = oc.dk(cm)
    u = oc.affine(shape, [1e-3, 0, 0, 0, 0, 0])
    g = np.zeros(3 * ns)
    r = oc.dembx(shape, pix, dk, u, g, 0.0)
    assert r["status"] == oc.OC_HAH_BREAK and r["lstep"] == 1
    assert np.array_equal(r["u"], u)


# ---------------------------------------------------------------- stress (f90:2149-2408, 2639-2833)
def test_average_strain_equals_applied_strain_for_any_field():
    shape = (5, 4, 3)
    ns = int(np.prod(shape))
    pix, _, cm, _ = _random_case(shape, 3, seed=60)
    e6 = [0.8e-3, -0.2e-3, 0.3e-3, 0.15e-3, -0.05e-3, 0.4e-3]
    rng = np.random.default_rng(61)
    u = 1e-2 * rng.standard_normal(3 * ns)      # periodic fluctuation, arbitrary
    s, e = oc.stress(shape, pix, cm, e6, u)
    want = [e6[0], e6[1], e6[2], 2 * e6[3], 2 * e6[4], 2 * e6[5]]
    assert np.allclose(e, want, rtol=1e-12, atol=1e-16)


def test_fullfield_vm_and_means():
    shape = (4, 4, 5)
    ns = int(np.prod(shape))
    pix, _, cm, _ = _random_case(shape, 3, seed=62)
    e6 = [0, 0, 1e-3, 0, 0, 0]
    rng = np.random.default_rng(63)
    u = oc.affine(shape, e6) + 1e-4 * rng.standard_normal(3 * ns)
    sf, vm = oc.fullfield(shape, pix, cm, e6, u)
    s = sf.T
    dev = s[:, :3] - s[:, :3].mean(axis=1, keepdims=True)
    j2 = 0.5 * np.sum(dev ** 2, axis=1) + np.sum(s[:, 3:] ** 2, axis=1)
    assert _rel(vm, np.sqrt(3.0 * j2)) < 1e-12
    st, _ = oc.stress(shape, pix, cm, e6, u)
    assert _rel(sf.mean(axis=1), st) < 1e-12


def test_hill_condition_at_convergence():
    shape = (6, 6, 5)
    ns = int(np.prod(shape))
    pix, rod, _, _ = _random_case(shape, 4, seed=70, pore_frac=0.03)
    e6 = [oc.AML_F32, 0, 0, 0, 0, 0]
    r = oc.solve(shape, 5, pix, rod, C11, C12, C44, e6, tol=1e-12)
    # utot/ns = 1/2 <sigma> : <eps> at equilibrium
    assert abs(r["utot"] / ns - 0.5 * r["stress6"][0] * e6[0]) < 1e-6 * abs(r["utot"] / ns)


def test_assig_fractions():
    pix = np.array([1, 1, 2, 3, 3, 3, 9], np.int32)
    prob = oc.assig(pix, 3)
    assert np.array_equal(prob, np.array([2, 1, 3]) / 7.0)


# ---------------------------------------------------------------- effective stiffness
def _oracle_ceff(shape, nphase, pix, rod, magnitude=1e-3):
    from _cases import ceff_from_cases, unit_strains
    S, E = [], []
    for e6 in unit_strains(magnitude):
        r = oc.solve(shape, nphase, pix, rod, C11, C12, C44, list(e6), want_fields=False)
        S.append(r["stress6"])
        E.append(r["strain6"])
    return ceff_from_cases(S, E)


def test_effective_stiffness_single_crystal_is_cubic():
    """One grain without pores: the affine field is the solution, C_eff = the cubic matrix
    (xz, yz, xy shear order)."""
    shape = (4, 3, 5)
    ns = int(np.prod(shape))
    ce = _oracle_ceff(shape, 2, np.ones(ns, np.int32), np.zeros(3))
    c = np.zeros((6, 6))
    c[:3, :3] = C12
    c[[0, 1, 2], [0, 1, 2]] = C11
    c[[3, 4, 5], [3, 4, 5]] = C44
    assert _rel(ce, c) < 1e-12


def test_effective_stiffness_between_reuss_and_voigt():
    """A pore-free polycrystal: C_eff is symmetric and lies between the Reuss and Voigt
    averages of the grain stiffnesses (the affine field is admissible, and the discrete
    solution is stiffer than the continuum one)."""
    from _cases import voronoi_case
    shape = (8, 7, 6)
    ng = 4
    pix, rod = voronoi_case(shape, ng, seed=700, pore_frac=0.0)
    ce = _oracle_ceff(shape, ng + 1, pix, rod)
    # averages use voxel-centre strains (f90:2149-2408), not the energy-consistent
    # average, so C_eff is symmetric only to the discretisation (~4e-5 here)
    assert _rel(ce, ce.T) < 2e-4
    rot = oc.rotate_stiffness(rod, C11, C12, C44).reshape(ng, 6, 6)
    frac = np.bincount(pix, minlength=ng + 1)[1:] / pix.size
    voigt = np.einsum("g,gij->ij", frac, rot)
    reuss = np.linalg.inv(np.einsum("g,gij->ij", frac, np.linalg.inv(rot)))
    sym = 0.5 * (ce + ce.T)
    scale = np.max(np.abs(voigt))
    # cubic grains share one bulk modulus: along the hydrostatic strain Voigt is exact
    assert np.min(np.linalg.eigvalsh(voigt - sym)) > -1e-6 * scale
    assert np.min(np.linalg.eigvalsh(sym - reuss)) > -1e-6 * scale
    assert np.max(np.linalg.eigvalsh(voigt - sym)) > 1e-3 * scale      # strictly inside


# ---------------------------------------------------------------- hotspots (postprocessing.m:86-243)
def test_hotspot_plateaus_and_borders():
    shape, vm, at = hot_fixture()
    pix = np.ones(vm.size, np.int32)
    # sigma 0.05: the off-centre weights are exp(-200), so the smoothing leaves the field
    # unchanged and the plateaus stay exact
    idx, val, n = oc.hotspots(shape, pix, vm, 1, sigma=0.05, thr_mpa=250.0, npeaks=5)
    assert n == 4
    assert list(idx[:4]) == [at(2, 2, 2), at(3, 2, 2), at(7, 6, 5), at(0, 0, 0)]
    assert list(val[:4]) == [np.float32(300e6)] * 2 + [np.float32(290e6), np.float32(260e6)]
    assert idx[4] == -1
    # a constant field is one regional maximum of every voxel; ties by index
    idx, _, n = oc.hotspots(shape, pix, np.full(vm.size, 300e6), 1, sigma=0.05, thr_mpa=250.0)
    assert n == vm.size and list(idx) == [0, 1, 2]


def test_hotspot_smoothing_equals_scipy_masked_gaussian():
    """Masked normalised smoothing against scipy.ndimage.gaussian_filter (zero padding,
    radius ceil(2 sigma)), in double: the oracle's single-precision quotient agrees to
    float rounding; pores (label > nsolid) read 0."""
    from scipy.ndimage import gaussian_filter
    shape = (11, 9, 8)
    nx, ny, nz = shape
    rng = np.random.default_rng(60)
    vm = rng.uniform(50e6, 400e6, int(np.prod(shape)))
    pix = rng.integers(1, 4, vm.size).astype(np.int32)
    pix[rng.random(vm.size) < 0.1] = 4
    for sigma in (0.4, 0.9, 1.3):
        _, _, _, sm = oc.hotspots(shape, pix, vm, 3, sigma=sigma, thr_mpa=250.0, want_smooth=True)
        solid = (pix <= 3).reshape(nz, ny, nx)
        v3 = vm.reshape(nz, ny, nx)
        kw = dict(mode="constant", cval=0.0, truncate=np.ceil(2 * sigma) / sigma)
        num = gaussian_filter(np.where(solid, v3, 0.0), sigma, **kw)
        den = gaussian_filter(solid.astype(float), sigma, **kw)
        want = np.where(solid, num / np.where(den < 1e-6, 1.0, den), 0.0).ravel()
        assert np.max(np.abs(sm - want)) < 3e-7 * np.max(want)
        assert np.all(sm[pix > 3] == 0.0)


def test_hotspot_maxima_equal_maximum_filter_on_a_tie_free_field():
    from scipy.ndimage import maximum_filter
    shape = (12, 10, 9)
    nx, ny, nz = shape
    rng = np.random.default_rng(61)
    vm = rng.uniform(0.0, 500e6, int(np.prod(shape)))
    pix = np.ones(vm.size, np.int32)
    idx, val, n = oc.hotspots(shape, pix, vm, 1, sigma=0.4, thr_mpa=250.0, npeaks=8, want_smooth=True)[:3]
    sm = oc.hotspots(shape, pix, vm, 1, sigma=0.4, thr_mpa=250.0, want_smooth=True)[3]
    s3 = sm.reshape(nz, ny, nx)
    mx = maximum_filter(s3, size=3, mode="constant", cval=-np.inf)
    cand = np.flatnonzero(((s3 == mx) & (s3 / np.float32(1e6) >= np.float32(250.0))).ravel())
    assert len(np.unique(sm)) == sm.size                 # no ties: the filter test is exact
    assert n == cand.size
    order = cand[np.argsort(-sm[cand], kind="stable")]
    assert list(idx) == list(order[:8])
    assert np.array_equal(val, sm[order[:8]])


def test_top_peaks_text_format():
    from elas3dxtal.io import format_top_peaks
    shape = (4, 3, 2)
    txt = format_top_peaks(shape, [5, 17], [np.float32(301.2345e6), np.float32(250.0e6)],
                           xvec=[-1.5, -0.5, 0.5, 1.5], yvec=[0, 10, 20], zvec=[7.126, 8.0])
    # node 5 = (1, 1, 0), node 17 = (1, 1, 1)
    assert txt == ("Top 2 Local Maxima (Von Mises Stress):\n"
                   "Peak #1: 301.234 MPa at (-0.50, 10.00, 7.13)\n"
                   "Peak #2: 250.000 MPa at (-0.50, 10.00, 8.00)\n")
    with pytest.raises(ValueError):
        format_top_peaks(shape, [], [])
