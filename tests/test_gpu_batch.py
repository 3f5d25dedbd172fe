"""Batches of separate problems on one grid (e3x_create_batch, the serving loop's
co-batching): every member's results are bit for bit those of its problem solved alone,
and, in FP64, the oracle's."""
import numpy as np
import pytest

import _oracle as oc
from _cases import voronoi_case
from elas3dxtal import _lib as L
from elas3dxtal.solver import Elas3D, solve, solve_batch

pytestmark = pytest.mark.gpu

C12, C44 = 147.3e9, 124.7e9


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _members(shape, n, seed):
    """n problems on one grid: own grains, porosity, Zener ratio and one unit load each."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        ng = int(rng.integers(3, 12))
        pix, rod = voronoi_case(shape, ng, seed=seed * 10 + k, pore_frac=float(rng.uniform(0.0, 0.05)))
        a = float(rng.uniform(0.5, 5.0))
        e6 = np.zeros(6)
        j = int(rng.integers(0, 6))
        e6[j] = oc.AML_F32 if j < 3 else 0.5 * oc.AML_F32
        out.append(dict(pix=pix, orientation=rod, nphase=ng + 1, c11=C12 + 2 * C44 / a, c12=C12, c44=C44, strain6=e6))
    return out


@pytest.mark.parametrize("shape,n,precision", [((16, 12, 8), 3, L.PREC_FP64), ((16, 12, 8), 6, L.PREC_MIXED),
                                               ((24, 24, 24), 4, L.PREC_FP64), ((9, 7, 5), 2, L.PREC_MIXED)])
def test_batch_members_equal_solves_alone(shape, n, precision):
    mem = _members(shape, n, seed=70 + n)
    batch, s = solve_batch(shape, mem, precision=precision, replace_every=7, keep=True)
    try:
        u_b = [s.get(L.FIELD_U, r) for r in range(n)]
        vm_b = [s.stress_fullfield(r)[1] for r in range(n)]
    finally:
        s.close()
    for r, m in enumerate(mem):
        one, s1 = solve(m["pix"], m["orientation"], shape, m["nphase"], m["c11"], m["c12"], m["c44"], m["strain6"],
                        precision=precision, replace_every=7, keep=True)
        try:
            u1 = s1.get(L.FIELD_U)
            vm1 = s1.stress_fullfield(0)[1]
        finally:
            s1.close()
        assert batch.exit[r] == one.exit[0] == L.PCG_CONVERGED
        assert batch.lstep[r] == one.lstep[0]
        assert batch.utot0[r] == one.utot0[0] and batch.utot[r] == one.utot[0]
        assert np.array_equal(batch.stress6[r], one.stress6[0])
        assert np.array_equal(batch.history[r], one.history[0])
        assert np.array_equal(u_b[r], u1)
        assert np.array_equal(vm_b[r], vm1)


def test_batch_member_matches_oracle():
    """A member of an FP64 batch against the oracle (the same bar as a single solve)."""
    shape = (20, 18, 16)
    mem = _members(shape, 3, seed=91)
    batch, s = solve_batch(shape, mem, keep=True)
    try:
        u = [s.get(L.FIELD_U, r) for r in range(3)]
    finally:
        s.close()
    for r, m in enumerate(mem):
        ro = oc.solve(shape, m["nphase"], m["pix"], m["orientation"], m["c11"], m["c12"], m["c44"], m["strain6"])
        K = int(batch.lstep[r])
        assert abs(K - ro["lstep"]) <= max(1, int(0.01 * ro["lstep"]))
        if K != ro["lstep"]:
            ro = oc.solve(shape, m["nphase"], m["pix"], m["orientation"], m["c11"], m["c12"], m["c44"], m["strain6"],
                          tol=0.0, ldemb=K)
        assert abs(batch.utot[r] - ro["utot"]) < 1e-10 * abs(ro["utot"])
        assert _rel(batch.stress6[r], ro["stress6"]) < 1e-10
        assert _rel(u[r], ro["u"]) < 1e-10


def test_batch_calls_and_labels():
    shape = (6, 6, 6)
    mem = _members(shape, 2, seed=93)
    with Elas3D(shape, batch_nphase=[mem[0]["nphase"], mem[1]["nphase"]]) as s:
        with pytest.raises(L.E3xError) as e:       # the single-problem calls are not for batches
            s.ppixel(mem[0]["pix"], None)
        assert e.value.code == L.E3X_ESTATE
        with pytest.raises(L.E3xError) as e:       # not every member set yet
            s.energy()
        assert e.value.code == L.E3X_ESTATE
        bad = mem[1]["pix"].copy()
        bad[3] = mem[1]["nphase"] + 1              # beyond this member's phases (f90:2986-2989)
        with pytest.raises(L.E3xError) as e:
            s.set_member(1, bad, mem[1]["orientation"], mem[1]["c11"], C12, C44, mem[1]["strain6"])
        assert e.value.code == L.E3X_EPHASE
    with pytest.raises(L.E3xError) as e:
        Elas3D(shape, batch_nphase=[3])            # a batch has 2..6 members
    assert e.value.code == L.E3X_EINVAL


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_poisoned_allocations_give_the_same_bits(monkeypatch, precision):
    """Every device allocation filled with 0xFF bytes (NaN in FP32 and FP64, the all-ones
    phase id; E3X_POISON_ALLOC=1) changes no bit of a batch or of a single solve: no kernel
    reads a buffer before something wrote it."""
    shape = (16, 12, 8)
    mem = _members(shape, 4, seed=95)
    m = mem[0]

    def run():
        batch, s = solve_batch(shape, mem, precision=precision, replace_every=7, keep=True)
        try:
            ub = [s.get(L.FIELD_U, r) for r in range(len(mem))]
        finally:
            s.close()
        one, s1 = solve(m["pix"], m["orientation"], shape, m["nphase"], m["c11"], m["c12"], m["c44"], m["strain6"],
                        precision=precision, replace_every=7, keep=True)
        try:
            u1 = s1.get(L.FIELD_U)
        finally:
            s1.close()
        return batch, ub, one, u1

    clean = run()
    monkeypatch.setenv("E3X_POISON_ALLOC", "1")
    poisoned = run()
    monkeypatch.delenv("E3X_POISON_ALLOC")
    for res in (clean, poisoned):
        assert all(int(e) == L.PCG_CONVERGED for e in res[0].exit) and int(res[2].exit[0]) == L.PCG_CONVERGED
    (b0, u0, o0, v0), (b1, u1, o1, v1) = clean, poisoned
    assert np.array_equal(b0.lstep, b1.lstep) and np.array_equal(b0.utot, b1.utot)
    assert np.array_equal(b0.stress6, b1.stress6)
    assert all(np.array_equal(a, b) for a, b in zip(b0.history, b1.history))  # one array per member
    assert all(np.array_equal(a, b) for a, b in zip(u0, u1))
    assert o0.lstep[0] == o1.lstep[0] and o0.utot[0] == o1.utot[0] and np.array_equal(v0, v1)
    assert np.array_equal(o0.stress6, o1.stress6)
