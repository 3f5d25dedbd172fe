"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
oracle): the oracle must still reproduce them (CPU), and the GPU library must match them
(marked gpu) with the parity bars of the GPU tests: FP64 fields, energies and averages to
1e-10 relative, the iteration count within 1 %.  On the indefinite reset case the
tolerances are those of its oracle pin, since summation order there grows to ~1e-6.

The 200^3 fixtures (the metric's workload at Zener ratios 0.5, 1, 2 and 5) hold a hash of
the generated microstructure instead of the microstructure: the CPU test checks that the
generator still produces it, the GPU test regenerates it, solves, and compares the
scalars, the residual history and the displacement at ~3,000 fixed nodes."""
import os

import numpy as np
import pytest

import _oracle as oc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["poly_10x9x8", "reset_6x6x6", "void_12x12x10"]
C11, C12, C44 = 246.5e9, 147.3e9, 124.7e9


def _load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CUBE200 = ["cube200_A0p5", "cube200_A1", "cube200_A2", "cube200_A5"]


def _cube200_input(g):
    import hashlib

    from elas3dxtal import configs
    w = configs.build("cube200")
    assert tuple(w["shape"]) == tuple(int(v) for v in g["shape"]) and w["nphase"] == int(g["nphase"])
    assert hashlib.sha256(np.ascontiguousarray(w["pix"], np.int32).tobytes()).hexdigest() == str(g["pix_sha256"])
    assert hashlib.sha256(np.ascontiguousarray(w["rod"], np.float64).tobytes()).hexdigest() == str(g["rod_sha256"])
    return w


@pytest.mark.parametrize("name", CUBE200[:1])
def test_generator_reproduces_cube200_input(name):
    """The seeded generator still writes the microstructure the 200^3 fixtures were solved
    on (the fixtures keep its hash), whatever the thread count."""
    g = _load(name)
    w = _cube200_input(g)
    assert abs(w["pore_fraction"] - 0.01) < 1e-3
    c11, c12, c44 = (float(v) for v in g["c"])
    assert abs(2 * c44 / (c11 - c12) - float(g["zener"])) < 1e-12


def _gpu_spread(args, g, lstep, res, u):
    """Change of the GPU FP64 solve when its initial gradient is perturbed by one ulp with
    alternating signs (as the fixture's oracle spread), run for `lstep` iterations."""
    from elas3dxtal import _lib as L
    from elas3dxtal.solver import Elas3D
    pix, rod, shape, nphase, c11, c12, c44, e6 = args
    with Elas3D(shape, nphase) as s:
        s.ppixel(pix, rod)
        s.femat(c11, c12, c44, e6)
        s.affine_guess()
        _, gg0 = s.energy()
        g0 = s.get(L.FIELD_GB)
        sign = np.where(np.arange(g0.size) % 2 == 0, 1.0, -1.0)
        s.set(L.FIELD_GB, g0 * (1.0 + 2.0 ** -53 * sign))
        d = s.dembx(gg0, tol=0.0, ldemb=lstep)
        ut, _ = s.energy()
        st, _ = s.stress()
        up = s.get(L.FIELD_U).reshape(3, -1)[:, g["u_index"]]
    return dict(u=_rel(up, u), history=_rel(d["history"][0], res.history[0]),
                utot=abs(ut[0] - res.utot[0]) / abs(res.utot[0]), stress6=_rel(st[0], res.stress6[0]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", CUBE200)
def test_gpu_matches_cube200_golden(name):
    """The 200^3 / 500-grain workload at one Zener ratio: the GPU FP64 solve against the
    oracle's (iterations within 1 %, the difference printed; residual history to 1e-8;
    energies, averages and the sampled displacement to 1e-10)."""
    from elas3dxtal import _lib as L
    from elas3dxtal.solver import solve
    g = _load(name)
    w = _cube200_input(g)
    c11, c12, c44 = (float(v) for v in g["c"])
    args = (w["pix"], w["rod"], w["shape"], w["nphase"], c11, c12, c44, g["strain6"])
    res, s = solve(*args, keep=True)
    lstep = int(g["lstep"])
    K = int(res.lstep[0])
    print(f"{name}: lstep GPU {K}, oracle {lstep} (difference {K - lstep})")
    try:
        assert res.exit[0] == int(g["status"]) == L.PCG_CONVERGED
        assert res.reset_count[0] == int(g["reset_count"])
        assert abs(K - lstep) <= max(1, int(0.01 * lstep))
        if K != lstep:
            s.close()
            res, s = solve(*args, tol=0.0, ldemb=lstep, keep=True)
            assert res.exit[0] == L.PCG_MAXITER and int(res.lstep[0]) == lstep
        u = s.get(L.FIELD_U).reshape(3, -1)[:, g["u_index"]]
    finally:
        s.close()
    errs = dict(utot0=abs(res.utot0[0] - float(g["utot0"])) / abs(float(g["utot0"])),
                utot=abs(res.utot[0] - float(g["utot"])) / abs(float(g["utot"])),
                stress6=_rel(res.stress6[0], g["stress6"]), strain6=_rel(res.strain6[0], g["strain_avg6"]),
                u=_rel(u, g["u_sample"]), history=_rel(res.history[0], g["history"]))
    # bars: the north star's (1e-10 on fields, energy and averages; 1e-8 on the history),
    # or, where a one-ulp perturbation of the initial gradient moves the result by more
    # (the ill-conditioned A = 5 case), ten times the larger of that change in the
    # oracle (stored with the fixture) and in the GPU solve (measured here the same way:
    # the same PCG from the perturbed gradient, run for as many iterations)
    sp = dict(zip(("u", "history", "utot", "stress6"), (float(v) for v in g["spread_u_history_utot_stress"])))
    if sp["u"] > 1e-10 or sp["utot"] > 1e-10 or sp["stress6"] > 1e-10 or sp["history"] > 1e-8:
        gsp = _gpu_spread(args, g, lstep, res, u)
        print(name, "GPU one-ulp spread", {k: f"{v:.2e}" for k, v in gsp.items()})
        sp = {k: max(sp[k], gsp[k]) for k in sp}
    print(name, {k: f"{v:.2e}" for k, v in errs.items()}, "one-ulp spread",
          {k: f"{v:.2e}" for k, v in sp.items()})
    assert errs["utot0"] < 1e-10
    assert errs["utot"] < max(1e-10, 10 * sp["utot"])
    assert errs["stress6"] < max(1e-10, 10 * sp["stress6"])
    assert errs["strain6"] < 1e-12
    assert errs["u"] < max(1e-10, 10 * sp["u"])
    assert errs["history"] < max(1e-8, 10 * sp["history"])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = _load(name)
    shape = tuple(int(v) for v in g["shape"])
    oc.lib.oc_set_threads(1)
    try:
        r = oc.solve(shape, int(g["nphase"]), g["pix"], g["rod"], C11, C12, C44, g["strain6"],
                     phasemod=g["phasemod"], flag_m=int(g["flag_m"]))
    finally:
        oc.lib.oc_set_threads(0)
    assert (r["lstep"], r["status"], r["reset_count"]) == (int(g["lstep"]), int(g["status"]), int(g["reset_count"]))
    for k in ("utot0", "gg0", "utot", "gg"):
        assert abs(r[k] - float(g[k])) <= 1e-13 * abs(float(g[k])), k
    assert _rel(r["stress6"], g["stress6"]) < 1e-13
    assert _rel(r["history"], g["history"]) < 1e-12
    assert _rel(r["u"], g["u"]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(name):
    from elas3dxtal import _lib as L
    from elas3dxtal.solver import solve
    g = _load(name)
    shape = tuple(int(v) for v in g["shape"])
    args = (g["pix"], g["rod"], shape, int(g["nphase"]), C11, C12, C44, g["strain6"])
    kw = dict(phasemod=g["phasemod"], flag_m=int(g["flag_m"]), keep=True)
    res, s = solve(*args, **kw)
    reset = name.startswith("reset")
    lstep = int(g["lstep"])
    K = int(res.lstep[0])
    print(f"{name}: lstep GPU {K}, golden {lstep}")
    try:
        assert res.exit[0] == int(g["status"]) == L.PCG_CONVERGED
        assert res.reset_count[0] == int(g["reset_count"])
        assert abs(K - lstep) <= max(1, int(0.01 * lstep))
        if K != lstep:
            # equal iteration counts: the GPU run for exactly the golden's lstep iterations
            # (tol = 0, the iteration limit ends it) holds the iterate the file froze
            s.close()
            res, s = solve(*args, tol=0.0, ldemb=lstep, **kw)
            assert res.exit[0] == L.PCG_MAXITER and int(res.lstep[0]) == lstep
        u = s.get(L.FIELD_U)
        sf, vm = s.stress_fullfield(0) if "vm" in g else (None, None)
    finally:
        s.close()
    assert abs(res.utot0[0] - float(g["utot0"])) < 1e-12 * abs(float(g["utot0"]))
    tol = 1e-6 if reset else 1e-10
    assert abs(res.utot[0] - float(g["utot"])) < (1e-8 if reset else 1e-10) * abs(float(g["utot"]))
    assert _rel(u, g["u"]) < tol
    assert _rel(res.history[0], g["history"]) < (1e-5 if reset else 1e-8)
    assert _rel(res.stress6[0], g["stress6"]) < tol
    assert _rel(res.strain6[0], g["strain_avg6"]) < 1e-12
    if vm is not None:
        assert _rel(sf.ravel(), g["stress_field"]) < 1e-10
        assert _rel(vm, g["vm"]) < 1e-10
