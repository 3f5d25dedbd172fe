"""Writes the golden fixtures of this directory from the CPU oracle (oracle/, a
restatement of src/elas3dxtal_pcg.f90 pinned by tests/test_oracle_pins.py).

The reference ships no golden outputs, and it cannot be built in this image.  These
files freeze the oracle's outputs on three small seeded cases, so that every later test
run checks against the same numbers: the GPU tests against the files directly, and a
CPU test checks that the oracle still reproduces them.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import _oracle as oc                      # noqa: E402
from _cases import RESET_CASE, voronoi_case  # noqa: E402

C11, C12, C44 = 246.5e9, 147.3e9, 124.7e9


def case_poly():
    """10x9x8 grid, 5 Voronoi grains + 3 % single-voxel pores, exx (the main program's
    sequence, f90:326-540), with u and the full field (f90:2639-2833)."""
    shape = (10, 9, 8)
    pix, rod = voronoi_case(shape, 5, seed=900, pore_frac=0.03)
    e6 = np.array([oc.AML_F32, 0, 0, 0, 0, 0])
    r = oc.solve(shape, 6, pix, rod, C11, C12, C44, e6)
    cm = oc.cmod(6, oc.rotate_stiffness(rod, C11, C12, C44), [0.0, 0.0])
    sf, vm = oc.fullfield(shape, pix, cm, e6, r["u"])
    return dict(shape=np.array(shape), nphase=6, pix=pix.astype(np.int32), rod=rod, strain6=e6,
                phasemod=np.zeros(2), flag_m=1, lstep=r["lstep"], status=r["status"],
                reset_count=r["reset_count"], utot0=r["utot0"], gg0=r["gg0"], utot=r["utot"], gg=r["gg"],
                stress6=r["stress6"], strain_avg6=r["strain6"], history=r["history"], u=r["u"],
                stress_field=sf.ravel(), vm=vm)


def case_reset():
    """The indefinite case whose PCG resets its direction (f90:1961-1981)."""
    rc = RESET_CASE
    shape = rc["shape"]
    pix, rod = voronoi_case(shape, rc["ngrain"], rc["seed"], rc["pore_frac"])
    e6 = np.array([oc.AML_F32, 0, 0, 0, 0, 0])
    r = oc.solve(shape, rc["ngrain"] + 1, pix, rod, C11, C12, C44, e6, phasemod=rc["phasemod"])
    return dict(shape=np.array(shape), nphase=rc["ngrain"] + 1, pix=pix.astype(np.int32), rod=rod, strain6=e6,
                phasemod=np.array(rc["phasemod"]), flag_m=1, lstep=r["lstep"], status=r["status"],
                reset_count=r["reset_count"], utot0=r["utot0"], gg0=r["gg0"], utot=r["utot"], gg=r["gg"],
                stress6=r["stress6"], strain_avg6=r["strain6"], history=r["history"], u=r["u"])


def case_void():
    """flag_m = 0 (applications/sec3_1/pcg): isotropic matrix with the Voigt moduli of the
    cubic constants and a centred spherical void, ezz."""
    shape = (12, 12, 10)
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r2 = (i - nx / 2 + 0.5) ** 2 + (j - ny / 2 + 0.5) ** 2 + (k - nz / 2 + 0.5) ** 2
    pix = np.where(r2.ravel() <= 3.0 ** 2, 2, 1).astype(np.int32)
    pm = np.array([(C11 + 2 * C12) / 3.0, (C11 - C12 + 3 * C44) / 5.0, 0.0, 0.0])
    e6 = np.array([0, 0, oc.AML_F32, 0, 0, 0])
    r = oc.solve(shape, 2, pix, np.zeros((1, 3)), C11, C12, C44, e6, phasemod=pm, flag_m=0)
    return dict(shape=np.array(shape), nphase=2, pix=pix, rod=np.zeros((1, 3)), strain6=e6, phasemod=pm,
                flag_m=0, lstep=r["lstep"], status=r["status"], reset_count=r["reset_count"], utot0=r["utot0"],
                gg0=r["gg0"], utot=r["utot"], gg=r["gg"], stress6=r["stress6"], strain_avg6=r["strain6"],
                history=r["history"], u=r["u"])


# Nodes at which the 200^3 fixtures keep the displacement: every CUBE200_STRIDE-th node
# (2,998 of 8,000,000, spread over every z plane).
CUBE200_STRIDE = 2671


def case_cube200(zener):
    """The 200^3 workload at Zener ratio `zener`: 500 periodic Voronoi grains from seeded
    centres (the generator of elas3d-xtal_amd, seed 2001), uniform SO(3) orientations,
    C12 = 147.3, C44 = 124.7 GPa, C11 = C12 + 2 C44 / A, 1 % lognormal spherical pores,
    exx = float32(1e-3), the reference's PCG controls.  Kept: a hash of the generated
    microstructure (the test regenerates it), the solve's scalars and residual history,
    and the displacement at a fixed sample of nodes."""
    import hashlib

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "elas3d-xtal_amd"))
    from elas3dxtal import configs

    w = configs.build("cube200")
    c11, c12, c44 = configs.materials("cube200", zener=zener)
    shape, pix, rod = w["shape"], w["pix"], w["rod"]
    e6 = np.array([oc.AML_F32, 0, 0, 0, 0, 0])
    r = oc.solve(shape, w["nphase"], pix, rod, c11, c12, c44, e6)
    ns = int(np.prod(shape))
    idx = np.arange(0, ns, CUBE200_STRIDE)
    u = r["u"].reshape(3, ns)[:, idx]
    # The oracle's own sensitivity to rounding: the same PCG from its initial gradient
    # perturbed by one ulp (alternating signs), run for the same number of iterations.  On
    # ill-conditioned cases (A = 5) the iterates then move far above 1e-10, and any other
    # summation order (the GPU's) moves them as much.
    cm = oc.cmod(w["nphase"], oc.rotate_stiffness(rod, c11, c12, c44), [0.0, 0.0])
    dk = oc.dk(cm)
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk, b, cval, u0)
    sign = np.where(np.arange(g0.size) % 2 == 0, 1.0, -1.0)
    rp = oc.dembx(shape, pix, dk, u0, g0 * (1.0 + 2.0 ** -53 * sign), oc.sumsq(g0), tol=0.0, ldemb=r["lstep"])
    _, ut_p = oc.energy(shape, pix, dk, b, cval, rp["u"])
    st_p, _ = oc.stress(shape, pix, cm, e6, rp["u"])

    def rel(a, b_):
        return float(np.max(np.abs(np.asarray(a) - np.asarray(b_))) / np.max(np.abs(np.asarray(b_))))

    spread = np.array([rel(rp["u"].reshape(3, ns)[:, idx], u), rel(rp["history"], r["history"]),
                       abs(ut_p - r["utot"]) / abs(r["utot"]), rel(st_p, r["stress6"])])
    return dict(shape=np.array(shape), nphase=w["nphase"], zener=zener, c=np.array([c11, c12, c44]),
                pix_sha256=hashlib.sha256(np.ascontiguousarray(pix, np.int32).tobytes()).hexdigest(),
                rod_sha256=hashlib.sha256(np.ascontiguousarray(rod, np.float64).tobytes()).hexdigest(),
                strain6=e6, lstep=r["lstep"], status=r["status"], reset_count=r["reset_count"],
                utot0=r["utot0"], gg0=r["gg0"], utot=r["utot"], gg=r["gg"], stress6=r["stress6"],
                strain_avg6=r["strain6"], history=r["history"], u_index=idx, u_sample=u,
                spread_u_history_utot_stress=spread)


CASES = {"poly_10x9x8": case_poly, "reset_6x6x6": case_reset, "void_12x12x10": case_void}
CUBE200 = {f"cube200_A{a:g}".replace(".", "p"): a for a in (0.5, 1.0, 2.0, 5.0)}

if __name__ == "__main__":
    # usage: make_golden.py            the small cases (one thread: one fixed order)
    #        make_golden.py cube200    the four 200^3 cases (the oracle's reductions are
    #                                  fixed-order chunked sums, the same bits for any
    #                                  thread count; all threads, ~2-4 min per case)
    if len(sys.argv) > 1 and sys.argv[1] == "cube200":
        todo = {name: (lambda a=a: case_cube200(a)) for name, a in CUBE200.items()}
        if len(sys.argv) > 2:
            todo = {k: v for k, v in todo.items() if k in sys.argv[2:]}
    else:
        oc.lib.oc_set_threads(1)
        todo = CASES
    for name, fn in todo.items():
        d = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in d.items()})
        print(name, "lstep", int(d["lstep"]), "status", int(d["status"]), os.path.getsize(path), "bytes", flush=True)
