"""Host-side mirror of the reference's solver routines over the HIP C ABI.

`Elas3D` holds one device-resident problem (the reference's `e3d` state,
src/elas3dxtal_pcg.f90:155-181) and exposes the reference's subroutines under
their own names: ppixel / assig / femat / energy / dembx / stress /
stress_fullfield.  `solve()` runs them in the order of `program elas3d`
(f90:326-540).  Arrays use the reference's layout: vectors (3, ns) per
right-hand side == Fortran u(ns,3); node m = i + nx*(j + ny*k).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

TOL_DEFAULT = float(np.float32(1e-8))   # `real, parameter :: tol = 1.0d-8` (f90:148)
AML_DEFAULT = float(np.float32(1e-3))   # `aml = 1.0e-3` single-precision literal (f90:125)
LDEMB_DEFAULT = 10000                   # f90:141


def unit_strain(k, magnitude=AML_DEFAULT):
    """Tensor strain6 {exx, eyy, ezz, exz, eyz, exy} of load case k (0..5): three normal
    cases, then engineering shears xz, yz, xy of `magnitude` (tensor component = half)."""
    e = np.zeros(6)
    e[k] = magnitude if k < 3 else 0.5 * magnitude
    return e


class Elas3D:
    """One microstructure on one GPU; nrhs load cases solved together.  With
    batch_nphase (a list of 2..6 phase counts) the handle is instead a batch of that many
    separate problems on one grid (e3x_create_batch): right-hand side r is member r, set
    by set_member."""

    def __init__(self, shape, nphase=None, nrhs=1, device=0, batch_nphase=None):
        lib = L.load()
        self.shape = tuple(int(v) for v in shape)
        self.nx, self.ny, self.nz = self.shape
        self.ns = self.nx * self.ny * self.nz
        h = C.c_void_p()
        if batch_nphase is not None:
            bn = np.ascontiguousarray(batch_nphase, dtype=np.int32)
            self.nphase = int(bn.sum())
            self.nrhs = int(bn.size)
            self.batch_nphase = bn.copy()
            L.check(lib.e3x_create_batch(self.nx, self.ny, self.nz, self.nrhs, L.iptr(bn), device, C.byref(h)))
        else:
            self.nphase = int(nphase)
            self.nrhs = int(nrhs)
            self.batch_nphase = None
            L.check(lib.e3x_create(self.nx, self.ny, self.nz, self.nphase, self.nrhs, device, C.byref(h)))
        self._h = h
        self._lib = lib
        self.strain6 = np.zeros((self.nrhs, 6))
        self.C = np.zeros(self.nrhs)
        self.stop_rule = L.STOP_REFERENCE
        self.pcg_variant = L.PCG_REFERENCE
        self._pix_max = 0

    # -- lifetime -------------------------------------------------------------------
    def close(self):
        if self._h is not None and self._h.value:
            self._lib.e3x_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def memory_bytes(self):
        b = C.c_int64()
        L.check(self._lib.e3x_memory_bytes(self._h, C.byref(b)))
        return b.value

    # -- reference routines ---------------------------------------------------------
    def ppixel(self, pix, orientation=None):
        """Data part of ppixel_hdf5 (f90:2879-2991): pix labels 1..nphase, Rodrigues
        orientation[3*(nphase-1)]."""
        pix = np.ascontiguousarray(pix, dtype=np.int32).ravel()
        if pix.size != self.ns:
            raise L.E3xError(L.E3X_EINVAL, f"pix size {pix.size} does not match ns {self.ns}")
        ori = None
        if orientation is not None:
            ori = np.ascontiguousarray(orientation, dtype=np.float64).ravel()
            if ori.size != 3 * (self.nphase - 1):
                raise L.E3xError(L.E3X_EINVAL, f"orientation size {ori.size} does not match "
                                               f"3*(nphase-1) = {3 * (self.nphase - 1)}")
        L.check(self._lib.e3x_set_microstructure(self._h, L.i32ptr(pix), L.dptr(ori)))
        self._pix_max = int(pix.max())

    def set_member(self, member, pix, orientation, c11, c12, c44, strain6, phasemod=(0.0, 0.0), flag_m=1):
        """Member `member` of a batch: microstructure, materials and load (e3x_set_member)."""
        pix = np.ascontiguousarray(pix, dtype=np.int32).ravel()
        if pix.size != self.ns:
            raise L.E3xError(L.E3X_EINVAL, f"pix size {pix.size} does not match ns {self.ns}")
        ori = None if orientation is None else np.ascontiguousarray(orientation, dtype=np.float64).ravel()
        pm = np.zeros(4)
        pmv = np.asarray(phasemod, float).ravel()
        pm[:pmv.size] = pmv
        e = np.ascontiguousarray(np.asarray(strain6, float).ravel()[:6])
        L.check(self._lib.e3x_set_member(self._h, member, L.i32ptr(pix), L.dptr(ori), c11, c12, c44, L.dptr(pm),
                                         flag_m, L.dptr(e)))
        self.strain6[member] = e
        self._pix_max = max(self._pix_max, int(pix.max()))

    def assig(self):
        prob = np.empty(self.nphase)
        L.check(self._lib.e3x_assig(self._h, L.dptr(prob)))
        return prob

    def femat(self, c11, c12, c44, strain6, phasemod=(0.0, 0.0), flag_m=1):
        """femat_OpenMP (f90:555-1104); strain6 is (nrhs, 6) {exx,eyy,ezz,exz,eyz,exy}."""
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(strain6, float), (self.nrhs, 6)))
        pm = np.zeros(4)
        pmv = np.asarray(phasemod, float).ravel()
        pm[:pmv.size] = pmv
        cout = np.zeros(self.nrhs)
        L.check(self._lib.e3x_femat(self._h, c11, c12, c44, L.dptr(pm), flag_m, L.dptr(e), L.dptr(cout)))
        self.strain6 = e.copy()
        self.C = cout
        return cout

    def affine_guess(self):
        L.check(self._lib.e3x_affine_guess(self._h))

    def energy(self):
        """energy_OpenMP (f90:1107-1353) -> (utot[nrhs], gg[nrhs] = sum(gb^2))."""
        ut, gg = np.zeros(self.nrhs), np.zeros(self.nrhs)
        L.check(self._lib.e3x_energy(self._h, L.dptr(ut), L.dptr(gg)))
        return ut, gg

    def set_stop_rule(self, rule):
        """Stop test of the following PCG runs (e3x_set_stop_rule): L.STOP_REFERENCE, the
        reference's relative residual rel_res < tol (f90:1996-2004, the default), or
        L.STOP_NIST_ABS, NIST ELAS3D's absolute test sum(g^2) < 1e-8 * ns on the
        unpreconditioned gradient (tol is then unused)."""
        L.check(self._lib.e3x_set_stop_rule(self._h, rule))
        self.stop_rule = rule

    def set_pcg_variant(self, variant):
        """PCG loop of the following runs (e3x_set_pcg_variant): L.PCG_REFERENCE, the loop
        of src/elas3dxtal_pcg.f90 (default); L.PCG_SWEEP, the anisotropy-sweep programs'
        loop (no divergence reset, stop test before the direction update, A_2_0:1518);
        L.PCG_PLAIN_CG, the plain-CG programs' loop (no preconditioner either, cg:915)."""
        L.check(self._lib.e3x_set_pcg_variant(self._h, variant))
        self.pcg_variant = variant

    def dembx(self, gg0, tol=TOL_DEFAULT, ldemb=LDEMB_DEFAULT, kkk=1, precision=L.PREC_FP64,
              replace_every=50, want_history=True):
        """dembx_OpenMP (f90:1356-2043).  Returns a dict of per-rhs arrays."""
        gg0 = np.ascontiguousarray(np.broadcast_to(np.asarray(gg0, float), (self.nrhs,)))
        lstep = np.zeros(self.nrhs, np.int32)
        gg = np.zeros(self.nrhs)
        ex = np.zeros(self.nrhs, np.int32)
        rc = np.zeros(self.nrhs, np.int32)
        hist = np.zeros(self.nrhs * max(ldemb, 1)) if want_history else None
        L.check(self._lib.e3x_dembx_ex(self._h, precision, replace_every if precision else 0,
                                       L.dptr(gg0), C.c_float(tol), ldemb, kkk, L.iptr(lstep),
                                       L.dptr(gg), L.dptr(hist), L.iptr(ex), L.iptr(rc)))
        pre, pcg = C.c_double(), C.c_double()
        L.check(self._lib.e3x_last_timing(self._h, C.byref(pre), C.byref(pcg)))
        out = dict(lstep=lstep, gg=gg, exit=ex, reset_count=rc, precond_ms=pre.value, pcg_ms=pcg.value)
        if want_history:
            hh = hist.reshape(self.nrhs, max(ldemb, 1))
            out["history"] = [hh[r, :lstep[r]].copy() for r in range(self.nrhs)]
        return out

    def stress(self):
        """stress_OpenMP (f90:2149-2408) -> (stress6, strain6), each (nrhs, 6)."""
        s, e = np.zeros(6 * self.nrhs), np.zeros(6 * self.nrhs)
        L.check(self._lib.e3x_stress(self._h, L.dptr(s), L.dptr(e)))
        return s.reshape(self.nrhs, 6), e.reshape(self.nrhs, 6)

    def effective_stiffness(self, c11, c12, c44, magnitude=AML_DEFAULT, phasemod=(0.0, 0.0), flag_m=1,
                            precision=L.PREC_FP64, replace_every=50, tol=TOL_DEFAULT, ldemb=LDEMB_DEFAULT):
        """6x6 effective stiffness from the six unit load cases solved as six right-hand
        sides (e3x_effective_stiffness; the handle needs nrhs = 6).  Returns (C_eff,
        lstep[6], exit[6]); Voigt order xx yy zz xz yz xy, engineering shear strains."""
        pm = np.zeros(4)
        pmv = np.asarray(phasemod, float).ravel()
        pm[:pmv.size] = pmv
        ce = np.zeros(36)
        lstep = np.zeros(6, np.int32)
        ex = np.zeros(6, np.int32)
        L.check(self._lib.e3x_effective_stiffness(self._h, c11, c12, c44, L.dptr(pm), flag_m, magnitude, precision,
                                                  replace_every if precision else 0, C.c_float(tol), ldemb,
                                                  L.dptr(ce), L.iptr(lstep), L.iptr(ex)))
        return ce.reshape(6, 6), lstep, ex

    def stress_fullfield(self, rhs=0):
        """stress_fullfield_OpenMP (f90:2639-2833) -> stress (6, ns), vm (ns)."""
        sf, vm = np.zeros(6 * self.ns), np.zeros(self.ns)
        L.check(self._lib.e3x_stress_fullfield(self._h, rhs, L.dptr(sf), L.dptr(vm)))
        return sf.reshape(6, self.ns), vm

    def hotspots(self, rhs=0, vm=None, nsolid=None, sigma=0.4, threshold_mpa=250.0, npeaks=3):
        """Hotspots of the von Mises field as the reference's post-processing finds them
        for top_peaks.txt (applications/sec3_2/A_2_0/elas3dxtal_postprocessing.m:86-243),
        on the device (e3x_hotspots).  vm: a field in node order, or None for this
        handle's own (right-hand side rhs).  nsolid None: the largest phase label, as the
        scripts take it (.m:67).  Returns (index[k] 0-based node numbers, value[k] smoothed
        von Mises in Pa, found)."""
        if nsolid is None:
            nsolid = int(self._pix_max)
        vmv = None
        if vm is not None:
            vmv = np.ascontiguousarray(vm, dtype=np.float64).ravel()
            if vmv.size != self.ns:
                raise L.E3xError(L.E3X_EINVAL, "vm size must be ns")
        idx = np.zeros(max(npeaks, 1), np.int64)
        val = np.zeros(max(npeaks, 1))
        found = C.c_int64()
        L.check(self._lib.e3x_hotspots(self._h, rhs, L.dptr(vmv), nsolid, sigma, threshold_mpa, npeaks,
                                       idx.ctypes.data_as(C.POINTER(C.c_int64)), L.dptr(val), C.byref(found)))
        n = min(npeaks, found.value)
        return idx[:n].copy(), val[:n].copy(), found.value

    # -- fields ---------------------------------------------------------------------
    def get(self, which, rhs=0):
        out = np.zeros(3 * self.ns)
        L.check(self._lib.e3x_get_field(self._h, which, rhs, L.dptr(out)))
        return out

    def set(self, which, values, rhs=0):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if v.size != 3 * self.ns:
            raise L.E3xError(L.E3X_EINVAL, "field size must be 3*ns")
        L.check(self._lib.e3x_set_field(self._h, which, rhs, L.dptr(v)))

    def time_apply(self, precision=L.PREC_FP64, reps=20):
        """Average GPU time (ms) of the PCG's stencil-apply kernel, HIP events on the
        solver's stream (e3x_time_apply)."""
        ms = C.c_double()
        L.check(self._lib.e3x_time_apply(self._h, precision, reps, C.byref(ms)))
        return ms.value

    def set_tuning(self, precision=L.PREC_FP64, waves=0):
        """Launch variant of the PCG apply (e3x_set_tuning): 0 measures and keeps the
        fastest at the next PCG; else FP64 5/6/8, FP32 6/8 waves per SIMD."""
        L.check(self._lib.e3x_set_tuning(self._h, precision, waves))

    def tuning(self, precision=L.PREC_FP64):
        """-> (waves per SIMD in use, measured mean apply ms)."""
        w, ms = C.c_int(), C.c_double()
        L.check(self._lib.e3x_get_tuning(self._h, precision, C.byref(w), C.byref(ms)))
        return w.value, ms.value

    def set_apply_timing(self, on=True):
        """Time the stencil applies of the following PCG runs in the loop itself (HIP
        events on the solver's stream; e3x_set_apply_timing)."""
        L.check(self._lib.e3x_set_apply_timing(self._h, 1 if on else 0))

    def last_apply_timing(self):
        """-> (launches timed, mean ms) of the last PCG's fused applies."""
        n, ms = C.c_int(), C.c_double()
        L.check(self._lib.e3x_last_apply_timing(self._h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def tables(self):
        cm = np.zeros(36 * self.nphase)
        dk = np.zeros(576 * self.nphase)
        L.check(self._lib.e3x_get_tables(self._h, L.dptr(cm), L.dptr(dk)))
        return cm.reshape(self.nphase, 6, 6), dk.reshape(self.nphase, 8, 3, 8, 3)


@dataclass
class SolveResult:
    """What program elas3d logs and writes (f90:449-519)."""
    utot0: np.ndarray
    gg0: np.ndarray
    utot: np.ndarray
    gg: np.ndarray
    rel_res: np.ndarray
    lstep: np.ndarray
    exit: np.ndarray
    reset_count: np.ndarray
    C: np.ndarray
    stress6: np.ndarray
    strain6: np.ndarray
    history: list = field(default_factory=list)
    prob: np.ndarray = None
    timing: dict = field(default_factory=dict)


def solve(pix, orientation, shape, nphase, c11, c12, c44, strain6, phasemod=(0.0, 0.0),
          flag_m=1, tol=TOL_DEFAULT, ldemb=LDEMB_DEFAULT, precision=L.PREC_FP64,
          replace_every=50, device=0, keep=False, stop_rule=L.STOP_REFERENCE, variant=L.PCG_REFERENCE):
    """The main program's sequence (f90:326-540) for one microstructure.

    strain6 may be (6,) or (nrhs, 6); with nrhs > 1 the load cases are solved in one
    batched PCG.  stop_rule selects the PCG's stop test (Elas3D.set_stop_rule), variant
    its loop (Elas3D.set_pcg_variant).
    Returns SolveResult (and the live Elas3D if keep=True).
    """
    e = np.atleast_2d(np.asarray(strain6, float))
    s = Elas3D(shape, nphase, nrhs=e.shape[0], device=device)
    try:
        if stop_rule != L.STOP_REFERENCE:
            s.set_stop_rule(stop_rule)
        if variant != L.PCG_REFERENCE:
            s.set_pcg_variant(variant)
        s.ppixel(pix, orientation)
        prob = s.assig()
        cval = s.femat(c11, c12, c44, e, phasemod=phasemod, flag_m=flag_m)
        s.affine_guess()
        ut0, gg0 = s.energy()
        d = s.dembx(gg0, tol=tol, ldemb=ldemb, kkk=1, precision=precision, replace_every=replace_every)
        ut1, _ = s.energy()
        st, sn = s.stress()
        res = SolveResult(utot0=ut0, gg0=gg0, utot=ut1, gg=d["gg"], rel_res=np.sqrt(d["gg"] / gg0),
                          lstep=d["lstep"], exit=d["exit"], reset_count=d["reset_count"], C=cval,
                          stress6=st, strain6=sn, history=d["history"], prob=prob,
                          timing=dict(precond_ms=d["precond_ms"], pcg_ms=d["pcg_ms"]))
    except Exception:
        s.close()
        raise
    if keep:
        return res, s
    s.close()
    return res


def solve_batch(shape, members, tol=TOL_DEFAULT, ldemb=LDEMB_DEFAULT, precision=L.PREC_FP64, replace_every=50,
                device=0, keep=False):
    """Several problems on one grid solved as one batch (e3x_create_batch): the main
    program's sequence (f90:326-540) for every member at once.  members: list of dicts
    with pix, orientation, nphase, c11, c12, c44, strain6 and optionally phasemod, flag_m.
    Returns SolveResult with one row per member (and the live Elas3D if keep=True)."""
    s = Elas3D(shape, batch_nphase=[int(m["nphase"]) for m in members], device=device)
    try:
        for r, m in enumerate(members):
            s.set_member(r, m["pix"], m.get("orientation"), m["c11"], m["c12"], m["c44"], m["strain6"],
                         phasemod=m.get("phasemod", (0.0, 0.0)), flag_m=m.get("flag_m", 1))
        s.affine_guess()
        ut0, gg0 = s.energy()
        d = s.dembx(gg0, tol=tol, ldemb=ldemb, kkk=1, precision=precision, replace_every=replace_every)
        ut1, _ = s.energy()
        st, sn = s.stress()
        res = SolveResult(utot0=ut0, gg0=gg0, utot=ut1, gg=d["gg"], rel_res=np.sqrt(d["gg"] / gg0),
                          lstep=d["lstep"], exit=d["exit"], reset_count=d["reset_count"], C=np.zeros(s.nrhs),
                          stress6=st, strain6=sn, history=d["history"],
                          timing=dict(precond_ms=d["precond_ms"], pcg_ms=d["pcg_ms"]))
    except Exception:
        s.close()
        raise
    if keep:
        return res, s
    s.close()
    return res


def effective_stiffness(pix, orientation, shape, nphase, c11, c12, c44, magnitude=AML_DEFAULT,
                        phasemod=(0.0, 0.0), flag_m=1, precision=L.PREC_FP64, replace_every=50,
                        tol=TOL_DEFAULT, ldemb=LDEMB_DEFAULT, device=0):
    """Effective 6x6 stiffness of one microstructure: the six unit macro strains as six
    right-hand sides of one batched PCG.  Returns (C_eff, lstep[6], exit[6])."""
    with Elas3D(shape, nphase, nrhs=6, device=device) as s:
        s.ppixel(pix, orientation)
        return s.effective_stiffness(c11, c12, c44, magnitude=magnitude, phasemod=phasemod, flag_m=flag_m,
                                     precision=precision, replace_every=replace_every, tol=tol, ldemb=ldemb)
