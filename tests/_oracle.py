"""ctypes binding of the CPU oracle (oracle/libe3x_oracle.so) for the tests.

The oracle is the parity checker: tests call it to produce expected values; the
product under test is the HIP library reached through include/e3x.h.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libe3x_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

OC_RUNNING, OC_CONVERGED, OC_HAH_BREAK, OC_MAXITER = 0, 1, 2, 3
TOL_F32 = float(np.float32(1e-8))      # `real, parameter :: tol = 1.0d-8` (f90:148)
AML_F32 = float(np.float32(1e-3))      # `aml = 1.0e-3` single-precision literal (f90:125)


class OcResult(C.Structure):
    _fields_ = [("utot0", C.c_double), ("gg0", C.c_double), ("utot", C.c_double),
                ("gg", C.c_double), ("rel_res", C.c_double), ("lstep", C.c_int),
                ("status", C.c_int), ("reset_count", C.c_int), ("C", C.c_double),
                ("stress6", C.c_double * 6), ("strain6", C.c_double * 6)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle library missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    i, d, f, v = C.c_int, C.c_double, C.c_float, None
    sig = {
        "oc_build_ib": (v, [i, i, i, _ip]),
        "oc_assig": (v, [C.c_int64, i, _ip, _dp]),
        "oc_rodrigues_matrix": (v, [_dp, _dp]),
        "oc_rotate_stiffness": (v, [i, _dp, d, d, d, _dp]),
        "oc_cmod": (v, [i, i, _dp, _dp, _dp]),
        "oc_dk": (v, [i, _dp, _dp]),
        "oc_femat_rhs": (v, [i, i, i, _ip, _dp, _dp, _dp, C.POINTER(C.c_double)]),
        "oc_affine": (v, [i, i, i, _dp, _dp]),
        "oc_apply": (v, [i, i, i, _ip, _dp, _dp, _dp]),
        "oc_energy": (v, [i, i, i, _ip, _dp, _dp, d, _dp, _dp, C.POINTER(C.c_double)]),
        "oc_sumsq": (d, [C.c_int64, _dp]),
        "oc_invert3x3": (v, [_dp, _dp]),
        "oc_precond": (v, [i, i, i, _ip, _dp, _dp]),
        "oc_dembx": (i, [i, i, i, _ip, _dp, _dp, _dp, _dp, d, f, i, C.POINTER(C.c_int),
                         C.POINTER(C.c_double), _dp, C.POINTER(C.c_int)]),
        "oc_dembx_variant": (i, [i, i, i, _ip, _dp, _dp, _dp, _dp, d, f, i, C.POINTER(C.c_int),
                                 C.POINTER(C.c_double), _dp, C.POINTER(C.c_int), i]),
        "oc_hotspots": (C.c_int64, [i, i, i, _ip, _dp, i, d, d, i, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "oc_stress": (v, [i, i, i, _ip, _dp, _dp, _dp, _dp, _dp]),
        "oc_fullfield": (v, [i, i, i, _ip, _dp, _dp, _dp, _dp, _dp]),
        "oc_solve": (i, [i, i, i, i, i, _ip, _dp, d, d, d, _dp, _dp, f, i, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.POINTER(OcResult)]),
        "oc_set_threads": (v, [i]),
        "oc_get_threads": (i, []),
        "oc_last_loop_seconds": (d, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def build_ib(nx, ny, nz):
    ib = np.empty(27 * nx * ny * nz, np.int32)
    lib.oc_build_ib(nx, ny, nz, ib)
    return ib.reshape(27, -1)


def assig(pix, nphase):
    prob = np.empty(nphase)
    lib.oc_assig(pix.size, nphase, _i32(pix), prob)
    return prob


def rodrigues_matrix(r):
    q = np.empty(9)
    lib.oc_rodrigues_matrix(_f64(r), q)
    return q.reshape(3, 3)


def rotate_stiffness(rod, c11, c12, c44):
    rod = _f64(rod).ravel()
    ng = rod.size // 3
    out = np.empty(36 * max(ng, 1))
    lib.oc_rotate_stiffness(ng, rod, c11, c12, c44, out)
    return out[:36 * ng].reshape(ng, 6, 6)


def cmod(nphase, rot36, phasemod, flag_m=1):
    out = np.empty(36 * nphase)
    rot = _f64(rot36).ravel() if rot36 is not None and np.size(rot36) else np.zeros(36)
    lib.oc_cmod(nphase, flag_m, rot, _f64(phasemod).ravel(), out)
    return out.reshape(nphase, 6, 6)


def dk(cm):
    cm = _f64(cm)
    npf = cm.shape[0]
    out = np.empty(576 * npf)
    lib.oc_dk(npf, cm.ravel(), out)
    return out.reshape(npf, 8, 3, 8, 3)


def femat_rhs(shape, pix, dkm, strain6):
    nx, ny, nz = shape
    b = np.empty(3 * nx * ny * nz)
    c = C.c_double()
    lib.oc_femat_rhs(nx, ny, nz, _i32(pix), _f64(dkm).ravel(), _f64(strain6), b, C.byref(c))
    return b, c.value


def affine(shape, strain6):
    nx, ny, nz = shape
    u = np.empty(3 * nx * ny * nz)
    lib.oc_affine(nx, ny, nz, _f64(strain6), u)
    return u


def apply(shape, pix, dkm, x):
    nx, ny, nz = shape
    y = np.empty(3 * nx * ny * nz)
    lib.oc_apply(nx, ny, nz, _i32(pix), _f64(dkm).ravel(), _f64(x), y)
    return y


def energy(shape, pix, dkm, b, cval, u):
    nx, ny, nz = shape
    gb = np.empty(3 * nx * ny * nz)
    ut = C.c_double()
    lib.oc_energy(nx, ny, nz, _i32(pix), _f64(dkm).ravel(), _f64(b), cval, _f64(u), gb, C.byref(ut))
    return gb, ut.value


def sumsq(v):
    v = _f64(v)
    return lib.oc_sumsq(v.size, v)


def invert3x3(a):
    out = np.empty(9)
    lib.oc_invert3x3(_f64(a).ravel(), out)
    return out.reshape(3, 3)


def precond(shape, pix, dkm):
    nx, ny, nz = shape
    out = np.empty(9 * nx * ny * nz)
    lib.oc_precond(nx, ny, nz, _i32(pix), _f64(dkm).ravel(), out)
    return out.reshape(-1, 3, 3)


PCG_REFERENCE, PCG_SWEEP, PCG_PLAIN_CG = 0, 1, 2


def dembx(shape, pix, dkm, u, gb, gg0, tol=TOL_F32, ldemb=10000, variant=PCG_REFERENCE):
    nx, ny, nz = shape
    u = _f64(u).copy()
    gb = _f64(gb).copy()
    h = np.zeros_like(u)
    hist = np.zeros(max(ldemb, 1))
    L, gg, rc = C.c_int(), C.c_double(), C.c_int()
    st = lib.oc_dembx_variant(nx, ny, nz, _i32(pix), _f64(dkm).ravel(), u, gb, h, gg0, tol, ldemb,
                              C.byref(L), C.byref(gg), hist, C.byref(rc), variant)
    return dict(status=st, lstep=L.value, gg=gg.value, u=u, gb=gb, h=h,
                history=hist[:L.value].copy(), reset_count=rc.value)


def stress(shape, pix, cm, strain6, u):
    nx, ny, nz = shape
    s, e = np.empty(6), np.empty(6)
    lib.oc_stress(nx, ny, nz, _i32(pix), _f64(cm).ravel(), _f64(strain6), _f64(u), s, e)
    return s, e


def fullfield(shape, pix, cm, strain6, u):
    nx, ny, nz = shape
    ns = nx * ny * nz
    sf, vm = np.empty(6 * ns), np.empty(ns)
    lib.oc_fullfield(nx, ny, nz, _i32(pix), _f64(cm).ravel(), _f64(strain6), _f64(u), sf, vm)
    return sf.reshape(6, ns), vm


def solve(shape, nphase, pix, rod, c11, c12, c44, strain6, phasemod=(0.0, 0.0), flag_m=1,
          tol=TOL_F32, ldemb=10000, want_fields=True):
    nx, ny, nz = shape
    ns = nx * ny * nz
    res = OcResult()
    u = np.empty(3 * ns) if want_fields else None
    gb = np.empty(3 * ns) if want_fields else None
    hist = np.empty(ldemb)
    rod = _f64(rod).ravel() if rod is not None and np.size(rod) else np.zeros(3)
    rc = lib.oc_solve(nx, ny, nz, nphase, flag_m, _i32(pix), rod, c11, c12, c44,
                      _f64(phasemod).ravel(), _f64(strain6), tol, ldemb,
                      u.ctypes.data if u is not None else None,
                      gb.ctypes.data if gb is not None else None, hist.ctypes.data, C.byref(res))
    if rc != 0:
        raise ValueError("oc_solve rejected the input (phase label out of range)")
    out = dict(utot0=res.utot0, gg0=res.gg0, utot=res.utot, gg=res.gg, rel_res=res.rel_res,
               lstep=res.lstep, status=res.status, reset_count=res.reset_count, C=res.C,
               stress6=np.array(res.stress6[:]), strain6=np.array(res.strain6[:]),
               history=hist[:res.lstep].copy())
    if want_fields:
        out["u"], out["gb"] = u, gb
    return out


def hotspots(shape, pix, vm, nsolid, sigma=0.4, thr_mpa=250.0, npeaks=3, want_smooth=False):
    """-> (index[npeaks], value[npeaks] float32, found[, smoothed field float32])."""
    nx, ny, nz = shape
    ns = nx * ny * nz
    idx = np.zeros(npeaks, np.int64)
    val = np.zeros(npeaks, np.float32)
    sm = np.zeros(ns, np.float32) if want_smooth else None
    fp = C.POINTER(C.c_float)
    n = lib.oc_hotspots(nx, ny, nz, _i32(pix), _f64(vm), nsolid, sigma, thr_mpa, npeaks,
                        idx.ctypes.data_as(C.POINTER(C.c_int64)), val.ctypes.data_as(fp),
                        sm.ctypes.data_as(fp) if want_smooth else None)
    return (idx, val, int(n), sm) if want_smooth else (idx, val, int(n))
