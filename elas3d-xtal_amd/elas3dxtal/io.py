"""Bindings of libe3x_io.so (include/e3x_io.h): HDF5 in the reference's schema and the
seeded periodic microstructure generator."""
import ctypes as C
import json
import os

import numpy as np

from . import _lib as L

PORES_NONE, PORES_VOXELS, PORES_SPHERES = 0, 1, 2


class GenParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ngrain", C.c_int),
                ("seed", C.c_uint64), ("pore_mode", C.c_int), ("pore_fraction", C.c_double),
                ("pore_median", C.c_double), ("pore_sigma", C.c_double),
                ("lof_fraction", C.c_double), ("lof_axes", C.c_double * 3)]


_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
SIGNATURES = {
    "e3x_io_last_error": (C.c_char_p, []),
    "e3x_h5_probe_input": (C.c_int, [C.c_char_p, _i64p, _i64p, _i32p]),
    "e3x_h5_read_input": (C.c_int, [C.c_char_p, C.c_int64, C.c_int, _i32p, _dp]),
    "e3x_h5_write_input": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _i32p, C.c_int, _dp]),
    "e3x_h5_write_fullfield": (C.c_int, [C.c_char_p, C.c_int64, _i32p, _dp, _dp]),
    "e3x_h5_read_fullfield": (C.c_int, [C.c_char_p, C.c_int64, _i32p, _dp, _dp]),
    "e3x_generate": (C.c_int, [C.POINTER(GenParams), _i32p, _dp, _dp, _dp]),
    "e3x_rodrigues_from_bunge": (C.c_int, [C.c_int, _dp, _dp]),
    "e3x_bunge_from_rodrigues": (C.c_int, [C.c_int, _dp, _dp]),
    "e3x_request_from_files": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(L.Request), C.POINTER(_i32p),
                                         C.POINTER(_dp), C.c_char_p, C.c_int]),
    "e3x_request_free": (None, [_i32p, _dp]),
}
_io = None


def load():
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(L.IO_LIB_PATH):
        raise RuntimeError(f"{L.IO_LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(L.IO_LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _io = lib
    return lib


def _check(rc):
    if rc != 0:
        raise L.E3xError(rc, load().e3x_io_last_error().decode(errors="replace"))


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_i32p)


def generate(shape, ngrain, seed, pore_mode=PORES_NONE, pore_fraction=0.0, pore_median=3.0,
             pore_sigma=0.4, lof_fraction=0.0, lof_axes=(20.0, 5.0, 3.0)):
    """-> (pix int32[ns], rodrigues (ngrain,3), euler (ngrain,3), achieved pore fraction)."""
    nx, ny, nz = shape
    p = GenParams(nx, ny, nz, ngrain, seed, pore_mode, pore_fraction, pore_median, pore_sigma,
                  lof_fraction, (C.c_double * 3)(*lof_axes))
    pix = np.empty(nx * ny * nz, np.int32)
    rod = np.empty(3 * ngrain)
    eul = np.empty(3 * ngrain)
    frac = C.c_double()
    _check(load().e3x_generate(C.byref(p), _i(pix), _d(rod), _d(eul), C.byref(frac)))
    return pix, rod.reshape(ngrain, 3), eul.reshape(ngrain, 3), frac.value


def rodrigues_from_bunge(euler):
    e = np.ascontiguousarray(euler, float).reshape(-1, 3)
    out = np.empty_like(e)
    _check(load().e3x_rodrigues_from_bunge(e.shape[0], _d(e), _d(out)))
    return out


def bunge_from_rodrigues(rod):
    r = np.ascontiguousarray(rod, float).reshape(-1, 3)
    out = np.empty_like(r)
    _check(load().e3x_bunge_from_rodrigues(r.shape[0], _d(r), _d(out)))
    return out


def probe_input(path):
    npix, nori = C.c_int64(), C.c_int64()
    dims = np.zeros(3, np.int32)
    _check(load().e3x_h5_probe_input(path.encode(), C.byref(npix), C.byref(nori), _i(dims)))
    return npix.value, nori.value, tuple(int(v) for v in dims)


def read_input(path, ns, nphase):
    """ppixel_hdf5 (src/elas3dxtal_pcg.f90:2879-2991)."""
    pix = np.empty(ns, np.int32)
    ori = np.empty(max(3 * (nphase - 1), 1))
    _check(load().e3x_h5_read_input(path.encode(), ns, nphase, _i(pix), _d(ori)))
    return pix, ori[:3 * (nphase - 1)]


def write_input(path, shape, pix, rodrigues):
    pix = np.ascontiguousarray(pix, np.int32)
    rod = np.ascontiguousarray(rodrigues, float).ravel()
    _check(load().e3x_h5_write_input(path.encode(), *shape, _i(pix), rod.size // 3, _d(rod)))


def write_fullfield(path, pix, vm, stress):
    pix = np.ascontiguousarray(pix, np.int32)
    vm = np.ascontiguousarray(vm, float)
    st = np.ascontiguousarray(stress, float)
    _check(load().e3x_h5_write_fullfield(path.encode(), pix.size, _i(pix), _d(vm), _d(st)))


def read_fullfield(path, ns):
    pix, vm, st = np.empty(ns, np.int32), np.empty(ns), np.empty(6 * ns)
    _check(load().e3x_h5_read_fullfield(path.encode(), ns, _i(pix), _d(vm), _d(st)))
    return pix, vm, st.reshape(6, ns)


def format_top_peaks(shape, index, value, xvec=None, yvec=None, zvec=None, nsave=3):
    """top_peaks.txt of the reference's post-processing (applications/sec3_2/A_2_0/
    elas3dxtal_postprocessing.m:223-243): the first min(3, n) peaks, value in MPa (the
    single-precision smoothed field divided by 1e6 in single, .m:207) and the voxel
    coordinates xvec(i), yvec(j), zvec(k) (default: the 0-based indices).  Raises, as the
    script does (.m:117), when no peak was found."""
    nx, ny, nz = shape
    if len(index) == 0:
        raise ValueError("No hotspots found exceeding the threshold!")
    xs = np.arange(nx, dtype=float) if xvec is None else np.asarray(xvec, float)
    ys = np.arange(ny, dtype=float) if yvec is None else np.asarray(yvec, float)
    zs = np.arange(nz, dtype=float) if zvec is None else np.asarray(zvec, float)
    n = min(nsave, len(index))
    lines = [f"Top {n} Local Maxima (Von Mises Stress):"]
    for k in range(n):
        m = int(index[k])
        i, j, kk = m % nx, (m // nx) % ny, m // (nx * ny)
        mpa = float(np.float32(value[k]) / np.float32(1e6))
        lines.append(f"Peak #{k + 1}: {mpa:.3f} MPa at ({xs[i]:.2f}, {ys[j]:.2f}, {zs[kk]:.2f})")
    return "\n".join(lines) + "\n"


def write_top_peaks(path, shape, index, value, xvec=None, yvec=None, zvec=None):
    with open(path, "w") as f:
        f.write(format_top_peaks(shape, index, value, xvec, yvec, zvec))


def request_from_files(h5_path, controls=""):
    """A serving request from an input file in the reference's schema and JSON controls
    (e3x_request_from_files).  Returns (Request, pix pointer, orientation pointer,
    full-field path or None); free the pointers with request_free once answered."""
    rq = L.Request()
    pix, ori = _i32p(), _dp()
    path = C.create_string_buffer(4096)
    text = controls if isinstance(controls, str) else json.dumps(controls)
    _check(load().e3x_request_from_files(h5_path.encode(), text.encode(), C.byref(rq), C.byref(pix), C.byref(ori),
                                         path, 4096))
    return rq, pix, ori, (path.value.decode() or None)


def request_free(pix, ori):
    load().e3x_request_free(pix, ori)
