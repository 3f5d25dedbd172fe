"""ctypes binding of libe3x.so (include/e3x.h).

The HIP library is the only compute path: if it is missing this module raises, it
never falls back to a CPU implementation.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(PKG_DIR), "csrc")
LIB_PATH = os.path.join(LIB_DIR, "libe3x.so")
IO_LIB_PATH = os.path.join(LIB_DIR, "libe3x_io.so")

E3X_OK, E3X_EINVAL, E3X_EDEVICE, E3X_ENOMEM, E3X_EPHASE, E3X_ESTATE, E3X_EIO = range(7)
PCG_RUNNING, PCG_CONVERGED, PCG_HAH_BREAK, PCG_MAXITER = range(4)
FIELD_U, FIELD_GB, FIELD_B, FIELD_H = range(4)
PREC_FP64, PREC_MIXED = 0, 1
STOP_REFERENCE, STOP_NIST_ABS = 0, 1
PCG_REFERENCE, PCG_SWEEP, PCG_PLAIN_CG = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p


class E3xError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"e3x error {code}: {msg}")
        self.code = code


# name -> (restype, argtypes); mirrors include/e3x.h
SIGNATURES = {
    "e3x_last_error": (C.c_char_p, []),
    "e3x_version": (C.c_int, []),
    "e3x_device_count": (C.c_int, [_ip]),
    "e3x_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "e3x_destroy": (C.c_int, [_vp]),
    "e3x_create_batch": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.POINTER(_vp)]),
    "e3x_set_member": (C.c_int, [_vp, C.c_int, _i32p, _dp, C.c_double, C.c_double, C.c_double, _dp, C.c_int,
                                 _dp]),
    "e3x_memory_bytes": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "e3x_set_microstructure": (C.c_int, [_vp, _i32p, _dp]),
    "e3x_assig": (C.c_int, [_vp, _dp]),
    "e3x_femat": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _dp, C.c_int, _dp, _dp]),
    "e3x_affine_guess": (C.c_int, [_vp]),
    "e3x_energy": (C.c_int, [_vp, _dp, _dp]),
    "e3x_dembx": (C.c_int, [_vp, _dp, C.c_float, C.c_int, C.c_int, _ip, _dp, _dp, _ip, _ip]),
    "e3x_dembx_ex": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_float, C.c_int, C.c_int, _ip, _dp,
                               _dp, _ip, _ip]),
    "e3x_set_stop_rule": (C.c_int, [_vp, C.c_int]),
    "e3x_set_pcg_variant": (C.c_int, [_vp, C.c_int]),
    "e3x_dembx_begin": (C.c_int, [_vp, C.c_int, C.c_int, _dp, C.c_float, C.c_int, C.c_int]),
    "e3x_dembx_advance": (C.c_int, [_vp, C.c_int, _ip]),
    "e3x_dembx_end": (C.c_int, [_vp, _ip, _dp, _dp, _ip, _ip]),
    "e3x_stress": (C.c_int, [_vp, _dp, _dp]),
    "e3x_hotspots": (C.c_int, [_vp, C.c_int, _dp, C.c_int, C.c_double, C.c_double, C.c_int,
                               C.POINTER(C.c_int64), _dp, C.POINTER(C.c_int64)]),
    "e3x_effective_stiffness": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _dp, C.c_int, C.c_double,
                                          C.c_int, C.c_int, C.c_float, C.c_int, _dp, _ip, _ip]),
    "e3x_stress_fullfield": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "e3x_get_field": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "e3x_set_field": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "e3x_get_tables": (C.c_int, [_vp, _dp, _dp]),
    "e3x_last_timing": (C.c_int, [_vp, _dp, _dp]),
    "e3x_time_apply": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "e3x_stress_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _dp, _vp, _vp, _vp]),
    "e3x_set_apply_timing": (C.c_int, [_vp, C.c_int]),
    "e3x_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "e3x_get_tuning": (C.c_int, [_vp, C.c_int, _ip, _dp]),
    "e3x_last_apply_timing": (C.c_int, [_vp, _ip, _dp]),
}



class Request(C.Structure):
    """struct e3x_request (include/e3x_serve.h)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nphase", C.c_int),
                ("pix", _i32p), ("orientation", _dp),
                ("c11", C.c_double), ("c12", C.c_double), ("c44", C.c_double),
                ("phasemod", C.c_double * 4), ("flag_m", C.c_int), ("strain6", C.c_double * 6),
                ("precision", C.c_int), ("replace_every", C.c_int), ("tol", C.c_float), ("ldemb", C.c_int),
                ("stress_field", _dp), ("vm", _dp)]


class Response(C.Structure):
    """struct e3x_response (include/e3x_serve.h)."""
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("lstep", C.c_int), ("exit_flag", C.c_int),
                ("reset_count", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("utot0", C.c_double), ("utot", C.c_double), ("gg", C.c_double),
                ("stress6", C.c_double * 6), ("strain6", C.c_double * 6),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


# mirrors include/e3x_serve.h
SERVE_SIGNATURES = {
    "e3x_server_create": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(_vp)]),
    "e3x_server_destroy": (C.c_int, [_vp]),
    "e3x_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "e3x_server_progress": (C.c_int, [_vp, _ip, _ip, _ip]),
    "e3x_server_take": (C.c_int, [_vp, C.POINTER(Response), _ip]),
    "e3x_server_drain": (C.c_int, [_vp]),
    "e3x_server_set_batching": (C.c_int, [_vp, C.c_int, C.c_int64]),
    "e3x_server_set_small_slots": (C.c_int, [_vp, C.c_int]),
    "e3x_server_clock": (C.c_int, [_vp, _dp]),
    "e3x_server_memory": (C.c_int, [_vp, C.POINTER(C.c_int64), _ip]),
}

_lib = None


def load():
    """Load libe3x.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in list(SIGNATURES.items()) + list(SERVE_SIGNATURES.items()):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != E3X_OK:
        msg = load().e3x_last_error().decode(errors="replace")
        raise E3xError(rc, msg)


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def i32ptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)


def device_count():
    n = C.c_int()
    check(load().e3x_device_count(C.byref(n)))
    return n.value
