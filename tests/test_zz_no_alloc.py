"""e3x_stress_dev, the entry point that allocates nothing (its block partials live in a
scratch array the library holds per GPU), against e3x_stress on the same displacement:
the same kernel and summation order, so the same bits.  Two calls in a row, the second
on a smaller grid with fewer right-hand sides, so the second reads scratch the first left
behind.  Device memory for the caller's arrays comes from the HIP runtime the library
itself is linked against (libamdhip64), not from the library."""
import ctypes as C

import numpy as np
import pytest

from _cases import voronoi_case
from elas3dxtal import _lib as L
from elas3dxtal.solver import Elas3D

pytestmark = pytest.mark.gpu

C11, C12, C44 = 246.5e9, 147.3e9, 124.7e9
H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, n):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _case(shape, nrhs, seed):
    ng = 5
    pix, rod = voronoi_case(shape, ng, seed=seed, pore_frac=0.03)
    rng = np.random.default_rng(seed)
    e6 = 1e-3 * rng.standard_normal((nrhs, 6))
    with Elas3D(shape, ng + 1, nrhs=nrhs) as s:
        s.ppixel(pix, rod)
        s.femat(C11, C12, C44, e6)
        s.affine_guess()
        _, gg0 = s.energy()
        s.dembx(gg0, ldemb=7, want_history=False)   # a displacement with a PCG correction
        st, sn = s.stress()
        cm, _ = s.tables()
        u = np.stack([s.get(L.FIELD_U, r) for r in range(nrhs)])
    return dict(shape=shape, nphase=ng + 1, pix=pix.astype(np.int32), cmod=cm, e6=e6, u=u, stress=st, strain=sn)


def _dev_averages(lib, hip, c):
    nx, ny, nz = c["shape"]
    nrhs = c["u"].shape[0]
    bufs = [DevBuf(hip, c["pix"].nbytes), DevBuf(hip, c["cmod"].nbytes), DevBuf(hip, c["u"].nbytes),
            DevBuf(hip, 8 * 12 * nrhs)]
    try:
        bufs[0].put(c["pix"])
        bufs[1].put(c["cmod"])
        bufs[2].put(c["u"])
        e6 = np.ascontiguousarray(c["e6"], dtype=np.float64).ravel()
        L.check(lib.e3x_stress_dev(nx, ny, nz, c["nphase"], nrhs, bufs[0].p, bufs[1].p, L.dptr(e6), bufs[2].p,
                                   bufs[3].p, None))
        assert hip.hipDeviceSynchronize() == 0
        return bufs[3].get(np.float64, 12 * nrhs).reshape(nrhs, 12)
    finally:
        for b in bufs:
            b.free()


def test_stress_dev_equals_stress_two_calls():
    lib, hip = L.load(), _hip()
    big = _case((24, 20, 16), 3, seed=61)
    small = _case((10, 9, 7), 1, seed=62)
    for c in (big, small):
        avg = _dev_averages(lib, hip, c)
        assert np.array_equal(avg[:, :6], c["stress"]), (avg[:, :6], c["stress"])
        assert np.array_equal(avg[:, 6:], c["strain"])


def test_stress_dev_rejects_more_rhs_than_its_scratch_holds():
    lib = L.load()
    e6 = np.zeros(6 * 7)
    with pytest.raises(L.E3xError) as e:
        L.check(lib.e3x_stress_dev(8, 8, 8, 2, 7, None, None, L.dptr(e6), None, None, None))
    assert e.value.code == L.E3X_EINVAL and "at most 6" in str(e.value)
