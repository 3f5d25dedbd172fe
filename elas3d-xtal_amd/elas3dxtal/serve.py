"""Request loop over the HIP solver (include/e3x_serve.h): continuous batching of solves
of different grid sizes on one GPU, admitted by device memory.

    with Server(max_active=8) as sv:
        rid = sv.submit(pix, rod, (64, 64, 64), nphase, c11, c12, c44, strain6)
        while ...:
            sv.progress()                 # retire, admit, advance; never waits
            for r in sv.take():           # finished responses
                ...
        sv.drain()

Each admitted request runs the reference's main sequence (ppixel, femat, affine guess,
energy, dembx, energy, stress: src/elas3dxtal_pcg.f90:380-519) on a device-resident
handle, alone or, for small grids, as a member of a batch with other queued requests on
the same grid; its results do not depend on what else is running.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .solver import LDEMB_DEFAULT, TOL_DEFAULT


def _response_dict(r):
    return dict(id=int(r.id), status=int(r.status), lstep=int(r.lstep), exit=int(r.exit_flag),
                reset_count=int(r.reset_count), shape=(int(r.nx), int(r.ny), int(r.nz)),
                utot0=float(r.utot0), utot=float(r.utot), gg=float(r.gg),
                stress6=np.array(r.stress6[:]), strain6=np.array(r.strain6[:]),
                t_submit=float(r.t_submit), t_admit=float(r.t_admit), t_done=float(r.t_done))


class Server:
    def __init__(self, device=0, hbm_budget=0, max_active=8, batch_members=None, batch_max_nodes=None,
                 small_slots=0):
        self._lib = L.load()
        h = C.c_void_p()
        L.check(self._lib.e3x_server_create(device, int(hbm_budget), int(max_active), C.byref(h)))
        self._h = h
        self._files = {}   # id -> arrays and full-field buffers of requests from files
        if batch_members is not None or batch_max_nodes is not None:
            self.set_batching(6 if batch_members is None else batch_members,
                              128 ** 3 if batch_max_nodes is None else batch_max_nodes)
        if small_slots:
            self.set_small_slots(small_slots)

    def set_batching(self, max_members, max_nodes):
        """Co-batching of small grids (e3x_server_set_batching): up to max_members queued
        requests on one grid of at most max_nodes nodes (same precision and PCG controls)
        share one batch handle; max_members = 1 turns it off."""
        L.check(self._lib.e3x_server_set_batching(self._h, int(max_members), int(max_nodes)))

    def set_small_slots(self, slots):
        """Slots of their own for the solves of small grids (e3x_server_set_small_slots):
        up to `slots` beside the max_active ones; 0 shares max_active."""
        L.check(self._lib.e3x_server_set_small_slots(self._h, int(slots)))

    def close(self):
        if self._h is not None and self._h.value:
            self._lib.e3x_server_destroy(self._h)
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

    def submit(self, pix, orientation, shape, nphase, c11, c12, c44, strain6, phasemod=(0.0, 0.0), flag_m=1,
               precision=L.PREC_MIXED, replace_every=50, tol=TOL_DEFAULT, ldemb=LDEMB_DEFAULT):
        """Queue one solve (inputs are copied); returns its id."""
        nx, ny, nz = (int(v) for v in shape)
        pix = np.ascontiguousarray(pix, dtype=np.int32).ravel()
        if pix.size != nx * ny * nz:
            raise L.E3xError(L.E3X_EINVAL, "pix size does not match the grid")
        ori = None
        if orientation is not None:
            ori = np.ascontiguousarray(orientation, dtype=np.float64).ravel()
        rq = L.Request()
        rq.nx, rq.ny, rq.nz, rq.nphase = nx, ny, nz, int(nphase)
        rq.pix = L.i32ptr(pix)
        rq.orientation = L.dptr(ori)
        rq.c11, rq.c12, rq.c44 = float(c11), float(c12), float(c44)
        pm = np.zeros(4)
        pmv = np.asarray(phasemod, float).ravel()
        pm[:pmv.size] = pmv
        rq.phasemod[:] = pm.tolist()
        rq.flag_m = int(flag_m)
        rq.strain6[:] = np.asarray(strain6, float).ravel().tolist()
        rq.precision = int(precision)
        rq.replace_every = int(replace_every) if precision == L.PREC_MIXED else 0
        rq.tol = float(tol)
        rq.ldemb = int(ldemb)
        rid = C.c_int64()
        L.check(self._lib.e3x_server_submit(self._h, C.byref(rq), C.byref(rid)))
        return rid.value

    def submit_files(self, h5_path, controls=""):
        """Queue a request read from an input file in the reference's schema with JSON
        controls (a str, or a dict to encode; keys in include/e3x_io.h,
        e3x_request_from_files).  With a "fullfield" path the response comes with the
        full field, written there as fullfield_poly.h5 (the reference's output schema)
        when take() returns it."""
        from . import io
        rq, pix, ori, ff = io.request_from_files(h5_path, controls)
        ns = rq.nx * rq.ny * rq.nz
        keep = dict(pix=pix, ori=ori, path=ff, ns=ns)
        if ff:
            keep["stress"], keep["vm"] = np.zeros(6 * ns), np.zeros(ns)
            rq.stress_field = L.dptr(keep["stress"])
            rq.vm = L.dptr(keep["vm"])
        rid = C.c_int64()
        try:
            L.check(self._lib.e3x_server_submit(self._h, C.byref(rq), C.byref(rid)))
        except Exception:
            io.request_free(pix, ori)
            raise
        self._files[rid.value] = keep
        return rid.value

    def _finish_file(self, out):
        from . import io
        keep = self._files.pop(out["id"])
        if keep["path"] and out["status"] == L.E3X_OK:
            pix = np.ctypeslib.as_array(keep["pix"], shape=(keep["ns"],)).copy()
            io.write_fullfield(keep["path"], pix, keep["vm"], keep["stress"].reshape(6, -1))
            out["fullfield"] = keep["path"]
        io.request_free(keep["pix"], keep["ori"])

    def progress(self):
        """One scheduling pass; returns (active, queued, finished-not-taken)."""
        a, q, d = C.c_int(), C.c_int(), C.c_int()
        L.check(self._lib.e3x_server_progress(self._h, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def take(self):
        """All finished responses not yet taken, oldest first (dicts)."""
        out = []
        r, got = L.Response(), C.c_int()
        while True:
            L.check(self._lib.e3x_server_take(self._h, C.byref(r), C.byref(got)))
            if not got.value:
                return out
            d = _response_dict(r)
            if d["id"] in self._files:
                self._finish_file(d)
            out.append(d)

    def drain(self):
        L.check(self._lib.e3x_server_drain(self._h))

    def clock(self):
        t = C.c_double()
        L.check(self._lib.e3x_server_clock(self._h, C.byref(t)))
        return t.value

    def memory(self):
        b, n = C.c_int64(), C.c_int()
        L.check(self._lib.e3x_server_memory(self._h, C.byref(b), C.byref(n)))
        return b.value, n.value
