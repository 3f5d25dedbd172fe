"""The request loop (include/e3x_serve.h) on the GPU: every result is the same served
in a mix of concurrent solves (and co-batched with other requests on its grid) as alone,
under memory-based admission, with requests arriving while others run."""
import numpy as np
import pytest

import _oracle as oc
from _cases import voronoi_case
from elas3dxtal import _lib as L
from elas3dxtal.serve import Server
from elas3dxtal.solver import solve

pytestmark = pytest.mark.gpu

C12, C44 = 147.3e9, 124.7e9


def _requests(n, seed):
    """Seeded requests of different grid sizes, grains, Zener ratios, precisions and load
    cases (one unit macro strain each)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        shape = tuple(int(v) for v in rng.choice([6, 8, 10, 12, 16], 3))
        ng = int(rng.integers(2, 9))
        pix, rod = voronoi_case(shape, ng, seed=seed * 100 + k, pore_frac=float(rng.uniform(0, 0.05)))
        a = float(rng.uniform(0.5, 5.0))
        e6 = np.zeros(6)
        j = int(rng.integers(0, 6))
        e6[j] = oc.AML_F32 if j < 3 else 0.5 * oc.AML_F32
        prec = L.PREC_MIXED if k % 2 else L.PREC_FP64
        out.append(dict(pix=pix, rod=rod, shape=shape, nphase=ng + 1, c11=C12 + 2 * C44 / a, e6=e6, prec=prec))
    return out


def _alone(q):
    r = solve(q["pix"], q["rod"], q["shape"], q["nphase"], q["c11"], C12, C44, q["e6"], precision=q["prec"])
    return dict(lstep=int(r.lstep[0]), utot=float(r.utot[0]), stress6=r.stress6[0], exit=int(r.exit[0]))


def _submit(sv, q):
    return sv.submit(q["pix"], q["rod"], q["shape"], q["nphase"], q["c11"], C12, C44, q["e6"], precision=q["prec"])


def _check_same(served, alone):
    assert served["status"] == L.E3X_OK
    assert served["exit"] == alone["exit"] == L.PCG_CONVERGED
    assert served["lstep"] == alone["lstep"]
    # fixed-order reductions: the same bits whatever runs beside it
    assert served["utot"] == alone["utot"]
    assert np.array_equal(served["stress6"], alone["stress6"])


def test_served_results_equal_solves_alone():
    qs = _requests(8, seed=11)
    with Server(max_active=4) as sv:
        ids = [_submit(sv, q) for q in qs]
        sv.drain()
        got = {r["id"]: r for r in sv.take()}
    assert sorted(got) == ids
    for i, q in zip(ids, qs):
        _check_same(got[i], _alone(q))
        assert got[i]["t_submit"] <= got[i]["t_admit"] <= got[i]["t_done"]


def test_cobatched_requests_equal_solves_alone():
    """Requests on the same small grids queued together are admitted as batches (two grid
    sizes, both precisions, some with a tighter tolerance, so they batch in groups);
    every result is bit for bit its solve alone."""
    qs = []
    for k, q in enumerate(_requests(14, seed=15)):
        shape = (12, 10, 8) if k % 2 else (16, 16, 8)
        q = dict(q, shape=shape)
        q["pix"], q["rod"] = voronoi_case(shape, q["nphase"] - 1, seed=1500 + k, pore_frac=0.02)
        q["prec"] = L.PREC_MIXED if k % 3 else L.PREC_FP64
        qs.append(q)
    with Server(max_active=3, batch_members=4) as sv:
        ids = [_submit(sv, q) for q in qs]     # all queued before the first progress pass
        sv.drain()
        got = {r["id"]: r for r in sv.take()}
    assert sorted(got) == ids
    # admissions together: the batches' members share their admission time
    admits = [got[i]["t_admit"] for i in ids]
    assert len(set(admits)) < len(admits)
    for i, q in zip(ids, qs):
        _check_same(got[i], _alone(q))


def test_serving_stream_of_production_grids():
    """The serving workload's request mix at production grid sizes (edges 64-200, 10-800
    grains, Zener ratio 0.5-5, 0-5 % spherical pores, one unit strain each, mixed
    precision with replacement every 50): the serving benchmark's 256 requests, queued at
    once and served with co-batching of the small grids, all converge; every fourth equals
    its solve alone bit for bit (iterations, energy, average stress and strain)."""
    from elas3dxtal import configs
    qs = configs.serving_requests(256)
    edges = sorted({q["shape"][0] for q in qs})
    print("edges", edges, "grains", min(q["nphase"] for q in qs) - 1, "-", max(q["nphase"] for q in qs) - 1,
          flush=True)
    assert len(edges) >= 4
    kw = dict(precision=L.PREC_MIXED, replace_every=50)
    with Server(max_active=3, batch_members=6) as sv:
        ids = [sv.submit(q["pix"], q["rod"], q["shape"], q["nphase"], q["c11"], q["c12"], q["c44"], q["e6"], **kw)
               for q in qs]
        sv.drain()
        got = {r["id"]: r for r in sv.take()}
    assert sorted(got) == ids
    for i in ids:
        assert got[i]["status"] == L.E3X_OK and got[i]["exit"] == L.PCG_CONVERGED, i
    for i, q in list(zip(ids, qs))[::4]:
        r = solve(q["pix"], q["rod"], q["shape"], q["nphase"], q["c11"], q["c12"], q["c44"], q["e6"], **kw)
        g = got[i]
        assert g["status"] == L.E3X_OK and g["exit"] == int(r.exit[0]) == L.PCG_CONVERGED, (i, q["shape"])
        assert g["lstep"] == int(r.lstep[0]) and g["utot"] == float(r.utot[0]), (i, q["shape"])
        assert np.array_equal(g["stress6"], r.stress6[0]) and np.array_equal(g["strain6"], r.strain6[0])
        assert np.all(np.isfinite(g["stress6"])) and g["utot"] > 0


def test_batching_off_gives_the_same_results():
    qs = [dict(q, shape=(12, 12, 12)) for q in _requests(4, seed=16)]
    for k, q in enumerate(qs):
        q["pix"], q["rod"] = voronoi_case(q["shape"], q["nphase"] - 1, seed=1600 + k, pore_frac=0.02)
        q["prec"] = L.PREC_MIXED
    res = []
    for members in (1, 4):
        with Server(max_active=2, batch_members=members) as sv:
            ids = [_submit(sv, q) for q in qs]
            sv.drain()
            got = {r["id"]: r for r in sv.take()}
        res.append([got[i] for i in ids])
    for a, b in zip(*res):
        assert a["lstep"] == b["lstep"] and a["utot"] == b["utot"]
        assert np.array_equal(a["stress6"], b["stress6"])
    assert len({r["t_admit"] for r in res[1]}) == 1      # one batch of four


def _handle_bytes(shape, mixed, ldemb=10000, phases=1024):
    """Device bytes the server books for a one-rhs handle (its admission estimate)."""
    ns = int(np.prod(shape))
    b = ns * 2 + 7 * 3 * ns * 8 + 6 * ns * 8 + ns * 2
    b += -(-shape[0] // 8) * -(-shape[1] // 8) * -(-shape[2] // 4) * 256   # stencil lane order
    b += phases * ((36 + 576 + 126 + 3 + 6) * 8 + (576 + 126 + 6) * 4)
    tiles = -(-shape[0] // 8) * -(-shape[1] // 8) * -(-shape[2] // 4)
    b += max(12 * 2048, 2 * tiles) * 8 + (1 << 16) + ldemb * 8
    if mixed:
        b += 6 * 3 * ns * 4 + 6 * ns * 4     # g h h_alt z Ah du, M^-1 (FP32)
    return b


def test_admission_by_memory_and_arrivals_during_work():
    """A budget that holds two solver handles: two of three requests start, the third
    waits; three more arrive while the first ones run; handles are reused."""
    qs = [dict(q, shape=(12, 12, 12)) for q in _requests(6, seed=12)]
    for k, q in enumerate(qs):
        q["pix"], q["rod"] = voronoi_case(q["shape"], q["nphase"] - 1, seed=1200 + k, pore_frac=0.02)
    budget = int(2.5 * _handle_bytes((12, 12, 12), mixed=True))
    with Server(hbm_budget=budget, max_active=6, batch_members=1) as sv:   # one handle per request
        ids = [_submit(sv, q) for q in qs[:3]]
        active, queued, _ = sv.progress()
        assert (active, queued) == (2, 1)
        while sv.progress()[2] == 0:          # until one has finished
            pass
        ids += [_submit(sv, q) for q in qs[3:]]
        sv.drain()
        got = {r["id"]: r for r in sv.take()}
        used, handles = sv.memory()
    assert handles == 2 and used <= budget
    assert sorted(got) == ids
    for i, q in zip(ids, qs):
        _check_same(got[i], _alone(q))


def test_small_grids_have_slots_of_their_own():
    """With small slots, requests on small grids (at most the co-batching size) take slots
    beside max_active: a large solve holds the only regular slot and a small request
    submitted after it starts at once; without small slots it waits.  Results equal the
    solves alone."""
    large = dict(_requests(1, seed=14)[0], shape=(24, 20, 16))
    large["pix"], large["rod"] = voronoi_case(large["shape"], large["nphase"] - 1, seed=1401, pore_frac=0.02)
    small = dict(_requests(2, seed=15)[1], shape=(8, 8, 8))
    small["pix"], small["rod"] = voronoi_case(small["shape"], small["nphase"] - 1, seed=1501, pore_frac=0.02)
    started = {}
    for slots in (0, 1):
        with Server(max_active=1, batch_members=6, batch_max_nodes=12 ** 3, small_slots=slots) as sv:
            ids = [_submit(sv, large), _submit(sv, small)]
            active, queued, _ = sv.progress()
            started[slots] = (active, queued)
            sv.drain()
            got = {r["id"]: r for r in sv.take()}
        _check_same(got[ids[0]], _alone(large))
        _check_same(got[ids[1]], _alone(small))
    assert started[0] == (1, 1) and started[1] == (2, 0)
    with Server(max_active=1) as sv, pytest.raises(L.E3xError):
        sv.set_small_slots(-1)


def test_bad_request_is_reported_not_fatal():
    qs = _requests(2, seed=13)
    bad = dict(qs[0])
    bad["pix"] = bad["pix"].copy()
    bad["pix"][0] = bad["nphase"] + 1            # label out of range (f90:2986-2989)
    with Server(max_active=2) as sv:
        i_bad = _submit(sv, bad)
        i_ok = _submit(sv, qs[1])
        sv.drain()
        got = {r["id"]: r for r in sv.take()}
    assert got[i_bad]["status"] == L.E3X_EPHASE
    _check_same(got[i_ok], _alone(qs[1]))


def test_requests_from_files_with_full_field_response(tmp_path):
    """The serving front-end on files: requests read from inputs in the reference's schema
    with JSON controls, one asking for the full field.  Each answer equals its solve alone;
    the full-field file has the reference's output datasets (f90:2840-2874) and holds that
    solve's stress and von Mises fields bit for bit."""
    from elas3dxtal import io
    qs = []
    for k, shape in enumerate([(12, 10, 8), (12, 10, 8), (10, 10, 10)]):
        ng = 4 + k
        pix, rod = voronoi_case(shape, ng, seed=1700 + k, pore_frac=0.02)
        path = str(tmp_path / f"in{k}.h5")
        io.write_input(path, shape, pix, rod)
        ctl = {"zener": 1.5 + k, "strain6": [0, 0, 0, 0, 0.5e-3, 0] if k == 1 else [1e-3, 0, 0, 0, 0, 0],
               "precision": "mixed" if k == 2 else "fp64"}
        if k == 0:
            ctl["fullfield"] = str(tmp_path / "fullfield_poly.h5")
        qs.append(dict(pix=pix, rod=rod, shape=shape, nphase=ng + 1, c11=C12 + 2 * C44 / ctl["zener"],
                       e6=np.array(ctl["strain6"]), prec=L.PREC_MIXED if k == 2 else L.PREC_FP64, path=path, ctl=ctl))
    with Server(max_active=2) as sv:
        ids = [sv.submit_files(q["path"], q["ctl"]) for q in qs]
        sv.drain()
        got = {r["id"]: r for r in sv.take()}
    assert sorted(got) == ids
    for i, q in zip(ids, qs):
        _check_same(got[i], _alone(q))
    assert got[ids[0]]["fullfield"] == qs[0]["ctl"]["fullfield"]
    assert "fullfield" not in got[ids[1]]
    q = qs[0]
    ns = int(np.prod(q["shape"]))
    pix2, vm, st = io.read_fullfield(q["ctl"]["fullfield"], ns)
    r, s = solve(q["pix"], q["rod"], q["shape"], q["nphase"], q["c11"], C12, C44, q["e6"], keep=True)
    try:
        st1, vm1 = s.stress_fullfield(0)
    finally:
        s.close()
    assert np.array_equal(pix2, q["pix"])
    assert np.array_equal(vm, vm1) and np.array_equal(st, st1)
