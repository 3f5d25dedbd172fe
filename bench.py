"""Benchmark of the elas3dxtal_pcg solve on one MI355X.

Workload: the 200^3 periodic cube with 500 Voronoi grains (seeded centres, uniform
SO(3) orientations), cubic C12 = 147.3, C44 = 124.7 GPa with Zener ratio A = 2
(C11 = 272.0 GPa), 1% porosity as seeded lognormal spheres (median radius 3 voxels,
sigma 0.4), macro strain exx = float32(1e-3), the reference's PCG controls
(tol = float32(1e-8), block-Jacobi, divergence reset).

One step = one complete solve on device-resident inputs: femat (tables, jump RHS),
affine guess, energy, block-Jacobi PCG to convergence, energy, averaged stress.

value = FP64 PCG DOF-updates/s = 3 * ns * Lstep / t_pcg (t_pcg: GPU time of the
PCG loop, HIP events on the solver's stream), summed over the K timed steps.
Also reported: solves/s and per-solve latency percentiles (host wall clock), the
same for the mixed-precision form, the six unit load cases of the effective stiffness
as six right-hand sides of one PCG (FP64 and mixed; each right-hand side counts as a
solve), a serving stream of mixed-size requests with Poisson arrivals (continuous
batching: solves/s and p50/p99 latency), the roofline of the stencil-apply kernel, and
the CPU baseline (the oracle's C/OpenMP restatement of the reference on this host).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "elas3d-xtal_amd"))

from elas3dxtal import _lib as L  # noqa: E402
from elas3dxtal import configs  # noqa: E402
from elas3dxtal.solver import Elas3D  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per node of the PCG apply kernel (k_stencil, fused form used from the
# second iteration on), averaged over the two forms the PCG alternates between: every
# launch reads z, h_prev and the uint16 phase id and writes h and Ah; every other launch
# also takes two u steps (reads u and h_{it-2}, writes u).  z, h, Ah are FP64 (FP32 in the
# mixed form); the u steps go to u itself (FP64) or, in the mixed form, to an FP32
# increment folded into u at each true-gradient replacement.  FP64: (98 + 170) / 2;
# mixed: (50 + 86) / 2.
def _apply_bytes(s):
    return 4 * 3 * s + 2 + (2 * 3 * s + 3 * s) / 2


APPLY_BYTES_PER_NODE = {L.PREC_FP64: _apply_bytes(8), L.PREC_MIXED: _apply_bytes(4)}


PMC_SUMMARY = os.path.join(ROOT, "profiles", "stencil_apply_cube200_pmc.txt")


def measured_traffic(key):
    """HBM bytes per launch of the apply kernel on the 200^3 A=2 workload, from the
    FETCH_SIZE / WRITE_SIZE passes summarised in profiles/ (None if absent)."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def solve_resident(s, c11, c12, c44, e6, precision):
    """One step on the resident microstructure: the main program's sequence
    (src/elas3dxtal_pcg.f90:417-519)."""
    t0 = time.perf_counter()
    s.femat(c11, c12, c44, e6)
    s.affine_guess()
    _, gg0 = s.energy()
    d = s.dembx(gg0, precision=precision, replace_every=50, want_history=False)
    ut, _ = s.energy()
    st, sn = s.stress()
    return time.perf_counter() - t0, d, ut, st, sn


def run_gpu(w, c11, c12, c44, e6, steps, warmup, precision):
    """`steps` timed solves on the resident microstructure.  The stencil applies of the
    timed PCG loops are bracketed by HIP events on the solver's stream (their mean is the
    roofline's launch time); after the timed steps the outputs of the last solve are
    copied back for the error table."""
    s = Elas3D(w["shape"], w["nphase"])
    try:
        s.ppixel(w["pix"], w["rod"])
        for _ in range(warmup):
            solve_resident(s, c11, c12, c44, e6, precision)
        s.set_apply_timing(True)
        walls, pcg_ms, lsteps, ut, apply_ms, apply_n = [], [], [], None, [], 0
        for _ in range(steps):
            wall, d, ut, st, sn = solve_resident(s, c11, c12, c44, e6, precision)
            walls.append(wall)
            pcg_ms.append(d["pcg_ms"])
            lsteps.append(int(d["lstep"][0]))
            exit_flag = int(d["exit"][0])
            n, ms = s.last_apply_timing()
            apply_ms.append(ms * n)
            apply_n += n
        s.set_apply_timing(False)
        replay_ms = s.time_apply(precision=precision, reps=20)
        tuned = s.tuning(precision)
        _, vm = s.stress_fullfield()
        u = s.get(L.FIELD_U)
        s.affine_guess()
        u_aff = s.get(L.FIELD_U)
        return dict(u_affine=u_aff, walls=walls, pcg_ms=pcg_ms, lsteps=lsteps, exit=exit_flag, utot=float(ut[0]),
                    stress6=st[0].copy(), strain6=sn[0].copy(), u=u, vm=vm,
                    apply_ms=sum(apply_ms) / max(apply_n, 1), apply_launches=apply_n, replay_ms=replay_ms,
                    tuned_waves=tuned[0])
    finally:
        s.close()


def output_errors(mx, ref):
    """Error of every output of the mixed solve against the FP64 solve of the same
    workload: energy, each averaged stress and strain (relative to the largest component
    of its tensor), and the largest pointwise error of the displacement and von Mises
    fields (relative to the field's largest magnitude)."""
    names = ["xx", "yy", "zz", "xz", "yz", "xy"]
    smax, emax = np.max(np.abs(ref["stress6"])), np.max(np.abs(ref["strain6"]))
    out = {"utot": abs(mx["utot"] - ref["utot"]) / abs(ref["utot"])}
    for k, n in enumerate(names):
        out["stress_" + n] = float(abs(mx["stress6"][k] - ref["stress6"][k]) / smax)
    for k, n in enumerate(names):
        out["strain_" + n] = float(abs(mx["strain6"][k] - ref["strain6"][k]) / emax)
    out["u_field_max"] = float(np.max(np.abs(mx["u"] - ref["u"])) / np.max(np.abs(ref["u"])))
    # the displacement less its affine part E x (f90:423-438), which both forms hold exactly
    corr = ref["u"] - ref["u_affine"]
    out["u_correction_max"] = float(np.max(np.abs(mx["u"] - ref["u"])) / np.max(np.abs(corr)))
    out["u_field_rms"] = float(np.sqrt(np.mean((mx["u"] - ref["u"]) ** 2)) / np.sqrt(np.mean(ref["u"] ** 2)))
    out["vm_field_max"] = float(np.max(np.abs(mx["vm"] - ref["vm"])) / np.max(np.abs(ref["vm"])))
    return out


def run_six(w, c11, c12, c44, steps, warmup, precision):
    """The six unit load cases of the effective stiffness as six right-hand sides of one
    PCG (e3x_effective_stiffness): one step = femat .. averages for all six, giving C_eff."""
    s = Elas3D(w["shape"], w["nphase"], nrhs=6)
    try:
        s.ppixel(w["pix"], w["rod"])
        for _ in range(warmup):
            s.effective_stiffness(c11, c12, c44, precision=precision)
        walls, pcg_ms, lsteps, ce = [], [], [], None
        for _ in range(steps):
            t0 = time.perf_counter()
            ce, lstep, ex = s.effective_stiffness(c11, c12, c44, precision=precision)
            walls.append(time.perf_counter() - t0)
            pre, pcg = C.c_double(), C.c_double()
            L.check(s._lib.e3x_last_timing(s._h, C.byref(pre), C.byref(pcg)))
            pcg_ms.append(pcg.value)
            lsteps.append(int(np.sum(lstep)))
        return dict(walls=np.array(walls), pcg_ms=pcg_ms, lsteps=lsteps, ceff=ce, lstep_max=int(np.max(lstep)))
    finally:
        s.close()


def serving_requests(n_requests, seed=5000):
    """The serving stream's requests (configs.serving_requests)."""
    return configs.serving_requests(n_requests, seed)


def run_serving(reqs, interarrival_ms, max_active, seed=5001, batch_members=6, small_slots=0, batch_max_nodes=None):
    """Serving stream over `reqs` with Poisson arrivals (mean gap interarrival_ms; 0: all
    at once).  One host thread submits at the arrival times and drives the server between
    them; latency is arrival to result."""
    from elas3dxtal.serve import Server

    n_requests = len(reqs)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(interarrival_ms * 1e-3, n_requests) if interarrival_ms > 0 else np.zeros(n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    with Server(max_active=max_active, batch_members=batch_members, batch_max_nodes=batch_max_nodes,
                small_slots=small_slots) as sv:
        t0 = time.perf_counter()
        done = []
        k = 0
        while len(done) < n_requests:
            now = time.perf_counter() - t0
            while k < n_requests and arrivals[k] <= now:
                q = reqs[k]
                sv.submit(q["pix"], q["rod"], q["shape"], q["nphase"], q["c11"], q["c12"],
                          q["c44"], q["e6"], precision=L.PREC_MIXED, replace_every=50)
                k += 1
            sv.progress()
            done += sv.take()
            if k < n_requests and not done:
                time.sleep(1e-4)
        wall = time.perf_counter() - t0
    lat = np.array([r["t_done"] - r["t_submit"] for r in done])
    wait = np.array([r["t_admit"] - r["t_submit"] for r in done])
    service = np.array([r["t_done"] - r["t_admit"] for r in done])
    ok = sum(r["status"] == 0 and r["exit"] == L.PCG_CONVERGED for r in done)
    dofups = sum(3.0 * np.prod(r["shape"]) * r["lstep"] for r in done)
    by_edge = {}
    for e in sorted({r["shape"][0] for r in done}):
        le = np.array([r["t_done"] - r["t_submit"] for r in done if r["shape"][0] == e])
        by_edge[str(e)] = dict(n=int(le.size), p50_ms=round(float(np.percentile(le, 50) * 1e3), 1),
                               p99_ms=round(float(np.percentile(le, 99) * 1e3), 1))
    return {
        "requests": n_requests, "converged": int(ok), "max_active": max_active, "small_slots": small_slots,
        "co_batching_members": batch_members,
        "latency_by_edge": by_edge,
        "mean_interarrival_ms": interarrival_ms,
        "solves_per_s": n_requests / wall,
        "latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
        "latency_ms_p99": float(np.percentile(lat, 99) * 1e3),
        "queue_wait_ms_p50_p99": [round(float(np.percentile(wait, q) * 1e3), 1) for q in (50, 99)],
        "admitted_to_done_ms_p50_p99": [round(float(np.percentile(service, q) * 1e3), 1) for q in (50, 99)],
        "dof_updates_per_s": float(dofups / wall),
        "wall_s": wall,
        "microstructures": "one per request, seeded, generated before the stream",
    }


def cpu_baseline(w, c11, c12, c44, e6, iters):
    """The oracle (C/OpenMP restatement of the reference) on this host: `iters` PCG
    iterations on the same 200^3 workload; timed is the iteration loop alone (the oracle's
    own clock around f90:1574-2012's loop), as the GPU's t_pcg is."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as oc

    shape, pix, rod = w["shape"], w["pix"], w["rod"]
    rot = oc.rotate_stiffness(rod, c11, c12, c44)
    cm = oc.cmod(w["nphase"], rot, [0.0, 0.0])
    dk = oc.dk(cm)
    b, cval = oc.femat_rhs(shape, pix, dk, e6)
    u0 = oc.affine(shape, e6)
    g0, _ = oc.energy(shape, pix, dk, b, cval, u0)
    gg0 = oc.sumsq(g0)
    r = oc.dembx(shape, pix, dk, u0, g0, gg0, ldemb=iters)
    dt = oc.lib.oc_last_loop_seconds()
    ns = int(np.prod(shape))
    return dict(value=3.0 * ns * r["lstep"] / dt, unit="DOF-updates/s", cores=int(oc.lib.oc_get_threads()),
                kind="port",
                sample=f"the PCG iteration loop alone (block-Jacobi setup excluded), {r['lstep']} iterations of "
                       f"the 200^3 A=2 workload in {dt:.1f} s: oracle C restatement of dembx_OpenMP, OpenMP")


def profile_run(n, zener):
    """Short fixed workload for hardware-counter passes: the 200^3 problem after a
    3-iteration PCG (so z and h hold PCG values), 7 launches of the fused apply kernel
    (FP64, then the mixed form's FP32), then a fill of 3*ns doubles (3*ns*8 bytes
    written) and 3 elementwise copies (3*ns*8 bytes read and written each), whose
    counter readings calibrate the bytes."""
    import torch

    w = configs.build("cube200", n=n)
    c11, c12, c44 = configs.materials("cube200", zener=zener)
    ns = int(np.prod(w["shape"]))
    with Elas3D(w["shape"], w["nphase"]) as s:
        s.ppixel(w["pix"], w["rod"])
        s.femat(c11, c12, c44, configs.exx_strain())
        s.affine_guess()
        _, gg0 = s.energy()
        s.dembx(gg0, ldemb=3, want_history=False)
        t64 = s.time_apply(precision=L.PREC_FP64, reps=6)
        s.dembx(gg0, ldemb=3, precision=L.PREC_MIXED, want_history=False)
        t32 = s.time_apply(precision=L.PREC_MIXED, reps=6)
    x = torch.ones(3 * ns, dtype=torch.float64, device="cuda")   # fill: writes 24*ns bytes
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)                                   # reads and writes 24*ns bytes
    torch.cuda.synchronize()
    print(json.dumps({"profile": "apply", "ns": ns, "apply_us_fp64": t64 * 1e3, "apply_us_mixed": t32 * 1e3,
                      "copy_bytes_each_way": 24 * ns}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200, help="grid edge (the workload is 200)")
    ap.add_argument("--zener", type=float, default=2.0)
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-six", action="store_true", help="skip the six-load-case (effective stiffness) runs")
    ap.add_argument("--six-steps", type=int, default=2)
    ap.add_argument("--serve-requests", type=int, default=256, help="serving stream length (0: skip)")
    ap.add_argument("--serve-interarrival-ms", type=float, default=70.0)
    ap.add_argument("--serve-max-active", type=int, default=2, help="serving slots of solves above the co-batching size")
    ap.add_argument("--serve-small-slots", type=int, default=2,
                    help="serving slots of their own for co-batched small grids (0: they share --serve-max-active)")
    ap.add_argument("--serve-batch-members", type=int, default=6, help="co-batching of small grids (1: off)")
    ap.add_argument("--serve-sweep-ms", type=float, nargs="*", default=[100.0, 85.0, 0.0],
                    help="further mean interarrival gaps for the latency-against-load sweep (0: all at once)")
    ap.add_argument("--profile", action="store_true",
                    help="short fixed run of the apply kernel and a calibration copy, for rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the solve on a single MI355X "
              f"(--gpus {args.gpus} is not supported; independent solves shard one per GPU).",
              file=sys.stderr)
        sys.exit(2)

    if args.profile:
        profile_run(args.n, args.zener)
        return
    w = configs.build("cube200", n=args.n)
    c11, c12, c44 = configs.materials("cube200", zener=args.zener)
    e6 = configs.exx_strain()
    ns = int(np.prod(w["shape"]))

    fp64 = run_gpu(w, c11, c12, c44, e6, args.steps, args.warmup, L.PREC_FP64)
    t_pcg = sum(fp64["pcg_ms"]) * 1e-3
    dofups = 3.0 * ns * sum(fp64["lsteps"]) / t_pcg
    walls = np.array(fp64["walls"])

    apply_s = fp64["apply_ms"] * 1e-3
    alg_bytes = APPLY_BYTES_PER_NODE[L.PREC_FP64] * ns
    achieved = alg_bytes / apply_s / 1e9
    default_workload = args.n == 200 and args.zener == 2.0
    traffic = measured_traffic("traffic_fp64_bytes_per_launch") if default_workload else None
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel="k_stencil<double, kApplyFused> (h = z + gamma h_prev, every other launch the "
                           "two pending u steps, Ah = K h, h.Ah partials)",
                    avg_launch_us=round(fp64["apply_ms"] * 1e3, 2),
                    launch_time_source=f"HIP events on the solver's stream around the {fp64['apply_launches']} "
                                       "fused applies of the timed PCG loops",
                    replay_avg_launch_us=round(fp64["replay_ms"] * 1e3, 2),
                    waves_per_simd=fp64["tuned_waves"],
                    traffic_source=("not measured in this run: FETCH_SIZE + WRITE_SIZE per launch from the "
                                    "rocprofv3 --pmc passes summarised in profiles/stencil_apply_cube200_pmc.txt")
                    if traffic else None,
                    algorithmic_bytes_per_launch=int(alg_bytes))

    out = {
        "metric": "PCG DOF-updates/s at 200^3 (FP64)",
        "value": dofups,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(walls.mean() * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded periodic Voronoi microstructure, generated on the host)",
        "config": {"workload": f"{args.n}^3 periodic cube, 500 Voronoi grains, Zener A={args.zener:g}, "
                               "1% lognormal spherical pores, exx=1e-3, reference PCG controls",
                   "grid": list(w["shape"]), "grains": 500, "zener": args.zener,
                   "pore_fraction": w["pore_fraction"], "dof": 3 * ns,
                   "stop_rule": "reference: sqrt(z.g/gg0) < float32(1e-8)"},
        "lstep": fp64["lsteps"][0],
        "exit": fp64["exit"],
        "solves_per_s": float(1.0 / walls.mean()),
        "latency_ms_p50": float(np.percentile(walls, 50) * 1e3),
        "latency_ms_p99": float(np.percentile(walls, 99) * 1e3),
        "pcg_ms_per_solve": float(np.mean(fp64["pcg_ms"])),
        "utot": fp64["utot"],
        "roofline": roofline,
    }
    if not args.no_mixed:
        mx = run_gpu(w, c11, c12, c44, e6, args.steps, args.warmup, L.PREC_MIXED)
        mwalls = np.array(mx["walls"])
        out["mixed"] = {
            "dof_updates_per_s": 3.0 * ns * sum(mx["lsteps"]) / (sum(mx["pcg_ms"]) * 1e-3),
            "solves_per_s": float(1.0 / mwalls.mean()),
            "latency_ms_p50": float(np.percentile(mwalls, 50) * 1e3),
            "latency_ms_p99": float(np.percentile(mwalls, 99) * 1e3),
            "lstep": mx["lsteps"][0],
            "errors_vs_fp64": output_errors(mx, fp64),
            "apply_us": round(mx["apply_ms"] * 1e3, 2),
            "apply_replay_us": round(mx["replay_ms"] * 1e3, 2),
            "apply_waves_per_simd": mx["tuned_waves"],
        }
    if not args.no_six:
        six = {}
        for name, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
            r = run_six(w, c11, c12, c44, args.six_steps, 1, prec)
            six[name] = {
                "solves_per_s": float(6.0 / r["walls"].mean()),
                "dof_updates_per_s": 3.0 * ns * sum(r["lsteps"]) / (sum(r["pcg_ms"]) * 1e-3),
                "batch_ms_p50": float(np.percentile(r["walls"], 50) * 1e3),
                "batch_ms_p99": float(np.percentile(r["walls"], 99) * 1e3),
                "lstep_max": r["lstep_max"],
                "ceff_GPa": np.round(r["ceff"] / 1e9, 4).tolist(),
            }
            six[name + "_ceff"] = r["ceff"]
        ce64, ce32 = six.pop("fp64_ceff"), six.pop("mixed_ceff")
        six["ceff_mixed_rel_err_vs_fp64"] = float(np.max(np.abs(ce32 - ce64)) / np.max(np.abs(ce64)))
        six["mixed_six_vs_fp64_single_solves_ratio"] = six["mixed"]["solves_per_s"] / out["solves_per_s"]
        out["six_rhs"] = six
    if args.serve_requests > 0:
        reqs = serving_requests(args.serve_requests)
        out["serving"] = run_serving(reqs, args.serve_interarrival_ms, args.serve_max_active,
                                     batch_members=args.serve_batch_members, small_slots=args.serve_small_slots)
        # the same requests and arrival draws at other offered loads, and all at once
        # (saturation throughput): latency against load
        sweep = []
        for ia in args.serve_sweep_ms:
            r = run_serving(reqs, ia, args.serve_max_active, batch_members=args.serve_batch_members,
                            small_slots=args.serve_small_slots)
            sweep.append({"mean_interarrival_ms": ia, "offered_per_s": (1e3 / ia) if ia > 0 else None,
                          "solves_per_s": round(r["solves_per_s"], 2), "converged": r["converged"],
                          "latency_ms_p50": round(r["latency_ms_p50"], 1), "latency_ms_p99": round(r["latency_ms_p99"], 1),
                          "queue_wait_ms_p50_p99": r["queue_wait_ms_p50_p99"]})
        out["serving"]["load_sweep"] = sweep
        del reqs
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, c11, c12, c44, e6, args.cpu_iters)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
