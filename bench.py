"""Frequency-point solves/s on one MI355X: closed box 2 x 1 x 1 m with 5 cm
quads (4000 cells, 8000 branches, 4002 nodes, N = 12002 MNA unknowns), full
retarded exp(-jkR)/R kernel, 200 log-spaced frequencies 10 kHz - 200 MHz,
one port across the long diagonal.

A step is one frequency point: retarded Lp(w)/P(w) assembly, MNA fill,
equilibration, LU, port solve and the component-wise backward error
(<= 1e-10 gate). `value` is the mixed-precision path (FP32 MFMA factor +
FP64 iterative refinement); the plain FP64 LU path is timed on the same
points and reported beside it, with each point's impedance error against
it. The static (frequency-independent) Lp/P assembly happens once per mesh
and is reported separately.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       python bench.py --workload serving|ragged [--meshes M] [--slots S]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"fp64_mfma": 78.6e12, "fp32_mfma": 157.3e12, "fp16_mfma": 2.5e15, "fp64_vector": 78.6e12, "hbm": 8.0e12}
C0 = 299792458.0


def sweep_frequencies():
    return [1e4 * (2e4) ** (k / 199.0) for k in range(200)]


def corner_nodes(xyz, lo, hi):
    a = int(np.argmin(np.linalg.norm(xyz - np.asarray(lo), axis=1)))
    b = int(np.argmin(np.linalg.norm(xyz - np.asarray(hi), axis=1)))
    return a, b


def cpu_baseline():
    """The reference's own per-point pipeline (compiled from its sources into
    oracle/_ref) at the benchmark size, timed on all host cores:
    harmonic_build_system, harmonic_factorize (equilibrated complex LU),
    harmonic_solve and harmonic_compute_residual for the benchmark port at
    1 MHz on the quasi-static 2 x 1 x 1 m box with 5 cm quads (N = 12002).
    Its static Lp/P come from the GPU engine (equal to the reference's to
    1e-12): assembling them on the CPU is once per mesh and takes far longer
    than the sample budget; the reference has no retarded kernel, so per-point
    assembly is absent from the figure. Without the reference build the same
    sequence runs through the oracle."""
    import ctypes as C

    from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
    from peeeeeeeeeeeeeeeeeeec_amd import lib, meshgen
    from tests import checker as ck

    cores = len(os.sched_getaffinity(0))
    xyz, quads = meshgen.box(2.0, 1.0, 1.0, 40, 20, 20)
    a, b = corner_nodes(xyz, (0, 0, 0), (2, 1, 1))
    f = 1.0e6
    with lib.Model(xyz, quads) as gm:
        eng = lib.Engine(gm, smooth_moments=0)
        Ls, Ps = eng.static_matrices()
        eng.close()
    n = Ls.shape[0] + Ps.shape[0]
    if ck.ref_available():
        import tests.golden.make_golden as mg
        kind = "reference"
        m, d = mg.ref_model(xyz=xyz, quads=quads)
        L = api.InductanceMatrix(n=Ls.shape[0], data=api.c_malloc_copy(Ls, C.c_double))
        P = api.PotentialMatrix(n=Ps.shape[0], data=api.c_malloc_copy(Ps, C.c_double))
        del Ls, Ps
        t0 = time.perf_counter()
        _, _, _, eta = mg.ref_port_solve(m, d, L, P, f, a, b, threads=cores)
        t_point = time.perf_counter() - t0
    else:
        kind = "port"
        o = ck.OracleModel(xyz, quads)
        t0 = time.perf_counter()
        _, _, _, etas = ck.oracle_port_solve(o, Ls, Ps, f, a, b)
        t_point = time.perf_counter() - t0
        eta = etas[2]
        o.close()
    return {
        "value": 1.0 / t_point,
        "unit": "freq-point solves/s",
        "cores": cores,
        "kind": kind,
        "sample": (f"one full-size point: quasi-static box 2x1x1 m at h=5 cm (N={n}), 1 MHz port, MNA build + "
                   f"equilibrated complex LU + solve + backward error ({eta:.1e}) in {t_point:.1f} s on {cores} host "
                   f"threads; static Lp/P supplied (once per mesh), no retarded kernel in the reference"),
    }


def calibrate_copy():
    """A plain device-to-device copy of 1 GiB, 5 times: the known-size kernel
    whose FETCH_SIZE / WRITE_SIZE readings scale the counters of the LU
    update (profiles/lu_update_traffic_c2.md)."""
    import torch

    assert torch.cuda.is_available(), "needs a GPU"
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": x.numel() * 4, "launches": 5}))


def traffic_from_profile(name="lu_update_rank256_traffic_c2.md"):
    """HBM bytes per launch of an LU update kernel measured by rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE runs scaled by the copy calibration), read
    from the committed summary; None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def assembly_roofline(static, t_static):
    """Static Lp/P assembly against the FP64 vector peak: work from the exact
    evaluation counts (12 flops per 1/R point, 140 per analytic triangle
    potential), times from HIP events around the near-field and far-tile
    kernels during the engine's construction."""
    t_k = (static["near_ms"] + static["far_ms"]) * 1e-3
    rate = static["flops"] / t_k if t_k > 0 else 0.0
    near = static["near_flops"] / (static["near_ms"] * 1e-3) if static["near_ms"] > 0 else 0.0
    far = static["far_flops"] / (static["far_ms"] * 1e-3) if static["far_ms"] > 0 else 0.0
    return {
        "kernel": "static assembly: near_entry_kernel (adaptive analytic near field) + far_tile_kernel (1/R tiles)",
        "bound": "fp64 valu",
        "achieved": rate / 1e12,
        "peak": PEAKS["fp64_vector"] / 1e12,
        "unit": "TFLOP/s",
        "frac": rate / PEAKS["fp64_vector"],
        "near_entry_tflops": near / 1e12,
        "far_tile_tflops": far / 1e12,
        "near_ms": static["near_ms"],
        "far_ms": static["far_ms"],
        "static_assembly_s": t_static,
        "evaluations": {k: static[k] for k in ("far_evaluations", "near_potential_evaluations",
                                                "near_pair_evaluations")},
    }


def _extreme_corner_port(xyz):
    s = xyz.sum(axis=1)
    return int(np.argmin(s)), int(np.argmax(s))


def _pctl(v, q):
    return 1e3 * float(np.percentile(np.asarray(v), q)) if len(v) else None


def bench_serving(args):
    """Serving stream: seeded box/plate meshes with 500-12000 unknowns
    registered on the request server (static Lp/P assembled once each);
    requests arrive as a Poisson stream at each swept rate, each with a
    frequency on a per-mesh grid of 16 log-spaced points in 10 kHz - 200 MHz
    (so repeated points reuse cached factors) and 1-64 random port pairs.
    Reports throughput and p50/p99 latency (submission to completion) per
    rate, and the answers of a sample of requests against FP64."""
    from peeeeeeeeeeeeeeeeeeec_amd import lib, meshgen

    rng = meshgen.Rng(args.seed)
    models = []
    while len(models) < args.meshes:
        # unknowns N = Ne + Nv ~ 3 * cells for a box, ~ 3 * cells for a plate
        target_n = rng.uniform(500, 12000)
        kind = "plate" if rng.uniform() < 0.5 else "box"
        a = rng.uniform(0.3, 3.0)
        b = a / rng.uniform(1.0, 4.0)
        c = a / rng.uniform(1.0, 4.0)
        cells = target_n / 3.0
        area = a * b if kind == "plate" else 2 * (a * b + a * c + b * c)
        h = math.sqrt(area / cells)
        if kind == "plate":
            xyz, quads = meshgen.plate(max(1, int(round(a / h))), max(1, int(round(b / h))), a, b)
        else:
            xyz, quads = meshgen.box_by_size(a, b, c, h)
        m = lib.Model(xyz, quads)
        if 500 <= m.ne + m.nv <= 12000:
            models.append(m)
    sv = lib.Server(slots=args.slots, retarded=True)
    t0 = time.perf_counter()
    ids = [sv.add_mesh(m) for m in models]
    t_static = time.perf_counter() - t0
    grids = [[1e4 * (2e4) ** (k / 15.0) for k in range(16)] for _ in models]
    print(f"[bench] serving: {len(models)} meshes, N = {[m.ne + m.nv for m in models]}, "
          f"static assembly {t_static:.1f} s", flush=True)
    per_rate = []
    sample = []
    for rate in args.rates:
        reqs, t = [], 0.0
        for _ in range(args.requests):
            t += -math.log(1.0 - rng.uniform()) / rate
            mi = int(rng.uniform() * len(models))
            f = grids[mi][int(rng.uniform() * 16)]
            nrhs = 1 + int(rng.uniform() * 64)
            nv = models[mi].nv
            ports = []
            for _ in range(nrhs):
                p = int(rng.uniform() * nv)
                q = (p + 1 + int(rng.uniform() * (nv - 1))) % nv
                ports.append((p, q))
            reqs.append((t, mi, f, ports))
        tickets = []
        start = time.perf_counter()
        for t_arr, mi, f, ports in reqs:
            delay = start + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(sv.submit(ids[mi], f, ports, lib.Server.MIXED))
        lat, cached, bad = [], 0, 0
        for tk, (t_arr, mi, f, ports) in zip(tickets, reqs):
            st, z, r = sv.wait(tk)
            bad += st != 0
            lat.append(r.latency_s)
            cached += r.report.factor_cached
            if len(sample) < args.fp64_sample and st == 0:
                sample.append((mi, f, ports, z))
        elapsed = time.perf_counter() - start
        rhs = sum(len(r[3]) for r in reqs)
        per_rate.append({"rate_req_s": rate, "requests": len(reqs), "rhs": rhs, "elapsed_s": elapsed,
                         "throughput_req_s": len(reqs) / elapsed, "throughput_rhs_s": rhs / elapsed,
                         "p50_latency_ms": _pctl(lat, 50), "p99_latency_ms": _pctl(lat, 99),
                         "factor_cache_hits": cached, "failed": bad})
        print(f"[bench] rate {rate}/s: {len(reqs) / elapsed:.2f} req/s, p50 {_pctl(lat, 50):.0f} ms, "
              f"p99 {_pctl(lat, 99):.0f} ms, cache hits {cached}", flush=True)
    errs = []
    for mi, f, ports, z in sample:
        st, z64, _ = sv.wait(sv.submit(ids[mi], f, ports, lib.Server.FP64))
        if st == 0:
            errs.append(float(np.max(np.abs(z - z64) / np.abs(z64))))
    sv.close()
    best = max(per_rate, key=lambda r: r["throughput_req_s"])
    print(json.dumps({
        "metric": "serving: solve requests/s with p50/p99 latency (mixed+IR, <=1e-10 backward error)",
        "value": best["throughput_req_s"], "unit": "requests/s", "n_gpus": 1, "higher_is_better": True,
        "data": "synthetic seeded meshes and Poisson arrivals",
        "config": {"workload": f"serving stream: {len(models)} box/plate meshes, N 500-12000, "
                               f"retarded, 1-64 RHS per request, {args.slots} slots",
                   "meshes_N": [m.ne + m.nv for m in models], "static_assembly_s": t_static},
        "per_rate": per_rate,
        "answer_rel_error_vs_fp64": {"max": max(errs) if errs else None, "count": len(errs)},
    }))


def bench_ragged(args):
    """Ragged batch: `--meshes` seeded plates/boxes (cells 300-3000, sides
    0.3-3 m, aspect 1-4), each with 8 log-uniform frequencies in 1-100 MHz
    and one port between extreme corners, all queued at once on the request
    server (engines overlap on the GPU, admission by HBM budget). Mixed and
    FP64 forms are timed on the same points and compared."""
    from peeeeeeeeeeeeeeeeeeec_amd import lib, meshgen

    batch = meshgen.ragged_batch(count=args.meshes, seed=args.seed)
    sv = lib.Server(slots=args.slots, retarded=True)
    t0 = time.perf_counter()
    models = [lib.Model(b["xyz"], b["quads"]) for b in batch]
    ids = [sv.add_mesh(m) for m in models]
    t_static = time.perf_counter() - t0
    work = [(ids[i], f, [_extreme_corner_port(batch[i]["xyz"])]) for i in range(len(batch))
            for f in batch[i]["frequencies"]]
    print(f"[bench] ragged: {len(models)} meshes, N {min(m.ne + m.nv for m in models)}-"
          f"{max(m.ne + m.nv for m in models)}, static assembly {t_static:.1f} s, {len(work)} points", flush=True)
    out = {}
    for mode, name in ((lib.Server.MIXED, "mixed"), (lib.Server.FP64, "fp64")):
        start = time.perf_counter()
        tickets = [sv.submit(m, f, p, mode) for m, f, p in work]
        res = [sv.wait(t) for t in tickets]
        el = time.perf_counter() - start
        out[name] = {"elapsed_s": el, "points_s": len(work) / el, "failed": sum(r[0] != 0 for r in res),
                     "z": np.array([r[1][0] for r in res]),
                     "p50_latency_ms": _pctl([r[2].latency_s for r in res], 50),
                     "p99_latency_ms": _pctl([r[2].latency_s for r in res], 99),
                     "worst_backward_error": max(r[2].report.backward_error for r in res)}
        print(f"[bench] ragged {name}: {len(work) / el:.2f} points/s", flush=True)
    sv.close()
    err = np.abs(out["mixed"]["z"] - out["fp64"]["z"]) / np.abs(out["fp64"]["z"])
    for v in out.values():
        v.pop("z")
    print(json.dumps({
        "metric": "ragged batch: frequency-point solves/s (mixed+IR vs FP64, <=1e-10 backward error)",
        "value": out["mixed"]["points_s"], "unit": "freq-point solves/s", "n_gpus": 1, "higher_is_better": True,
        "data": "synthetic seeded meshes (splitmix64)",
        "config": {"workload": f"ragged batch: {len(models)} plates/boxes, 8 frequencies each, retarded, "
                               f"{args.slots} slots", "static_assembly_s": t_static,
                   "N": [m.ne + m.nv for m in models]},
        "mixed": out["mixed"], "fp64": out["fp64"],
        "speedup_mixed_vs_fp64": out["mixed"]["points_s"] / out["fp64"]["points_s"],
        "impedance_rel_error_vs_fp64": {"max": float(err.max()), "median": float(np.median(err))},
    }))


def bench_c5(args):
    """Large body: closed PEC cylinder r = 1.5 m, length 20 m, flat caps,
    ~19000 quads (N ~ 57000 MNA unknowns), retarded. (b) monostatic RCS at
    100 MHz for 360 incidences (theta = 90 deg, phi = 0..359 deg, e_theta)
    on one factor, mixed and FP64, sigma error in dB; (a) a sample of the
    1024-point lightning sweep (port from one cap centre to the other; the
    full sweep time is the per-point time x 1024)."""
    from peeeeeeeeeeeeeeeeeeec_amd import lib, meshgen

    xyz, quads = meshgen.cylinder(1.5, 20.0, args.c5_m, args.c5_nz)
    model = lib.Model(xyz, quads)
    n = model.ne + model.nv
    print(f"[bench] c5: {len(quads)} quads, Ne={model.ne}, Nv={model.nv}, N={n}", flush=True)
    t0 = time.perf_counter()
    eng = lib.Engine(model, retarded=True)
    t_static = time.perf_counter() - t0
    print(f"[bench] c5: static assembly {t_static:.1f} s", flush=True)
    th = np.full(360, 90.0)
    ph = np.arange(360.0)
    rcs = {}
    for mode, name in ((lib.Engine.MIXED, "mixed"), (lib.Engine.FP64, "fp64")):
        t1 = time.perf_counter()
        st, res, _, rep = eng.monostatic_rcs(1e8, th, ph, mode)
        el = time.perf_counter() - t1
        if st != 0:
            raise SystemExit(f"c5 rcs {name} failed ({st})")
        rcs[name] = {"seconds": el, "sigma_dbsm": np.array([r.sigma_dbsm for r in res]),
                     "backward_error": rep.backward_error, "refinement_steps": rep.refinement_steps,
                     "ms_assembly": rep.ms_assembly, "ms_factor": rep.ms_factor, "ms_solve": rep.ms_solve}
        print(f"[bench] c5 rcs {name}: {el:.1f} s for 360 incidences (factor {rep.ms_factor / 1e3:.1f} s, "
              f"solves {rep.ms_solve / 1e3:.1f} s)", flush=True)
    err_db = float(np.max(np.abs(rcs["mixed"]["sigma_dbsm"] - rcs["fp64"]["sigma_dbsm"])))
    a = int(np.argmin(np.linalg.norm(xyz - np.array([0.0, 0.0, 0.0]), axis=1)))
    b = int(np.argmin(np.linalg.norm(xyz - np.array([0.0, 0.0, 20.0]), axis=1)))
    df = 2.0e3  # 1024 points up to 2.048 MHz
    sample = [df, 512 * df, 1024 * df]
    t1 = time.perf_counter()
    st, z, reps = eng.port_sweep(sample, (a, b), lib.Engine.MIXED)
    el = time.perf_counter() - t1
    if st != 0:
        raise SystemExit(f"c5 sweep sample failed ({st})")
    eng.close()
    for v in rcs.values():
        v["sigma_dbsm_range"] = [float(v["sigma_dbsm"].min()), float(v["sigma_dbsm"].max())]
        v.pop("sigma_dbsm")
    print(json.dumps({
        "metric": "C5 large body: monostatic RCS, 360 incidences on one factor (s); lightning sweep s/point",
        "value": rcs["mixed"]["seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "data": "synthetic cylinder mesh",
        "config": {"workload": f"cylinder r=1.5 m, L=20 m, {len(quads)} quads, N={n}, retarded, "
                               f"100 MHz, theta=90, phi=0..359, e_theta", "static_assembly_s": t_static},
        "rcs_mixed": rcs["mixed"], "rcs_fp64": rcs["fp64"], "rcs_error_db_max": err_db,
        "lightning_sample": {"frequencies": sample, "seconds_per_point": el / len(sample),
                             "full_1024_point_sweep_s_estimate": 1024 * el / len(sample),
                             "worst_backward_error": max(r.backward_error for r in reps)},
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=200, help="sweep: points timed (200 = the whole sweep)")
    ap.add_argument("--warmup", type=int, default=17,
                    help="sweep: untimed points (the first factor with pivoting, then batches covering the "
                         "autotuner's four factor-shape trials: 1 + 4 x batch)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=4,
                    help="sweep: mixed points factored together in shared launches (1 = one at a time)")
    ap.add_argument("--cells", type=float, default=0.05, help="quad size in metres (default 5 cm)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference kernel for HBM counter scaling)")
    ap.add_argument("--workload", choices=("sweep", "serving", "ragged", "c5"), default="sweep",
                    help="sweep: the headline frequency sweep (default); serving: Poisson request stream on the "
                         "request server; ragged: batch of seeded meshes; c5: 20 m cylinder RCS + sweep sample")
    ap.add_argument("--c5-m", type=int, default=23, help="c5: cap lattice cells per side (rim = 4m cells)")
    ap.add_argument("--c5-nz", type=int, default=195, help="c5: cells along the length")
    ap.add_argument("--meshes", type=int, default=None, help="serving: 6 meshes; ragged: 16 (64 for the full batch)")
    ap.add_argument("--slots", type=int, default=4, help="serving/ragged: requests running at once")
    ap.add_argument("--rates", type=float, nargs="+", default=[2.0, 5.0, 10.0], help="serving: arrival rates (req/s)")
    ap.add_argument("--requests", type=int, default=24, help="serving: requests per rate")
    ap.add_argument("--fp64-sample", type=int, default=8, help="serving: requests re-solved in FP64")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--refinement-target", type=float, default=None,
                    help="sweep: mixed-mode refinement stops at this backward error (engine default if omitted)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload != "sweep":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU", "n_gpus": args.gpus}))
            sys.exit(2)
        if args.meshes is None:
            args.meshes = 6 if args.workload == "serving" else 16
        {"serving": bench_serving, "ragged": bench_ragged, "c5": bench_c5}[args.workload](args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(independent frequency points per GPU)", "n_gpus": args.gpus}))
        sys.exit(2)

    from peeeeeeeeeeeeeeeeeeec_amd import lib, meshgen

    # the library's own HIP runtime only (no second GPU runtime in the
    # process): device 0, synchronised with hipDeviceSynchronize
    assert lib.device_count() > 0, "bench needs a GPU"
    sync = lib.device_synchronize
    h = args.cells
    nx, ny, nz = int(round(2 / h)), int(round(1 / h)), int(round(1 / h))
    xyz, quads = meshgen.box(2.0, 1.0, 1.0, nx, ny, nz)
    port = corner_nodes(xyz, (0, 0, 0), (2, 1, 1))
    freqs = sweep_frequencies()
    model = lib.Model(xyz, quads)
    sync()
    t0 = time.perf_counter()
    eng_opts = {} if args.refinement_target is None else {"refinement_target": args.refinement_target}
    eng = lib.Engine(model, retarded=True, **eng_opts)
    sync()
    t_static = time.perf_counter() - t0
    static = eng.static_stats()
    n = model.ne + model.nv
    n_cells, n_e, n_v = int(model.mesh.n_quads), model.ne, model.nv
    K, W = args.steps, args.warmup
    idx = np.unique(np.round(np.linspace(0, len(freqs) - 1, K)).astype(int)) if K <= len(freqs) else \
        np.arange(K) % len(freqs)
    timed = [freqs[i] for i in idx]
    warm = [freqs[(i * 37 + 11) % len(freqs)] for i in range(W)]
    B = max(1, args.batch)
    # the first warm-up point alone (its partial-pivoting factor sets the row
    # order), the rest like the timed points
    for i, f in enumerate(warm[:1]):
        st, *_ = eng.solve(f, [port], lib.Engine.MIXED, want_solution=False)
        assert st == 0, f"warmup point failed ({st})"
    rest = warm[1:]
    for i in range(0, len(rest), B):
        chunk = rest[i:i + B]
        if B == 1:
            st = eng.solve(chunk[0], [port], lib.Engine.MIXED, want_solution=False)[0]
        else:
            st = eng.solve_batch(chunk, [port])[0]
        assert st == 0, f"warmup point failed ({st})"
    print(f"[bench] mesh N={n} static assembly {t_static:.2f} s; timing {len(timed)} mixed points"
          f"{'' if B == 1 else f' in batches of {B}'}", flush=True)

    eng.set_profiling(True)
    sync()
    lat, zm, reports = [], [], []
    t_start = time.perf_counter()
    for i in range(0, len(timed), B):
        chunk = timed[i:i + B]
        t1 = time.perf_counter()
        if B == 1:
            st, z, _, _, rep = eng.solve(chunk[0], [port], lib.Engine.MIXED, want_solution=False)
            zs, reps, sts = [z], [rep], [st]
        else:
            _, sts, zs, reps = eng.solve_batch(chunk, [port])
        el = time.perf_counter() - t1
        for f, st, z, rep in zip(chunk, sts, zs, reps):
            if st != 0:
                raise SystemExit(f"mixed point f={f} failed with status {st}")
            # a batch member's answer is ready when its batch is done
            lat.append(el)
            zm.append(z[0])
            reports.append(rep)
    sync()
    t_mixed = time.perf_counter() - t_start
    gemm = eng.kernel_stats("lu_update")
    g16 = eng.kernel_stats("lu_update_split16")
    gmerged = eng.kernel_stats("lu_update_merged")
    gthin = eng.kernel_stats("lu_update_thin")
    panel = eng.kernel_stats("panel")
    smooth = eng.kernel_stats("smooth_tile")
    moments = eng.kernel_stats("smooth_moments")
    eng.set_profiling(False)
    print(f"[bench] mixed: {len(timed) / t_mixed:.3f} points/s", flush=True)

    # plain FP64 LU on the same points
    eng.set_profiling(True)
    z64, lat64 = [], []
    sync()
    t_start = time.perf_counter()
    for f in timed:
        t1 = time.perf_counter()
        st, z, _, _, rep = eng.solve(f, [port], lib.Engine.FP64, want_solution=False)
        lat64.append(time.perf_counter() - t1)
        if st != 0:
            raise SystemExit(f"fp64 point f={f} failed with status {st}")
        z64.append(z[0])
    sync()
    t_fp64 = time.perf_counter() - t_start
    gemm64 = eng.kernel_stats("lu_update")
    eng.set_profiling(False)
    print(f"[bench] fp64: {len(timed) / t_fp64:.3f} points/s", flush=True)
    eng.close()
    model.close()

    zm, z64 = np.array(zm), np.array(z64)
    zerr = np.abs(zm - z64) / np.abs(z64)
    value = len(timed) / t_mixed
    gemm_tf = gemm["flops"] / (gemm["ms"] * 1e-3) if gemm["ms"] > 0 else 0.0
    # the dominant kernel: the merged trailing updates (two panel pairs per
    # pass over C); the single-pair trailing updates near the end apart
    dom = gmerged if gmerged["launches"] else g16
    g16_gbs = dom["bytes"] / (dom["ms"] * 1e-3) if dom["ms"] > 0 else 0.0
    g16_tf = dom["flops"] / (dom["ms"] * 1e-3) if dom["ms"] > 0 else 0.0

    def rate(k):
        return {"launches": k["launches"], "ms_per_point": k["ms"] / len(timed),
                "tflops": k["flops"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else None}
    out = {
        "metric": "freq-point solves/sec at <=1e-10 rel residual (mixed+IR vs FP64 LU); p99 lat",
        "value": value,
        "unit": "freq-point solves/s",
        "n_gpus": 1,
        "steps": len(timed),
        "warmup": W,
        "ms_per_step": 1e3 * t_mixed / len(timed),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex fp32 factor (split-fp16 MFMA trailing update) + complex fp64 refinement/residual",
        "data": "synthetic mesh (closed box), no external inputs",
        "config": {
            "workload": (f"closed box 2x1x1 m, {h*100:.0f} cm quads ({n_cells} cells, Ne={n_e}, "
                         f"Nv={n_v}, N={n}), retarded exp(-jkR)/R, {len(timed)} of the 200 log-spaced points "
                         f"10 kHz-200 MHz, 1 port node {port[0]} -> {port[1]}"),
            "points": [float(f) for f in timed],
            "mixed_batch": B,
        },
        "latency_note": ("per point, submission to answer; the points of a mixed batch are answered together "
                         "when the batch is done" if B > 1 else "per point"),
        "p50_latency_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_latency_ms": 1e3 * float(np.percentile(lat, 99)),
        "fp64": {
            "value": len(timed) / t_fp64,
            "p50_latency_ms": 1e3 * float(np.percentile(lat64, 50)),
            "p99_latency_ms": 1e3 * float(np.percentile(lat64, 99)),
            "lu_update_tflops": gemm64["flops"] / (gemm64["ms"] * 1e-3) / 1e12 if gemm64["ms"] else None,
        },
        "speedup_mixed_vs_fp64": value / (len(timed) / t_fp64),
        "impedance_rel_error_vs_fp64": {"max": float(zerr.max()), "median": float(np.median(zerr))},
        "backward_error_max": float(max(r.backward_error for r in reports)),
        "refinement_steps_mean": float(np.mean([r.refinement_steps for r in reports])),
        "refinement_steps": [int(r.refinement_steps) for r in reports],
        "fp64_fallbacks": int(sum(r.used_fp64_fallback for r in reports)),
        "pivot_order_reused": int(sum(r.pivot_reuse and not r.pivot_refactor for r in reports)),
        "pivot_order_refactors": int(sum(r.pivot_refactor for r in reports)),
        "ordered_factor_block_widths": sorted({int(r.block_size) for r in reports if r.block_size}),
        "stage_ms_mean": {k: float(np.mean([getattr(r, "ms_" + k) for r in reports]))
                          for k in ("assembly", "fill", "factor", "solve", "check")},
        "stage_note": "a batch member's factor and solve times are its share of the batch's shared launches",
        "static_assembly_s": t_static,
        "sweep_amortized": {
            "points": len(freqs),
            "mixed_points_per_s": len(freqs) / (t_static + len(freqs) * t_mixed / len(timed)),
            "fp64_points_per_s": len(freqs) / (t_static + len(freqs) * t_fp64 / len(timed)),
            "note": ("static Lp/P assembly once plus every point of the sweep"
                     + ("" if len(timed) == len(freqs) else f", points extrapolated from {len(timed)} timed")),
        },
        "roofline": {
            "kernel": ("LU trailing update A22 -= L21 U12 of the reused-row-order factor, two panel pairs "
                       "(2 x 256 columns of L21, each pair with its own split-FP16 scales) merged in one pass "
                       "over C, complex FP32 storage, split-FP16 operands on v_mfma_f32_16x16x32_f16 "
                       "(gemm_tile_h3_kernel<16, 4, 4, 2, 2>: 128 x 128 tiles; the small updates at the end of "
                       "the factor <16, 2, 4, 4, 2>, 64 x 128)"
                       + (f"; each launch updates the {B} points of a batch (grid z)" if B > 1 else "")),
            "bound": "mfma",
            "achieved": g16_tf / 1e12,
            "peak": PEAKS["fp16_mfma"] / 1e12,
            "unit": "TFLOP/s",
            "frac": g16_tf / PEAKS["fp16_mfma"],
            "traffic": traffic_from_profile("lu_update_merged_traffic_c2.md"),
            "launches": dom["launches"],
            "avg_launch_ms": dom["ms"] / max(1, dom["launches"]),
            "algorithmic_flops_per_launch": dom["flops"] / max(1, dom["launches"]),
            "algorithmic_bytes_per_launch": dom["bytes"] / max(1, dom["launches"]),
            "mfma_issued_tflops": 3 * g16_tf / 1e12,
            "split_fp16_roof_tflops": PEAKS["fp16_mfma"] / 3 / 1e12,
            "frac_of_split_fp16_roof": 3 * g16_tf / PEAKS["fp16_mfma"],
            "hbm_gbs": g16_gbs / 1e9,
            "hbm_frac": g16_gbs / PEAKS["hbm"],
        },
        "roofline_assembly": assembly_roofline(static, t_static),
        "other_trailing_updates": {
            "single_pair": dict(rate(g16), kernel="gemm_tile_h3_kernel<16, 2, 4, 4, 1>: trailing updates of one "
                                                  "pair (near the end of the factor)"),
            "deferred_pair_rows_and_columns": dict(rate(gthin), kernel=(
                "gemm_next_h3_kernel: the next panel's / pair's rows, and the two block columns after a "
                "deferred pair")),
        },
        "next_column_update": {
            "kernel": ("look-ahead update of the next block column, split-FP16 (gemm_next_h3_kernel<8, 2> after "
                       "the first panel of a pair, <16, 2> after the second)"),
            "launches": gemm["launches"],
            "tflops": gemm_tf / 1e12,
            "ms_per_point": gemm["ms"] / len(timed),
        },
        "other_kernels": {
            "panel_ms_per_point": panel["ms"] / len(timed),
            "smooth_tile_ms_per_point": smooth["ms"] / len(timed),
            "smooth_tile_tflops": smooth["flops"] / (smooth["ms"] * 1e-3) / 1e12 if smooth["ms"] else None,
            "smooth_moments_launches": moments["launches"],
            "smooth_moments_ms_per_launch": moments["ms"] / moments["launches"] if moments["launches"] else None,
            "smooth_moments_gbs": moments["bytes"] / (moments["ms"] * 1e-3) / 1e9 if moments["ms"] else None,
        },
        "moment_terms_per_point": [int(r.moment_terms) for r in reports],
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as exc:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": repr(exc)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
