"""Benchmark: screened-Poisson solves/s on a closest-point tube, FP32 Krylov +
FP64 refinement vs plain FP64, with p50/p99 latency, the E-gather roofline and
the CPU oracle timed on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload torus-knot|sphere|surfaces|serving|rd]

A step is one pass of the hot path over one batch: `--batch` right-hand sides
on the workload's operator solved to a 1e-10 relative residual with FP32
BiCGStab inside FP64 iterative refinement, inputs resident in HBM.  The tube,
E and L are built by the GPU engine itself.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cpm-ibc_amd", "python"))

HBM_PEAK_GBS = 8000.0


def sphere_workload(ctx, cpmgpu, batch, seed=0):
    """Unit sphere, h = 0.05, p = 3, (1 - Laplace_S) u = f, f = 3z + 21xy at cp (+ seeded smooth RHS)."""
    tube = cpmgpu.Tube(ctx, "Sphere", 0.05)
    op = tube.operator(1.0, -1.0)
    cp = tube.export()["cp"]
    rng = np.random.default_rng(seed)
    rhs = [3 * cp[:, 2] + 21 * cp[:, 0] * cp[:, 1]]
    while len(rhs) < batch:
        a = rng.standard_normal(10)
        x, y, z = cp[:, 0], cp[:, 1], cp[:, 2]
        rhs.append(a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * y + a[5] * y * z + a[6] * x * z
                   + a[7] * (x * x - y * y) + a[8] * (3 * z * z - 1) + a[9] * x * y * z)
    return dict(name="unit sphere h=0.05 p=3 screened Poisson (1-Lap)u=f, f=3z+21xy", tube=tube, op=op,
                rhs=rhs, m=1.0, n=-1.0, surface=("Sphere", 0.05, None, None))


def torus_knot_workload(ctx, cpmgpu, batch, seed=0):
    """Laplace on the torus R=1, r=0.4 (h = 0.02, p = 3) with a 2000-segment (2,3) torus-knot polyline
    carrying 2nd-order Dirichlet data u = sin(3t) (+ sin(k t + phase) for extra RHS), t the knot
    parameter of the nearest sample t_i = 2 pi i / 2000: Surface::IBCParameter's rule for the
    reference's parameterised "Torus Knot" (src/Surface.cpp:2737-2751), computed on the GPU."""
    from cpmgpu.workloads import torus_knot_polyline
    dx = 0.02
    V = torus_knot_polyline(1.0, 0.4, 2000)
    tube = cpmgpu.Tube(ctx, "Torus", dx, params=[1.0, 0.4], bbox=(-1.4, 1.4))
    prob = cpmgpu.Problem(tube, [cpmgpu.IBC("Polyline", verts=V, boundary_order=2, on_ibc_tolerance=0.1 * dx * dx)])
    op = prob.operator(0.0, -1.0)
    pts, kind = prob.rhs_points()
    idr = np.flatnonzero(kind == 2)
    t = prob.ibc_parameter(0, pts[idr], vertex_param=2 * np.pi * np.arange(2000) / 2000)
    rng = np.random.default_rng(seed)
    rhs = []
    for k in range(batch):
        b = np.zeros(prob.D)
        freq, ph = (3, 0.0) if k == 0 else (int(rng.integers(1, 9)), float(rng.uniform(0, 2 * np.pi)))
        b[idr] = np.sin(freq * t + ph)
        rhs.append(b)
    return dict(name="torus R=1 r=0.4 h=0.02 p=3, Laplace with 2nd-order Dirichlet IBC on a 2000-segment (2,3) "
                     "torus-knot polyline, u=sin(3t)", tube=tube, prob=prob, op=op, rhs=rhs, m=0.0, n=-1.0,
                surface=("Torus", dx, [1.0, 0.4], (-1.4, 1.4)), ibc_verts=V, ibc_rows=idr, ibc_t=t)


def host_cpu():
    """nproc, the CPU model (lscpu) and the host threads the oracle may use: the box gives a GPU's
    processes a share of its cores, exported as OMP_NUM_THREADS."""
    import subprocess
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc))
    return dict(nproc=nproc, model=model, threads=max(1, min(threads, nproc)),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def cpu_rates(L, E, I, b, m, n, full_iters, args, what):
    """The oracle's BiCGStab (the C++ restatement of the reference solver: same E, L and diagonal
    preconditioner, FP64 to the reference's 1e-10 stop) timed on the host on the system (L, E, I),
    in the two variants of the reference's threading: (i) as shipped, OpenMP vector updates with a
    serial operator apply (OPENMP is never defined for GeneralizedBiCGStabSystem.h:71-72); (ii)
    OpenMP everywhere.  The faster one is the baseline.  One warm-up, then the median of
    `--cpu-samples` samples of the first `--cpu-iters` iterations (fewer for the slower as-shipped
    variant), extrapolated linearly to `full_iters` (the GPU FP64 count of the same algorithm), on
    every host thread this GPU's processes may use."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    host = host_cpu()
    threads = host["threads"]
    rates, notes = {}, []
    for name, (vec, apply), iters in (("as_shipped", (True, False), max(10, args.cpu_iters // 3)),
                                      ("openmp_everywhere", (True, True), args.cpu_iters)):
        orc.set_parallel(vec, apply, threads)
        orc.solve(L, E, I, b, m, n, tol=1e-10, max_it=3)   # warm-up
        per_iter = []
        for _ in range(args.cpu_samples):
            t0 = time.perf_counter()
            x, it, rr = orc.solve(L, E, I, b, m, n, tol=1e-10, max_it=min(full_iters, iters))
            per_iter.append((time.perf_counter() - t0) / max(it, 1))
        t_full = statistics.median(per_iter) * full_iters
        rates[name] = 1.0 / t_full
        notes.append(f"{name}: median of {args.cpu_samples} samples of {min(full_iters, iters)} iterations, "
                     f"{1000 * statistics.median(per_iter):.2f} ms per iteration "
                     f"(range {1000 * min(per_iter):.2f}-{1000 * max(per_iter):.2f})")
    orc.set_parallel(False, False, 0)
    best = max(rates, key=rates.get)
    return dict(value=rates[best], unit="solves/s", cores=threads, kind="port", variant=best,
                as_shipped_solves_per_s=rates["as_shipped"],
                openmp_everywhere_solves_per_s=rates["openmp_everywhere"],
                host_nproc=host["nproc"], host_cpu_model=host["model"], omp_num_threads=host["omp_num_threads"],
                sample=f"FP64 BiCGStab on {what}, {threads} host threads; "
                       f"{'; '.join(notes)}; extrapolated linearly to the full solve's {full_iters} iterations")


def cpu_baseline(work, args):
    """The oracle builds the same system as the GPU (tube, E, L, IBC rows) and cpu_rates times it."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    surf, dx, params, bbox = work["surface"]
    if "ibc_verts" in work:
        P = orc.IBCProblem(surf, dx, params=params, bbox=bbox,
                           ibcs=[orc.IBC("Polyline", verts=work["ibc_verts"], boundary_order=2,
                                         on_ibc_tolerance=0.1 * dx * dx)])
    else:
        P = orc.Problem(surf, dx, params=params, bbox=bbox)
    return cpu_rates(P.L(), P.E(), P.ident(), work["rhs"][0], work["m"], work["n"], work["fp64_iters"], args,
                     f"the same {P.D}-row system")


def egather_traffic(kernel, workload, groups=1):
    """HBM bytes per launch of `kernel` on `workload` (`groups` operator copies per launch) measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters (scaled by the same counters on a plain copy of known
    size), from the committed summary."""
    path = os.path.join(REPO, "profiles", "egather_traffic.txt")
    if not os.path.exists(path):
        return None
    rec = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                rec[k.strip()] = v.strip()
    if rec.get("kernel") != kernel or rec.get("workload") != workload or rec.get("groups", "1") != str(groups):
        return None
    try:
        return int(float(rec["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def surfaces_measure(ctx, cpmgpu, steps, warmup, update_delta, log=True):
    """The multi-tube batch: 8 surfaces (spheres r = 0.5..2, two tori, two seeded ellipsoids) x
    h in {0.05, 0.03, 0.02, 0.01} = 32 tubes, 8 spherical-harmonic right-hand sides each (a_lm ~ N(0, 1),
    l <= 8, seeds 1000 + instance), (1 - Laplace_S) u = f; a step is one batched solve of all 256
    systems (one cpm_solve_batch_device call), inputs resident in HBM: `steps` timed steps in
    FP32+refine, then one in plain FP64.  Frees its tubes and vectors before it returns."""
    from cpmgpu import workloads
    t0 = time.perf_counter()
    hs = (0.05, 0.03, 0.02, 0.01)
    tubes, keep, ops, d_rhs, d_x, rows = [], [], [], [], [], 0
    for ih, h in enumerate(hs):
        for js, (label, typ, par, bb) in enumerate(workloads.multi_surface_set()):
            t = cpmgpu.Tube(ctx, typ, h, params=par, bbox=bb)
            op = t.operator(1.0, -1.0)
            inst = ih * 8 + js
            A = np.stack([workloads.sh_coefficients(8 * inst + k) for k in range(8)], 1)
            F = workloads.sh_rhs(t.export()["cp"], A)
            tubes.append((label, h, t.N))
            keep.append((t, op))
            rows += t.N
            if log:
                print(f"tube {label} h={h}: {t.N} nodes ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
                      flush=True)
            for k in range(8):
                d = ctx.alloc(8 * t.N)
                ctx.h2d(d, np.ascontiguousarray(F[:, k]))
                ops.append(op)
                d_rhs.append(d)
                d_x.append(ctx.alloc(8 * t.N))
            del F
    setup_s = time.perf_counter() - t0
    o32 = cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP32_REFINE, update_delta=update_delta)
    for _ in range(warmup):
        infos = ctx.solve_batch_device(ops, d_rhs, d_x, o32)
        print("surfaces: warm-up step done", file=sys.stderr, flush=True)
    ctx.synchronize()
    # step 0 tries FP32 on every system; an operator whose solve fell back to FP64 is solved in FP64
    # from the start by the later steps (the library's FP64 routing), so they show the steady state
    step_s, first_fallbacks = [], None
    for k in range(steps):
        t0 = time.perf_counter()
        infos = ctx.solve_batch_device(ops, d_rhs, d_x, o32)
        ctx.synchronize()
        step_s.append(time.perf_counter() - t0)
        if first_fallbacks is None:
            first_fallbacks = sum(1 for i in infos if i["fallback_iters"] > 0)
        print(f"surfaces: FP32+refine step {k} done ({step_s[-1]:.1f} s)", file=sys.stderr, flush=True)
    t_run = sum(step_s)
    assert all(i["status"] == "converged" for i in infos), infos
    x32 = [ctx.d2h(np.zeros(ops[q].D), d_x[q]) for q in range(0, len(ops), 37)]
    n = len(ops)
    # the same batch in plain FP64 (the reference's arithmetic), one step
    t0 = time.perf_counter()
    infos64 = ctx.solve_batch_device(ops, d_rhs, d_x, cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP64))
    ctx.synchronize()
    t64 = time.perf_counter() - t0
    print(f"surfaces: FP64 step done ({t64:.1f} s)", file=sys.stderr, flush=True)
    assert all(i["status"] == "converged" for i in infos64), infos64
    err = max(float(np.linalg.norm(x - ctx.d2h(np.zeros(ops[q].D), d_x[q])) /
                    np.linalg.norm(ctx.d2h(np.zeros(ops[q].D), d_x[q])))
              for x, q in zip(x32, range(0, len(ops), 37)))
    big = max(range(n), key=lambda q: ops[q].D)
    eg_ms, eg_bytes = ops[big].profile_egather(cpmgpu.CPM_FP32_REFINE, 8, reps=20)
    iters = [i["iters"] for i in infos]
    by_h = {}
    for q, (i32, i64) in enumerate(zip(infos, infos64)):
        h = tubes[q // 8][1]
        e = by_h.setdefault(str(h), dict(systems=0, rows=0, iters_fp32=0, fallbacks=0, fallback_iters=0,
                                         iters_fp64=0))
        e["systems"] += 1
        e["rows"] += ops[q].D
        e["iters_fp32"] += i32["iters"]
        e["fallbacks"] += i32["fallback_iters"] > 0
        e["fallback_iters"] += i32["fallback_iters"]
        e["iters_fp64"] += i64["iters"]
    for h, e in by_h.items():
        print(f"surfaces h={h}: {e}", file=sys.stderr, flush=True)
    res = dict(solves=n, steps=steps, t_run=t_run, t64=t64, setup_s=setup_s, tubes=len(tubes), rows=rows,
               largest_rows=ops[big].D, fp32refine_solves_per_s=round(n / step_s[-1], 3),
               fp32refine_first_step_solves_per_s=round(n / step_s[0], 3),
               fp64_solves_per_s=round(n / t64, 3), fp32refine_vs_fp64=round((n / step_s[-1]) / (n / t64), 3),
               fp32refine_first_step_vs_fp64=round((n / step_s[0]) / (n / t64), 3),
               fallbacks_first_step=first_fallbacks,
               routed_to_fp64=sum(1 for i in infos if i["iters"] == 0 and i["fallback_iters"] > 0),
               iters_fp32refine_mean=round(float(np.mean(iters)), 1), iters_fp32refine_max=int(max(iters)),
               iters_fp64_mean=round(float(np.mean([i["iters"] for i in infos64])), 1),
               fp64_fallbacks=sum(1 for i in infos if i["fallback_iters"] > 0),
               max_rel_diff_fp32refine_vs_fp64_sample=err,
               egather_ms=eg_ms, egather_bytes=eg_bytes, by_h=by_h,
               egather_kernel=ops[big].gather_kernel(cpmgpu.CPM_FP32_REFINE, 8))
    for d in d_rhs + d_x:
        ctx.free(d)
    for t, op in keep:
        op.close()
        t.close()
    return res


def bench_surfaces(args, ctx, cpmgpu):
    """The multi-tube batch as its own benchmark line (see surfaces_measure)."""
    r = surfaces_measure(ctx, cpmgpu, args.steps, args.warmup, args.update_delta)
    achieved = r["egather_bytes"] / (r["egather_ms"] * 1e-3) / 1e9
    out = {
        "metric": "screened-Poisson solves/s, FP32+refine at 1e-10, many tubes in one batch",
        "value": r["fp32refine_solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * r["t_run"] / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 Krylov + f64 refinement",
        "data": "synthetic (GPU-built tubes of analytic surfaces; seeded spherical-harmonic RHS)",
        "config": {"workload": "8 surfaces x h {0.05,0.03,0.02,0.01} = 32 tubes x 8 RHS, (1-Lap)u=f, "
                               "lmax 8 spherical-harmonic RHS", "tubes": r["tubes"], "solves_per_step": r["solves"],
                   "tube_rows_total": r["rows"], "largest_tube_rows": r["largest_rows"], "tol": 1e-10,
                   "precond": "diagonal", "setup_s": round(r["setup_s"], 1)},
        "iters_mean": r["iters_fp32refine_mean"], "iters_max": r["iters_fp32refine_max"],
        "fp64_fallbacks": r["fp64_fallbacks"], "fp64_solves_per_s": r["fp64_solves_per_s"],
        "fp32refine_vs_fp64": r["fp32refine_vs_fp64"], "iters_fp64_mean": r["iters_fp64_mean"], "by_h": r["by_h"],
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, kernel=r["egather_kernel"],
                         avg_launch_ms=round(r["egather_ms"], 5), bytes_per_launch=int(r["egather_bytes"]),
                         on=f"largest tube ({r['largest_rows']} rows)"),
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    ctx.close()


def bench_serving(args, ctx, cpmgpu):
    """Serving loop: displaced icosphere (6 subdivisions, 81,920 triangles, radial displacement of
    amplitude 0.1 from degree <= 6 harmonics, seed 4), h = 0.01; tube and BVH stay resident.  Requests
    arrive as a Poisson process at --qps (exponential gaps, numpy seed 44); each carries 1..8 closed
    curves (plane sections of the mesh, 256 samples) with a random RGB colour as 2nd-order Dirichlet
    data, i.e. 3 right-hand sides of Laplace's equation.  On arrival the request's IBC problem and
    operator are built on the GPU; it then joins the running FP32+refine iteration (cpm_server) as soon
    as pool memory allows and leaves when its systems stop.  Latency = completion - arrival."""
    from cpmgpu import workloads
    h = 0.01
    V, F = workloads.displaced_icosphere(6, seed=4)
    t0 = time.perf_counter()
    tube = cpmgpu.Tube(ctx, "Triangulation", h, verts=V, faces=F)
    print(f"mesh tube: {tube.N} nodes ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    n = args.requests
    gaps = np.random.default_rng(44).exponential(1.0 / args.qps, n)
    arrive = np.cumsum(gaps) - gaps[0]
    payload = []
    for i in range(n):
        payload.append(workloads.serving_request(V, F, 4400 + i))
    print(f"{n} requests generated ({sum(len(c) for c, _ in payload)} curves)", file=sys.stderr, flush=True)
    prec = cpmgpu.CPM_FP64 if args.serve_precision == "fp64" else cpmgpu.CPM_FP32_REFINE
    srv = cpmgpu.Server(ctx, int(args.pool_gb * (1 << 30)), lanes=4,
                        o=cpmgpu.opts(tol=1e-10, precision=prec, update_delta=args.update_delta))
    # A build thread with its own context builds each request's problem, operator and right-hand sides
    # when it arrives (its kernels run on its own stream beside the iterations) and releases finished
    # requests; this thread admits built requests and drives the iterations.
    import queue
    import threading
    bctx = cpmgpu.Context(0)
    built, finished = queue.Queue(), queue.Queue()
    start = time.perf_counter()

    def release():
        """Frees finished requests; True once the loop has ended."""
        while True:
            try:
                item = finished.get_nowait()
            except queue.Empty:
                return False
            if item is None:
                return True
            P, op, dr, dx = item
            for d in dr + dx:
                bctx.free(d)
            op.close()
            P.close()

    def builder():
        for i in range(n):
            while time.perf_counter() - start < arrive[i]:
                release()
                time.sleep(0.002)
            tb = time.perf_counter()
            curves, cols = payload[i]
            P = cpmgpu.Problem(tube, [cpmgpu.IBC("Polyline", verts=c, boundary_order=2, on_ibc_tolerance=0.1 * h * h)
                                      for c in curves], ctx=bctx)
            op = P.operator(0.0, -1.0)
            dr = [bctx.alloc(8 * P.D) for _ in range(3)]
            dx = [bctx.alloc(8 * P.D) for _ in range(3)]
            for ch in range(3):                            # each curve's Dirichlet rows take its colour
                b = np.zeros(P.D)
                for c in range(len(curves)):
                    _, _, r0, nr = P.curve_rows(c)
                    b[r0:r0 + nr] = cols[c, ch]
                bctx.h2d(dr[ch], b)
            built.put((i, (P, op, dr, dx), 1000 * (time.perf_counter() - tb)))
        while not release():                               # releases until the loop ends
            time.sleep(0.005)

    th = threading.Thread(target=builder, daemon=True)
    th.start()
    live, waiting, rec = {}, [], {}
    while len(rec) < n:
        while True:
            try:
                i, req, bms = built.get_nowait()
            except queue.Empty:
                break
            live[i] = req + (bms,)
            waiting.append(i)
        while waiting:                                     # admission by pool memory
            i = waiting[0]
            if not srv.submit(live[i][1], live[i][2], live[i][3], i):
                break
            waiting.pop(0)
        done, active = srv.run(16)
        for tag, infos in done:
            t_done = time.perf_counter() - start
            P, op, dr, dx, bms = live.pop(tag)
            rec[tag] = dict(latency_ms=1000 * (t_done - arrive[tag]), build_ms=bms, rows=P.D,
                            iters=max(j["iters"] for j in infos),
                            fallback=sum(1 for j in infos if j["fallback_iters"] > 0),
                            routed=sum(1 for j in infos if j["fallback_iters"] > 0 and j["iters"] == 0),
                            ok=all(j["status"] == "converged" for j in infos))
            print(f"request {tag}: {len(payload[tag][0])} curves, {P.D} rows, latency {rec[tag]['latency_ms']:.0f} ms",
                  file=sys.stderr, flush=True)
            finished.put((P, op, dr, dx))
        if active == 0 and not waiting:
            time.sleep(0.002)
    makespan = time.perf_counter() - start
    finished.put(None)
    th.join()
    srv.close()
    bctx.close()
    # request 0's system again, on this context: the E gather the server's iterations run (FP32, 4
    # interleaved lanes) against the HBM roofline, its FP64 iteration count, and the oracle's solve of
    # the same matrices on the host
    curves, cols = payload[0]
    P0 = cpmgpu.Problem(tube, [cpmgpu.IBC("Polyline", verts=c, boundary_order=2, on_ibc_tolerance=0.1 * h * h)
                               for c in curves])
    op0 = P0.operator(0.0, -1.0)
    b0 = np.zeros(P0.D)
    for c in range(len(curves)):
        _, _, r0, nr = P0.curve_rows(c)
        b0[r0:r0 + nr] = cols[c, 0]
    _, inf0 = op0.solve(b0, cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP64))
    eg_ms, eg_bytes = op0.profile_egather(prec, 4, reps=30)
    achieved = eg_bytes / (eg_ms * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel=op0.gather_kernel(prec, 4), avg_launch_ms=round(eg_ms, 5),
                bytes_per_launch=int(eg_bytes), on=f"request 0 ({P0.D} rows)")
    cb = None
    if not args.no_cpu:
        L0, E0, I0 = P0.export()
        cb = cpu_rates(L0, E0, I0, b0, 0.0, -1.0, inf0["iters"], args,
                       f"request 0's {P0.D}-row system (matrices exported from the GPU build)")
    op0.close()
    P0.close()
    lat = sorted(r["latency_ms"] for r in rec.values())
    p99 = lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)]
    out = {
        "metric": f"serving: solves/s with p50/p99 request latency, "
                  f"{'FP64' if args.serve_precision == 'fp64' else 'FP32+refine'} at 1e-10",
        "value": round(3 * n / makespan, 3), "unit": "solves/s", "n_gpus": 1, "steps": n, "warmup": 0,
        "ms_per_step": round(1000 * makespan / n, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.serve_precision == "fp64" else "f32 Krylov + f64 refinement",
        "data": "synthetic (seeded displaced icosphere mesh, plane-section curves, random RGB)",
        "config": {"workload": "serving loop: displaced icosphere 81,920 triangles, h=0.01, Poisson arrivals, "
                               "1-8 curves x 3 RHS per request", "tube_nodes": tube.N, "requests": n,
                   "qps": args.qps, "pool_gb": args.pool_gb, "lanes": 4},
        "p50_ms": round(statistics.median(lat), 1), "p99_ms": round(p99, 1),
        "build_ms_mean": round(float(np.mean([r["build_ms"] for r in rec.values()])), 1),
        "rows_mean": int(np.mean([r["rows"] for r in rec.values()])),
        "iters_mean": round(float(np.mean([r["iters"] for r in rec.values()])), 1),
        "fp64_fallbacks": int(sum(r["fallback"] for r in rec.values())),
        "routed_to_fp64": int(sum(r["routed"] for r in rec.values())),
        "all_converged": all(r["ok"] for r in rec.values()),
        "fp64_iters_request0": inf0["iters"],
        "roofline": roof, "cpu_baseline": cb,
    }
    print(json.dumps(out))
    ctx.close()


def bench_rd(args, ctx, cpmgpu):
    """Large time-stepped batch: Gray-Scott (F 0.037, k 0.06, D_u 2e-5, D_v 1e-5) on a displaced icosphere
    (8 subdivisions, 1,310,720 triangles, amplitude 0.1, seed 5) at h = 0.005 with Neumann IBCs on 4 closed
    curves (parallel plane sections), implicit diffusion with dt = 1: each step solves
    (1 - dt D Lap) u' = u + dt R(u, v) for u and v of 16 seeded initial conditions, 32 systems in one
    FP32+refine batch (cpm_rd_step_device).  A step is one time step of all 16 states; --steps steps are
    timed (the full run is 2000 steps)."""
    from cpmgpu import workloads
    h = 0.005
    t0 = time.perf_counter()
    V, F = workloads.displaced_icosphere(8, seed=5)
    tube = cpmgpu.Tube(ctx, "Triangulation", h, verts=V, faces=F)
    print(f"mesh tube: {tube.N} nodes ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    curves = workloads.plane_section_curves(V, F, 4, seed=55)
    prob = cpmgpu.Problem(tube, [cpmgpu.IBC("Polyline", verts=c, boundary_order=2, boundary_type=(1, 1),
                                            on_ibc_tolerance=0.1 * h * h) for c in curves])
    op = prob.operator(1.0, -1.0)
    D = prob.D
    print(f"problem: {D} rows, {len(curves)} Neumann curves ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
          flush=True)
    p = cpmgpu.rd_params(h, method=1)
    p.F, p.k, p.nu_u, p.nu_v, p.dt = 0.037, 0.06, 2e-5, 1e-5, 1.0
    pts, kind = prob.rhs_points()
    tubemask = kind == 0
    n_ic = 16
    d_u = [ctx.alloc(8 * D) for _ in range(n_ic)]
    d_v = [ctx.alloc(8 * D) for _ in range(n_ic)]
    for i in range(n_ic):
        pert = np.random.default_rng(500 + i).uniform(0.0, 1.0, D)
        ctx.h2d(d_u[i], np.where(tubemask, 1.0 - 0.5 * pert, 0.0))
        ctx.h2d(d_v[i], np.where(tubemask, 0.25 * pert, 0.0))
    d_g = ctx.alloc(8 * D)
    ctx.h2d(d_g, np.zeros(D))
    setup_s = time.perf_counter() - t0
    o = cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP32_REFINE, update_delta=args.update_delta)
    for _ in range(args.warmup):
        op.rd_step_device(p, d_u, d_v, d_g, d_g, o)
    ctx.synchronize()
    iters = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        infos = op.rd_step_device(p, d_u, d_v, d_g, d_g, o)
        iters += [j["iters"] for j in infos]
        assert all(j["status"] == "converged" for j in infos), infos
        print(f"step {k} done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    ctx.synchronize()
    t_run = time.perf_counter() - t0
    # the same steps in plain FP64 from the current states (one step, for the ratio)
    o64 = cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP64)
    t1 = time.perf_counter()
    infos64 = op.rd_step_device(p, d_u, d_v, d_g, d_g, o64)
    ctx.synchronize()
    t64 = time.perf_counter() - t1
    u0 = ctx.d2h(np.zeros(D), d_u[0])
    out = {
        "metric": "time-stepped batch: Gray-Scott steps/s (32 screened-Poisson solves per step), FP32+refine",
        "value": round(args.steps / t_run, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * t_run / args.steps, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 Krylov + f64 refinement",
        "data": "synthetic (seeded displaced icosphere mesh, plane-section curves, seeded random initial states)",
        "config": {"workload": "Gray-Scott on a displaced icosphere (1,310,720 triangles), h=0.005, 4 Neumann curves, "
                               "dt=1, 16 initial states", "tube_nodes": tube.N, "rows": D, "states": n_ic,
                   "solves_per_step": 2 * n_ic, "setup_s": round(setup_s, 1)},
        "solves_per_s": round(2 * n_ic * args.steps / t_run, 2),
        "iters_mean": round(float(np.mean(iters)), 1), "iters_max": int(max(iters)),
        "fp64_step_ms": round(1000 * t64, 1), "iters_fp64_mean": round(float(np.mean([j["iters"] for j in infos64])), 1),
        "full_run_2000_steps_s": round(2000 * t_run / args.steps, 0),
        "u0_range": [float(u0[tubemask].min()), float(u0[tubemask].max())],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64,
                    help="streamed lanes (1, 2, 4, 8, or 16, 32 and 64 as groups of 8 in every launch)")
    ap.add_argument("--stream", type=int, default=128, help="right-hand sides streamed through the lanes per step")
    ap.add_argument("--workload", default="torus-knot", choices=["torus-knot", "sphere", "surfaces", "serving", "rd"])
    ap.add_argument("--requests", type=int, default=16, help="serving: requests")
    ap.add_argument("--qps", type=float, default=1.0, help="serving: mean arrival rate")
    ap.add_argument("--pool-gb", type=float, default=8.0, help="serving: workspace pool")
    ap.add_argument("--serve-precision", default="fp32refine", choices=["fp32refine", "fp64"],
                    help="serving: precision of the request loop")
    ap.add_argument("--cpu-iters", type=int, default=100, help="iterations per CPU-baseline sample")
    ap.add_argument("--cpu-samples", type=int, default=5, help="CPU-baseline samples (the median is reported)")
    ap.add_argument("--surfaces-steps", type=int, default=2,
                    help="torus-knot run: timed steps of the 256-system many-tubes batch (0: skip)")
    ap.add_argument("--update-delta", type=float, default=1e-2)
    ap.add_argument("--latency-samples", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (independent solves would shard "
              f"across GPUs by instance with no communication)", file=sys.stderr)
        sys.exit(2)

    import cpmgpu
    ctx = cpmgpu.Context(0)
    if args.workload == "surfaces":
        return bench_surfaces(args, ctx, cpmgpu)
    if args.workload == "serving":
        return bench_serving(args, ctx, cpmgpu)
    if args.workload == "rd":
        return bench_rd(args, ctx, cpmgpu)
    t0 = time.perf_counter()
    S = max(args.stream, args.batch)
    if args.workload == "sphere":
        work = sphere_workload(ctx, cpmgpu, S)
    else:
        work = torus_knot_workload(ctx, cpmgpu, S)
    setup_s = time.perf_counter() - t0
    op, rhs = work["op"], work["rhs"]
    D = op.D
    B = args.batch
    tol = 1e-10

    # resident inputs/outputs
    d_rhs = [ctx.alloc(8 * D) for _ in range(S)]
    d_x = [ctx.alloc(8 * D) for _ in range(S)]
    for d, r in zip(d_rhs, rhs):
        ctx.h2d(d, np.ascontiguousarray(r))
    o32 = cpmgpu.opts(tol=tol, precision=cpmgpu.CPM_FP32_REFINE, update_delta=args.update_delta)
    o64 = cpmgpu.opts(tol=tol, precision=cpmgpu.CPM_FP64)
    ops = [op] * B

    # the step: S right-hand sides streamed through B interleaved lanes (FP32 + FP64 refinement)
    for _ in range(args.warmup):
        infos = ctx.solve_stream_device(op, d_rhs, d_x, B, o32)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos = ctx.solve_stream_device(op, d_rhs, d_x, B, o32)
    ctx.synchronize()
    t_run = time.perf_counter() - t0
    value = S * args.steps / t_run
    ms_per_step = 1000 * t_run / args.steps
    x32 = [ctx.d2h(np.zeros(D), d) for d in d_x[:1]]
    info32 = infos[0]
    assert all(i["status"] == "converged" for i in infos), infos
    iters_all = [i["iters"] for i in infos]

    # one fixed batch of B lanes (the batch waits for its slowest system)
    nb_steps = max(2, args.steps // 4)
    ctx.solve_batch_device(ops, d_rhs[:B], d_x[:B], o32)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(nb_steps):
        ctx.solve_batch_device(ops, d_rhs[:B], d_x[:B], o32)
    ctx.synchronize()
    batch_rate = B * nb_steps / (time.perf_counter() - t0)

    # plain FP64 GPU path: one right-hand side per solve, reference algorithm
    ctx.solve_batch_device([op], d_rhs[:1], d_x[:1], o64)
    n64 = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for k in range(n64):
        inf64 = ctx.solve_batch_device([op], [d_rhs[k % B]], [d_x[k % B]], o64)[0]
    ctx.synchronize()
    fp64_rate = n64 / (time.perf_counter() - t0)
    inf64 = ctx.solve_batch_device([op], d_rhs[:1], d_x[:1], o64)[0]
    x64 = ctx.d2h(np.zeros(D), d_x[0])

    # FP64 with the same batching: the step's S right-hand sides as consecutive batches of 4, 8, 32
    # and 64 (FP64 groups 4 interleaved lanes per operator copy; it has no streamed mode: each batch
    # waits for its slowest system)
    fp64_batched = {}
    for lanes in (4, 8, 32, 64):
        ctx.solve_batch_device([op] * lanes, d_rhs[:lanes], d_x[:lanes], o64)
        ctx.synchronize()
        t0 = time.perf_counter()
        for k in range(0, S, lanes):
            ctx.solve_batch_device([op] * lanes, d_rhs[k:k + lanes], d_x[k:k + lanes], o64)
        ctx.synchronize()
        fp64_batched[lanes] = S / (time.perf_counter() - t0)
        print(f"fp64 batched {lanes} lanes: {fp64_batched[lanes]:.1f} solves/s", file=sys.stderr, flush=True)
    fp64_best = max(fp64_batched.values())

    # latency of a single request (one RHS, FP32+refine)
    lat = []
    for k in range(args.latency_samples):
        t0 = time.perf_counter()
        ctx.solve_batch_device([op], [d_rhs[k % B]], [d_x[k % B]], o32)
        lat.append(1000 * (time.perf_counter() - t0))
    lat.sort()
    p50 = statistics.median(lat)
    p99 = lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)]

    # roofline of the dominant kernel: E-gather, FP32, 8 interleaved RHS in B / 8 groups per launch (what
    # each streamed iteration runs)
    nr = 8 if B >= 8 else (4 if B >= 4 else (2 if B >= 2 else 1))
    groups = max(1, B // 8)
    eg_ms, eg_bytes = op.profile_egather(cpmgpu.CPM_FP32_REFINE, nr, reps=50, groups=groups)
    achieved = eg_bytes / (eg_ms * 1e-3) / 1e9
    kname = op.gather_kernel(cpmgpu.CPM_FP32_REFINE, nr)
    # bytes_per_launch counts the operator's data once per group (each group is one operator copy with
    # its nr right-hand sides); the groups of a streamed solve share one operator, so the launch's
    # distinct bytes are the operator's data once plus every group's x read and tmp written
    per_group_vec = 2.0 * op.D * nr * 4
    unique = eg_bytes / groups + (groups - 1) * per_group_vec
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=egather_traffic(kname, args.workload, groups),
                kernel=kname, groups_per_launch=groups, avg_launch_ms=round(eg_ms, 5),
                bytes_per_launch=int(eg_bytes), distinct_bytes_per_launch=int(unique),
                distinct_achieved=round(unique / (eg_ms * 1e-3) / 1e9, 1))
    copy_bytes = 1 << 29
    copy_ms = ctx.profile_copy(copy_bytes, 20)

    # one batch with host buffers (PCIe transfers of rhs and x included)
    t0 = time.perf_counter()
    ctx.solve_batch(ops, rhs[:B], o32)
    pcie_rate = B / (time.perf_counter() - t0)
    eg64_ms, eg64_bytes = op.profile_egather(cpmgpu.CPM_FP64, 1, reps=50)

    out = {
        "metric": "screened-Poisson solves/s + p50/p99 latency, FP32+refine vs FP64 at 1e-8",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Krylov + f64 refinement",
        "data": "synthetic (GPU-built tube of an analytic surface; analytic + seeded random RHS)",
        "config": {"workload": work["name"], "rows": D, "rhs_per_step": S, "lanes": B, "tol": tol,
                   "precond": "diagonal", "update_delta": args.update_delta,
                   "tube_build_ms": round(work["tube"].info["build_ms"], 2), "setup_s": round(setup_s, 3)},
        # FP32+refine against FP64 with the same batching (the fastest FP64 batch), and
        # against FP64 one right-hand side per solve (batching and precision together)
        "fp32refine_vs_fp64": round(value / fp64_best, 3),
        "fp64_batched_solves_per_s": {str(k): round(v, 3) for k, v in fp64_batched.items()},
        "fp64_solves_per_s": round(fp64_rate, 3),
        "fp32refine_vs_fp64_1rhs": round(value / fp64_rate, 3),
        "accuracy": "solves stop at relative residual 1e-10 (the reference's stop, src/Solver.cpp:110); "
                    "the metric's 1e-8 is the relative-error bar vs FP64, reported in err_*",
        "p50_ms": round(p50, 3),
        "p99_ms": round(p99, 3),
        "batched_solves_per_s": round(batch_rate, 3),
        "iters_fp32refine": info32["iters"], "updates": info32["updates"],
        "iters_fp32refine_mean": round(float(np.mean(iters_all)), 1), "iters_fp32refine_max": int(max(iters_all)),
        "iters_fp64": inf64["iters"],
        "egather_fp64_GBs": round(eg64_bytes / (eg64_ms * 1e-3) / 1e9, 1),
        "copy_GBs": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1),
        "pcie_inclusive_solves_per_s": round(pcie_rate, 3),
        "roofline": roof,
    }
    # accuracy of both precisions against a tight FP64 solve (the tests pin the FP64 GPU path to the oracle)
    tight = cpmgpu.opts(tol=1e-13, max_iters=50000)
    ctx.solve_batch_device([op], d_rhs[:1], d_x[:1], tight)
    x_tight = ctx.d2h(np.zeros(D), d_x[0])
    out["err_fp32refine_vs_fp64_tight"] = float(np.linalg.norm(x32[0] - x_tight) / np.linalg.norm(x_tight))
    out["err_fp64_vs_fp64_tight"] = float(np.linalg.norm(x64 - x_tight) / np.linalg.norm(x_tight))
    if args.workload == "sphere":
        # max surface error vs the exact u = z + 3xy at 200 x 200 (theta, phi) plot points, through the
        # interpolation rows of the tube (Helpers::GetPlotInterpMatrix, src/Helpers.cpp:97-103)
        th = np.repeat(np.arange(200) * (2 * np.pi / 199), 200)
        ph = np.tile(np.arange(200) * (np.pi / 199), 200)
        xp = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
        Ep = work["tube"].interp(xp)
        uex = xp[:, 2] + 3 * xp[:, 0] * xp[:, 1]
        out["max_surface_error_fp64"] = float(np.abs(Ep.matvec(x_tight) - uex).max())
        out["max_surface_error_fp32refine"] = float(np.abs(Ep.matvec(x32[0]) - uex).max())
    if "prob" in work:
        out["config"]["tube_nodes"] = work["prob"].info["tube_nodes"]
        out["config"]["ibc_dofs"] = work["prob"].info["ibc_dofs"]
        out["config"]["problem_build_ms"] = round(work["prob"].info["build_ms"], 2)
    if not args.no_cpu:
        work["fp64_iters"] = inf64["iters"]
        cb = cpu_baseline(work, args)
        out["cpu_baseline"] = cb
        out["gpu_vs_cpu"] = round(value / cb["value"], 1)
    if args.surfaces_steps > 0 and args.workload == "torus-knot":
        # the many-tubes screened-Poisson batch (256 systems) in the same run, FP32+refine and FP64
        for d in d_rhs + d_x:
            ctx.free(d)
        r = surfaces_measure(ctx, cpmgpu, args.surfaces_steps, 0, args.update_delta, log=False)
        out["surfaces_batch"] = {k: r[k] for k in (
            "solves", "tubes", "rows", "largest_rows", "fp32refine_solves_per_s", "fp64_solves_per_s",
            "fp32refine_vs_fp64", "fp32refine_first_step_solves_per_s", "fp32refine_first_step_vs_fp64",
            "fallbacks_first_step", "routed_to_fp64", "iters_fp32refine_mean", "iters_fp32refine_max",
            "iters_fp64_mean", "fp64_fallbacks", "max_rel_diff_fp32refine_vs_fp64_sample", "by_h")}
        out["surfaces_batch"]["workload"] = ("8 surfaces x h {0.05,0.03,0.02,0.01} = 32 tubes x 8 spherical-harmonic "
                                             "RHS, (1-Lap)u=f, tol 1e-10, one batched solve of 256 systems per step; "
                                             f"{r['steps']} FP32+refine steps (the first tries FP32 everywhere, the "
                                             "last is reported), then one FP64 step")
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
