"""Serving: continuous batching across operators (cpm_server_*).  Requests of
the mesh workload's kind -- interior-boundary curves cut by planes through a
triangle-mesh surface, one colour per curve, three right-hand sides -- run
alone and in a staggered mix through a pool that cannot hold them all at
once; every result is bit-identical in both, and within 1e-8 of FP64."""
import numpy as np
import pytest

import cpmgpu
from cpmgpu import workloads

pytestmark = pytest.mark.gpu

H = 0.05


def _requests(ctx, tube, n):
    V, F = workloads.icosphere(3)
    reqs = []
    for i in range(n):
        curves, cols = workloads.serving_request(V, F, 700 + i, max_curves=3, spacing=0.85)
        P = cpmgpu.Problem(tube, [cpmgpu.IBC("Polyline", verts=c, boundary_order=2, on_ibc_tolerance=0.1 * H * H)
                                  for c in curves])
        op = P.operator(0.0, -1.0)
        rhs = [np.zeros(P.D) for _ in range(3)]
        for c in range(len(curves)):                      # each curve's Dirichlet rows take its colour
            _, _, r0, nr = P.curve_rows(c)
            for ch in range(3):
                rhs[ch][r0:r0 + nr] = cols[c, ch]
        # the same rows as GetRHS's Dirichlet kind, and each row's point lies on its curve
        pts, kind = P.rhs_points()
        idr = np.flatnonzero(kind == 2)
        own = np.concatenate([np.full(P.curve_rows(c)[3], c) for c in range(len(curves))])
        assert np.array_equal(idr, np.concatenate([np.arange(P.curve_rows(c)[2], P.curve_rows(c)[2] + P.curve_rows(c)[3])
                                                   for c in range(len(curves))]))
        assert np.array_equal(workloads.curve_of_points(curves, pts[idr]), own)
        reqs.append((P, op, rhs))
    return reqs


def _serve(ctx, reqs, pool_bytes, order, o=None):
    """Submits requests in `order` (None = run 8 iterations), retrying the ones not admitted; returns
    {index: [x per rhs]}, their infos and how often a submit was refused."""
    srv = cpmgpu.Server(ctx, pool_bytes, lanes=4, o=o)
    dev = {}
    for i, (P, op, rhs) in enumerate(reqs):
        dr = [ctx.alloc(8 * P.D) for _ in rhs]
        dx = [ctx.alloc(8 * P.D) for _ in rhs]
        for d, b in zip(dr, rhs):
            ctx.h2d(d, b)
        dev[i] = (dr, dx)
    waiting, out, infos, busy = [], {}, {}, 0
    steps = list(order)
    while steps or waiting or len(out) < len([s for s in order if s is not None]):
        if steps:
            s = steps.pop(0)
            if s is not None:
                waiting.append(s)
        while waiting:
            i = waiting[0]
            if not srv.submit(reqs[i][1], dev[i][0], dev[i][1], i):
                busy += 1
                break
            waiting.pop(0)
        done, active = srv.run(8)
        for tag, inf in done:
            out[tag] = [ctx.d2h(np.zeros(reqs[tag][0].D), d) for d in dev[tag][1]]
            infos[tag] = inf
        if not steps and not waiting and active == 0 and len(out) < len([s for s in order if s is not None]):
            raise AssertionError("requests lost")
    srv.close()
    for dr, dx in dev.values():
        for d in dr + dx:
            ctx.free(d)
    return out, infos, busy


def test_results_same_alone_and_in_the_mix(gpu_ctx, monkeypatch):
    print("case serving mix", flush=True)
    # every request tries FP32 first: routing by the family's record would make a request's
    # precision depend on how the ones before it fared
    monkeypatch.setenv("CPMGPU_FP64_ROUTING", "0")
    V, F = workloads.icosphere(3)
    tube = cpmgpu.Tube(gpu_ctx, "Triangulation", H, verts=V, faces=F)
    reqs = _requests(gpu_ctx, tube, 4)
    per_req = max(((P.D * 4 + 63) // 64) * 64 for P, _, _ in reqs) * 80
    alone = {}
    for i in range(len(reqs)):
        x, inf, _ = _serve(gpu_ctx, reqs, per_req + 1024, [i])
        assert all(j["status"] == "converged" for j in inf[i]), inf[i]
        alone[i] = x[i]
    # pool for two of them: the later requests wait for memory while others run
    mix, infos, busy = _serve(gpu_ctx, reqs, int(2.3 * per_req), [0, 1, None, 2, 3, None, None])
    print(f"refused submits while the pool was full: {busy}", flush=True)
    assert busy >= 1
    for i in range(len(reqs)):
        for a, b in zip(alone[i], mix[i]):
            assert np.array_equal(a, b)                    # bit-identical alone and in the mix
    # accuracy against plain FP64 solves of the same systems
    for i, (P, op, rhs) in enumerate(reqs):
        xs, _ = gpu_ctx.solve_batch([op] * 3, rhs, cpmgpu.opts(tol=1e-12, precision=cpmgpu.CPM_FP64))
        for a, r in zip(mix[i], xs):
            assert np.linalg.norm(a - r) <= 1e-8 * np.linalg.norm(r)


def test_fp64_server_equals_batch_solves(gpu_ctx):
    """The FP64 server (opts precision CPM_FP64): each request's results are bit-identical alone, in a
    mix through a pool that holds two of four requests, and to a plain FP64 batch solve of the same
    operator with the same right-hand sides (the same kernels on the same segment), with the same
    iteration counts."""
    print("case fp64 serving", flush=True)
    V, F = workloads.icosphere(3)
    tube = cpmgpu.Tube(gpu_ctx, "Triangulation", H, verts=V, faces=F)
    reqs = _requests(gpu_ctx, tube, 4)
    o64 = cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP64)
    per_req = max(((P.D * 4 + 63) // 64) * 64 for P, _, _ in reqs) * 88
    alone = {}
    for i in range(len(reqs)):
        x, inf, _ = _serve(gpu_ctx, reqs, per_req + 1024, [i], o64)
        assert all(j["status"] == "converged" for j in inf[i]), inf[i]
        alone[i] = (x[i], inf[i])
    mix, infos, busy = _serve(gpu_ctx, reqs, int(2.3 * per_req), [0, 1, None, 2, 3, None, None], o64)
    print(f"refused submits while the pool was full: {busy}", flush=True)
    assert busy >= 1
    for i, (P, op, rhs) in enumerate(reqs):
        xs, bi = gpu_ctx.solve_batch([op] * 3, rhs, o64)
        assert [j["iters"] for j in infos[i]] == [j["iters"] for j in bi] == [j["iters"] for j in alone[i][1]]
        for a, m, r in zip(alone[i][0], mix[i], xs):
            assert np.array_equal(a, m) and np.array_equal(a, r)


def test_submit_rejections(gpu_ctx):
    print("case serving rejections", flush=True)
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", 0.2)
    op = tube.operator(1.0, -1.0)
    D = op.D
    srv = cpmgpu.Server(gpu_ctx, 80 * 4 * ((D + 63) // 64) * 64, lanes=2)
    d = [gpu_ctx.alloc(8 * D) for _ in range(6)]
    try:
        with pytest.raises(cpmgpu.CpmError, match="right-hand sides"):
            srv.submit(op, d[:3], d[3:6], 1)              # more right-hand sides than lanes
        jop = tube.operator(1.0, -1.0, precond=cpmgpu.PRECOND_JACOBI, precond_iters=2)
        with pytest.raises(cpmgpu.CpmError, match="preconditioner"):
            srv.submit(jop, d[:1], d[1:2], 2)
        assert srv.submit(op, d[:1], d[1:2], 3)
        assert srv.submit(op, d[2:3], d[3:4], 4)
        assert not srv.submit(op, d[4:5], d[5:6], 5)      # pool full: not admitted
    finally:
        srv.close()
        for p in d:
            gpu_ctx.free(p)


def test_problem_built_on_second_context(gpu_ctx):
    """A problem built on another context of the device (a build thread's) gives the same operator and
    solutions, solved or served on the first context."""
    print("case serving build context", flush=True)
    V, F = workloads.icosphere(3)
    tube = cpmgpu.Tube(gpu_ctx, "Triangulation", H, verts=V, faces=F)
    curves, cols = workloads.serving_request(V, F, 901, max_curves=2, spacing=0.85)
    ibcs = [cpmgpu.IBC("Polyline", verts=c, boundary_order=2, on_ibc_tolerance=0.1 * H * H) for c in curves]
    bctx = cpmgpu.Context(0)
    try:
        P1 = cpmgpu.Problem(tube, ibcs)
        P2 = cpmgpu.Problem(tube, ibcs, ctx=bctx)
        (L1, E1, I1), (L2, E2, I2) = P1.export(), P2.export()
        for u, v in ((L1, L2), (E1, E2)):
            assert np.array_equal(u.ptr, v.ptr) and np.array_equal(u.col, v.col) and np.array_equal(u.val, v.val)
        assert np.array_equal(I1, I2)
        op1, op2 = P1.operator(0.0, -1.0), P2.operator(0.0, -1.0)
        b = np.zeros(P1.D)
        _, _, r0, nr = P1.curve_rows(0)
        b[r0:r0 + nr] = 1.0
        o = cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE)
        x1, i1 = gpu_ctx.solve_batch([op1], [b], o)
        x2, i2 = gpu_ctx.solve_batch([op2], [b], o)      # operator of the other context, solved here
        assert i1[0]["status"] == i2[0]["status"] == "converged"
        assert np.array_equal(x1[0], x2[0])
        op1.close(), op2.close(), P1.close(), P2.close()
    finally:
        bctx.close()
