"""FP32+refine on a fine grid: the screened sphere system at h = 0.01 (1.04e6
rows, condition ~1e5) is one where FP32 BiCGStab steps stop reducing the
true FP64 residual.  Such systems are re-solved in FP64 from zero, the
reference algorithm (src/BiCGStab.h:30-108), so the result equals a plain
FP64 solve; systems that refine normally are not touched."""
import numpy as np
import pytest

import cpmgpu
from cpmgpu import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fine_sphere(gpu_ctx):
    t = cpmgpu.Tube(gpu_ctx, "Sphere", 0.01, params=[1.0], bbox=(-1.0, 1.0))
    op = t.operator(1.0, -1.0)
    cp = t.export()["cp"]
    A = np.stack([workloads.sh_coefficients(k) for k in range(3)], 1)
    F = workloads.sh_rhs(cp, A)
    return t, op, [np.ascontiguousarray(F[:, k]) for k in range(3)]


def test_fallback_equals_fp64(gpu_ctx, fine_sphere):
    print("case fallback sphere h=0.01", flush=True)
    _, op, rhs = fine_sphere
    x64, i64 = op.solve(rhs[0], cpmgpu.opts(precision=cpmgpu.CPM_FP64))
    x32, i32 = op.solve(rhs[0], cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE))
    print(i64, i32, flush=True)
    assert i64["status"] == i32["status"] == "converged"
    assert i32["fallback_iters"] == i64["iters"] > 0
    assert np.array_equal(x32, x64)                  # the same FP64 solve from zero
    assert i32["rel_residual"] == i64["rel_residual"]


def test_fallback_in_batches_and_streams(gpu_ctx, fine_sphere):
    print("case fallback batch + stream sphere h=0.01", flush=True)
    _, op, rhs = fine_sphere
    o32 = cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE)
    ref = [op.solve(b, cpmgpu.opts(precision=cpmgpu.CPM_FP64))[0] for b in rhs]
    xs, infos = gpu_ctx.solve_batch([op] * 3, rhs, o32)
    for x, r, inf in zip(xs, ref, infos):
        assert inf["status"] == "converged"
        if inf["fallback_iters"]:
            assert np.linalg.norm(x - r) <= 1e-9 * np.linalg.norm(r)
        else:
            assert np.linalg.norm(x - r) <= 1e-8 * np.linalg.norm(r)
    assert any(i["fallback_iters"] for i in infos)
    D = op.D
    d_rhs = [gpu_ctx.alloc(8 * D) for _ in rhs]
    d_x = [gpu_ctx.alloc(8 * D) for _ in rhs]
    try:
        for d, b in zip(d_rhs, rhs):
            gpu_ctx.h2d(d, b)
        infos = gpu_ctx.solve_stream_device(op, d_rhs, d_x, 2, o32)
        for d, r, inf in zip(d_x, ref, infos):
            x = gpu_ctx.d2h(np.zeros(D), d)
            assert inf["status"] == "converged"
            assert np.linalg.norm(x - r) <= 1e-8 * np.linalg.norm(r)
        assert any(i["fallback_iters"] for i in infos)
    finally:
        for d in d_rhs + d_x:
            gpu_ctx.free(d)


def test_no_fallback_where_refinement_works(gpu_ctx):
    print("case no fallback sphere h=0.05", flush=True)
    t = cpmgpu.Tube(gpu_ctx, "Sphere", 0.05)
    op = t.operator(1.0, -1.0)
    b = workloads.sh_rhs(t.export()["cp"], workloads.sh_coefficients(0)[:, None])[:, 0]
    x32, i32 = op.solve(b, cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE))
    x64, _ = op.solve(b, cpmgpu.opts(tol=1e-13, precision=cpmgpu.CPM_FP64))
    assert i32["status"] == "converged" and i32["fallback_iters"] == 0 and i32["updates"] >= 1
    assert np.linalg.norm(x32 - x64) <= 1e-8 * np.linalg.norm(x64)


def test_fallback_in_server(gpu_ctx, fine_sphere, monkeypatch):
    """Two requests that stall in FP32 finish in the same serving round: each is re-solved in FP64
    (FP32 first: the operator's record would route them to FP64 at submission)."""
    print("case fallback server sphere h=0.01", flush=True)
    monkeypatch.setenv("CPMGPU_FP64_ROUTING", "0")
    _, op, rhs = fine_sphere
    D = op.D
    ref = [op.solve(b, cpmgpu.opts(precision=cpmgpu.CPM_FP64))[0] for b in rhs[:2]]
    srv = cpmgpu.Server(gpu_ctx, 80 * 2 * 2 * ((D + 63) // 64 + 1) * 64, lanes=2)
    d_rhs = [gpu_ctx.alloc(8 * D) for _ in range(2)]
    d_x = [gpu_ctx.alloc(8 * D) for _ in range(2)]
    try:
        for d, b in zip(d_rhs, rhs[:2]):
            gpu_ctx.h2d(d, b)
        assert srv.submit(op, d_rhs[:1], d_x[:1], 0) and srv.submit(op, d_rhs[1:], d_x[1:], 1)
        got = {}
        for _ in range(2000):
            done, active = srv.run(16)
            for tag, inf in done:
                got[tag] = inf
            if len(got) == 2:
                break
        assert len(got) == 2
        for tag in (0, 1):
            assert got[tag][0]["status"] == "converged"
            x = gpu_ctx.d2h(np.zeros(D), d_x[tag])
            assert np.linalg.norm(x - ref[tag]) <= 1e-8 * np.linalg.norm(ref[tag])
        assert any(got[t][0]["fallback_iters"] for t in (0, 1))
    finally:
        srv.close()
        for d in d_rhs + d_x:
            gpu_ctx.free(d)


def test_fp64_routing_after_fallback(gpu_ctx, fine_sphere, monkeypatch):
    """Once an operator's FP32+refine solve has fallen back to FP64, later FP32+refine solves of it
    start in FP64 (no FP32 iterations; the result is the plain FP64 solve); CPMGPU_FP64_ROUTING=0
    tries FP32 first again."""
    print("case fp64 routing sphere h=0.01", flush=True)
    t, _, rhs = fine_sphere
    op = t.operator(1.0, -1.0)                       # a fresh operator: no fallback recorded
    o32 = cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE)
    x64, i64 = op.solve(rhs[1], cpmgpu.opts(precision=cpmgpu.CPM_FP64))
    monkeypatch.setenv("CPMGPU_FP64_ROUTING", "0")   # FP32 first (its family's record would route it)
    _, first = op.solve(rhs[1], o32)
    print("first", first, flush=True)
    assert first["iters"] > 0 and first["fallback_iters"] == i64["iters"]
    monkeypatch.delenv("CPMGPU_FP64_ROUTING")
    x2, second = op.solve(rhs[1], o32)
    print("second", second, flush=True)
    assert second["status"] == "converged" and second["iters"] == 0 and second["fallback_iters"] == i64["iters"]
    assert np.array_equal(x2, x64)
    xs, infos = gpu_ctx.solve_batch([op] * 2, rhs[:2], o32)
    assert all(i["iters"] == 0 and i["fallback_iters"] > 0 and i["status"] == "converged" for i in infos)
    monkeypatch.setenv("CPMGPU_FP64_ROUTING", "0")
    _, third = op.solve(rhs[1], o32)
    assert third["iters"] > 0 and third["fallback_iters"] == i64["iters"]


def test_fp64_routing_by_family(gpu_ctx, fine_sphere):
    """A fresh tensor-product operator has no record of its own: its FP32+refine solves follow its
    family's (the same m, n and L on the same grid spacing: the fine sphere's systems, which fell
    back above) and start in FP64, in solves and in a server (solved at submission, no FP32
    iterations, the plain FP64 result); an operator of another equation (m = 2) tries FP32 first."""
    print("case fp64 routing by family sphere h=0.01", flush=True)
    t, op, rhs = fine_sphere
    o32 = cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE)
    o64 = cpmgpu.opts(precision=cpmgpu.CPM_FP64)
    _, first = op.solve(rhs[0], o32)   # the family's record: this operator's fallbacks (or its routing)
    assert first["fallback_iters"] > 0
    x64, i64 = op.solve(rhs[2], o64)
    fresh = t.operator(1.0, -1.0)
    x, inf = fresh.solve(rhs[2], o32)
    print("fresh", inf, flush=True)
    assert inf["status"] == "converged" and inf["iters"] == 0 and inf["fallback_iters"] == i64["iters"]
    assert np.array_equal(x, x64)
    D = op.D
    served = t.operator(1.0, -1.0)
    srv = cpmgpu.Server(gpu_ctx, 80 * 2 * 2 * ((D + 63) // 64 + 1) * 64, lanes=2)
    d_b, d_x = gpu_ctx.alloc(8 * D), gpu_ctx.alloc(8 * D)
    try:
        gpu_ctx.h2d(d_b, rhs[2])
        assert srv.submit(served, [d_b], [d_x], 7)
        done, active = srv.run(1)
        print("served", done, active, flush=True)
        assert active == 0 and len(done) == 1 and done[0][0] == 7
        s = done[0][1][0]
        assert s["status"] == "converged" and s["iters"] == 0 and s["fallback_iters"] == i64["iters"]
        assert np.array_equal(gpu_ctx.d2h(np.zeros(D), d_x), x64)
    finally:
        srv.close()
        gpu_ctx.free(d_b)
        gpu_ctx.free(d_x)
    other = t.operator(2.0, -1.0)
    _, oinf = other.solve(rhs[2], o32)
    print("other family", oinf, flush=True)
    assert oinf["status"] == "converged" and oinf["iters"] > 0
