"""GPU parity: the HIP operator and BiCGStab against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import cpmgpu
import orc

pytestmark = pytest.mark.gpu


def _problem(surface="Sphere", dx=0.1, **kw):
    P = orc.Problem(surface, dx, **kw)
    return P, P.L(), P.E(), P.ident()


@pytest.mark.parametrize("surface,dx,m,n", [("Sphere", 0.2, 1.0, -1.0), ("Sphere", 0.1, 0.0, 1.0),
                                            ("Torus", 0.2, 1.0, -0.3)])
def test_multiply_matches_oracle(gpu_ctx, surface, dx, m, n):
    print(f"case multiply {surface} dx={dx} m={m} n={n}", flush=True)
    P, L, E, I = _problem(surface, dx)
    op = cpmgpu.Operator(gpu_ctx, L, E, I, m, n)
    x = np.random.default_rng(3).standard_normal(P.D)
    y_gpu = op.multiply(x, m, n)
    y_ref = orc.multiply(L, E, I, x, m, n)
    assert np.abs(y_gpu - y_ref).max() <= 1e-12 * np.abs(y_ref).max()


def test_multiply_identity_rows(gpu_ctx):
    print("case multiply hemisphere identity rows", flush=True)
    P = orc.Problem("Hemisphere", 0.1, bdy_order=1, which_bdy=(1,))
    L, E, I = P.L(), P.E(), P.ident()
    assert I.sum() > 0
    op = cpmgpu.Operator(gpu_ctx, L, E, I, 0.0, 1.0)
    x = np.random.default_rng(4).standard_normal(P.D)
    y = op.multiply(x, 0.0, 1.0)
    assert np.array_equal(y[I == 1], x[I == 1])
    assert np.abs(y - orc.multiply(L, E, I, x, 0.0, 1.0)).max() <= 1e-12 * np.abs(y).max()


@pytest.mark.parametrize("surface,dx,m,n", [("Sphere", 0.1, 1.0, -1.0), ("Sphere", 0.2, 1.0, 1.0),
                                            ("Torus", 0.2, 1.0, -1.0)])
def test_fp64_solve_matches_oracle(gpu_ctx, surface, dx, m, n):
    """FP64 mode: both sides solved to 1e-13 agree to 1e-10 relative; at the
    reference stop (1e-10) the GPU residual obeys the same bound."""
    print(f"case fp64 solve {surface} dx={dx}", flush=True)
    P, L, E, I = _problem(surface, dx)
    b = np.random.default_rng(5).standard_normal(P.D)
    op = cpmgpu.Operator(gpu_ctx, L, E, I, m, n)
    x_ref, it_ref, _ = orc.solve(L, E, I, b, m, n, tol=1e-13, max_it=20000)
    x, info = op.solve(b, cpmgpu.opts(tol=1e-13, max_iters=20000))
    assert info["status"] == "converged"
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert rel <= 1e-10, rel
    x10, info10 = op.solve(b)
    A = P.explicit_operator(m, n)
    assert np.linalg.norm(b - A.matvec(x10)) / np.linalg.norm(b) <= 2e-10
    _, it10, _ = orc.solve(L, E, I, b, m, n, tol=1e-10)
    assert abs(info10["iters"] - it10) <= max(3, 0.1 * it10)


def test_fp32_refine_solve_accuracy(gpu_ctx):
    print("case fp32+refine sphere dx=0.1", flush=True)
    P, L, E, I = _problem("Sphere", 0.1)
    b = np.random.default_rng(6).standard_normal(P.D)
    op = cpmgpu.Operator(gpu_ctx, L, E, I, 1.0, -1.0)
    x_ref, _, _ = orc.solve(L, E, I, b, 1.0, -1.0, tol=1e-14, max_it=20000)
    x, info = op.solve(b, cpmgpu.opts(tol=1e-12, precision=cpmgpu.CPM_FP32_REFINE))
    assert info["status"] == "converged" and info["updates"] >= 1
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert rel <= 1e-8, rel        # the FP32+refine accuracy bar vs the FP64 oracle


def test_batch_equals_single(gpu_ctx):
    """Multi-RHS + multi-operator batch: every instance equals its solo solve."""
    print("case batch", flush=True)
    P1, L1, E1, I1 = _problem("Sphere", 0.2)
    P2, L2, E2, I2 = _problem("Torus", 0.25)
    op1 = cpmgpu.Operator(gpu_ctx, L1, E1, I1, 1.0, -1.0)
    op2 = cpmgpu.Operator(gpu_ctx, L2, E2, I2, 1.0, -0.5)
    rng = np.random.default_rng(7)
    ops = [op1, op2, op1, op1, op2]
    rhs = [rng.standard_normal(o.D) for o in ops]
    for prec in (cpmgpu.CPM_FP64, cpmgpu.CPM_FP32_REFINE):
        o = cpmgpu.opts(tol=1e-11, precision=prec)
        xs, infos = gpu_ctx.solve_batch(ops, rhs, o)
        for op, b, x, inf in zip(ops, rhs, xs, infos):
            xs1, inf1 = op.solve(b, o)
            assert inf["status"] == "converged"
            assert np.linalg.norm(x - xs1) <= 1e-9 * np.linalg.norm(xs1)


def test_edge_cases(gpu_ctx):
    print("case edge", flush=True)
    P, L, E, I = _problem("Sphere", 0.2)
    op = cpmgpu.Operator(gpu_ctx, L, E, I, 1.0, -1.0)
    x, info = op.solve(np.zeros(P.D))
    assert info["status"] == "zero_rhs" and not x.any()
    b = np.random.default_rng(8).standard_normal(P.D)
    x, info = op.solve(b, cpmgpu.opts(max_iters=1))
    assert info["status"] == "maxit" and info["iters"] == 1
    # one iteration from x0 = 0 matches the oracle's single iteration
    x1, it1, _ = orc.solve(L, E, I, b, 1.0, -1.0, max_it=1)
    assert np.allclose(x, x1, rtol=1e-12, atol=1e-12)
    bn = b.copy()
    bn[5] = np.nan
    x, info = op.solve(bn)
    assert info["status"] == "breakdown"


@pytest.mark.parametrize("sweeps", [1, 3])
def test_jacobi_preconditioner_matches_oracle(gpu_ctx, sweeps):
    """Damped Jacobi preconditioning (GeneralizedBiCGStabSystem.h:360-380), as the
    geodesic pipeline uses it (problems/GeodesicDistance.cpp:118-119)."""
    print(f"case jacobi sweeps={sweeps}", flush=True)
    P, L, E, I = _problem("Sphere", 0.1)
    b = np.random.default_rng(9).standard_normal(P.D)
    op = cpmgpu.Operator(gpu_ctx, L, E, I, 1.0, -1.0, precond=cpmgpu.PRECOND_JACOBI, precond_iters=sweeps)
    x_ref, it_ref, _ = orc.solve(L, E, I, b, 1.0, -1.0, precond=2, precond_iters=sweeps, tol=1e-13, max_it=20000)
    x, info = op.solve(b, cpmgpu.opts(tol=1e-13, max_iters=20000))
    assert info["status"] == "converged"
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-10
    _, it10, _ = orc.solve(L, E, I, b, 1.0, -1.0, precond=2, precond_iters=sweeps, tol=1e-10)
    x10, info10 = op.solve(b)
    assert abs(info10["iters"] - it10) <= max(3, 0.1 * it10)
    x32, i32 = op.solve(b, cpmgpu.opts(tol=1e-11, precision=cpmgpu.CPM_FP32_REFINE))
    assert i32["status"] == "converged"
    assert np.linalg.norm(x32 - x_ref) / np.linalg.norm(x_ref) <= 1e-8


@pytest.mark.parametrize("precision,nb", [("fp32refine", 1), ("fp32refine", 8), ("fp64", 1), ("fp64", 4)])
def test_batch_graph_replays_equal_eager(gpu_ctx, monkeypatch, precision, nb):
    """Batch solves whose iterations are captured in a HIP graph (4 at a time) and replayed for every chunk
    between the host's polls (CPMGPU_GRAPH=1) end bit for bit where the eager launches end, with the same
    iteration counts, two calls in a row each, FP32+refine and FP64, one and several right-hand sides."""
    print(f"case batch graph {precision} nb={nb}", flush=True)
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", 0.05)
    op = tube.operator(1.0, -1.0)
    D = op.D
    rng = np.random.default_rng(13)
    B = [rng.standard_normal(D) * (1 + k) for k in range(nb)]
    prec = cpmgpu.CPM_FP32_REFINE if precision == "fp32refine" else cpmgpu.CPM_FP64
    o = cpmgpu.opts(tol=1e-10, precision=prec)
    monkeypatch.setenv("CPMGPU_FP64_ROUTING", "0")
    res = {}
    for mode in ("0", "1", "1"):
        monkeypatch.setenv("CPMGPU_GRAPH", mode)
        xs, infos = gpu_ctx.solve_batch([op] * nb, B, o)
        assert all(i["status"] == "converged" for i in infos)
        if mode in res:
            assert [i["iters"] for i in infos] == [i["iters"] for i in res[mode][1]]
            assert all(np.array_equal(a, b) for a, b in zip(xs, res[mode][0]))
        res[mode] = (xs, infos)
    assert [i["iters"] for i in res["1"][1]] == [i["iters"] for i in res["0"][1]]
    assert all(np.array_equal(a, b) for a, b in zip(res["1"][0], res["0"][0]))


@pytest.mark.parametrize("n,lanes", [(11, 8), (3, 8), (5, 1), (37, 16), (9, 16), (45, 32)])
def test_stream_solves(gpu_ctx, n, lanes):
    """Continuous batching (cpm_solve_stream_device): every system of the stream,
    whichever lane, group of 8 lanes (16 and 32 lanes) and refill it ran in,
    matches its own FP64 solve (both at tol 1e-10) to 1e-8; a zero right-hand
    side reports zero_rhs and x = 0."""
    print(f"case stream n={n} lanes={lanes}", flush=True)
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", 0.05)
    op = tube.operator(1.0, -1.0)
    D = op.D
    rng = np.random.default_rng(7)
    B = [rng.standard_normal(D) * (1 + 10 * (k % 3)) for k in range(n)]
    if n > 3:
        B[2] = np.zeros(D)
    d_b = [gpu_ctx.alloc(8 * D) for _ in B]
    d_x = [gpu_ctx.alloc(8 * D) for _ in B]
    for d, b in zip(d_b, B):
        gpu_ctx.h2d(d, b)
    infos = gpu_ctx.solve_stream_device(op, d_b, d_x, lanes, cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP32_REFINE))
    o64 = cpmgpu.opts(tol=1e-10)
    for k, (b, d, inf) in enumerate(zip(B, d_x, infos)):
        x = gpu_ctx.d2h(np.zeros(D), d)
        if not b.any():
            assert inf["status"] == "zero_rhs" and not x.any()
            continue
        assert inf["status"] == "converged", (k, inf)
        x64, i64 = op.solve(b, o64)
        rel = np.linalg.norm(x - x64) / np.linalg.norm(x64)
        assert rel < 1e-8, (k, rel, inf)
    for d in d_b + d_x:
        gpu_ctx.free(d)


def test_sphere_closed_form_surface_error(gpu_ctx):
    """Plain CPM on the unit sphere, (1 - Laplace_S) u = f with u = z + 3xy, f = 3z + 21xy, solved on the
    GPU-built tube in FP64 and FP32+refine: the max error at 200 x 200 plot points converges at second
    order (h 0.1 -> 0.05) and both precisions agree to 1e-8 of the solution."""
    print("case sphere closed form surface error", flush=True)
    th = np.repeat(np.arange(200) * (2 * np.pi / 199), 200)
    ph = np.tile(np.arange(200) * (np.pi / 199), 200)
    xp = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    uex = xp[:, 2] + 3 * xp[:, 0] * xp[:, 1]
    errs = []
    for dx in (0.1, 0.05):
        t = cpmgpu.Tube(gpu_ctx, "Sphere", dx)
        op = t.operator(1.0, -1.0)
        cp = t.export()["cp"]
        b = 3 * cp[:, 2] + 21 * cp[:, 0] * cp[:, 1]
        x64, i64 = op.solve(b, cpmgpu.opts(precision=cpmgpu.CPM_FP64))
        x32, i32 = op.solve(b, cpmgpu.opts(precision=cpmgpu.CPM_FP32_REFINE))
        assert i64["status"] == i32["status"] == "converged"
        assert np.linalg.norm(x32 - x64) <= 1e-8 * np.linalg.norm(x64)
        Ep = t.interp(xp)
        errs.append(np.abs(Ep.matvec(x64) - uex).max())
    print(f"max surface error h=0.1 {errs[0]:.2e}, h=0.05 {errs[1]:.2e}", flush=True)
    assert errs[1] < 2e-3 and errs[0] / errs[1] > 3.2


@pytest.mark.parametrize("lanes", [8, 16])
def test_stream_graph_replays_equal_eager(gpu_ctx, monkeypatch, lanes):
    """Streamed solves with their iterations captured in a HIP graph and replayed between the host's
    refills (CPMGPU_GRAPH=1) end bit for bit where the eager launches (CPMGPU_GRAPH=0) end, with the
    same iteration counts: every replay runs the same kernels on the same tables."""
    print(f"case stream graph lanes={lanes}", flush=True)
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", 0.05)
    op = tube.operator(1.0, -1.0)
    D = op.D
    rng = np.random.default_rng(11)
    B = [rng.standard_normal(D) * (1 + k) for k in range(2 * lanes + 3)]
    d_b = [gpu_ctx.alloc(8 * D) for _ in B]
    d_x = [gpu_ctx.alloc(8 * D) for _ in B]
    for d, b in zip(d_b, B):
        gpu_ctx.h2d(d, b)
    o = cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP32_REFINE)
    monkeypatch.setenv("CPMGPU_FP64_ROUTING", "0")   # every call runs FP32 (random data: some systems fall back)
    res = {}
    for mode in ("0", "1", "1"):
        monkeypatch.setenv("CPMGPU_GRAPH", mode)
        infos = gpu_ctx.solve_stream_device(op, d_b, d_x, lanes, o)
        xs = [gpu_ctx.d2h(np.zeros(D), d) for d in d_x]
        assert all(i["status"] == "converged" for i in infos)
        if mode in res:
            assert [i["iters"] for i in infos] == [i["iters"] for i in res[mode][1]]
            assert all(np.array_equal(a, b) for a, b in zip(xs, res[mode][0]))
        res[mode] = (xs, infos)
    assert [i["iters"] for i in res["1"][1]] == [i["iters"] for i in res["0"][1]]
    assert all(np.array_equal(a, b) for a, b in zip(res["1"][0], res["0"][0]))
    for d in d_b + d_x:
        gpu_ctx.free(d)
