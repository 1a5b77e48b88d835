"""Interior-boundary problems on triangle-mesh tubes against the oracle, at sizes
the oracle finishes in seconds: the serving workload's requests (closed
plane-section polylines with 2nd-order Dirichlet colours on a displaced
icosphere, three right-hand sides) and the time-stepped batch's (Neumann
plane-section curves, Gray-Scott steps).  Subsets of every curve, Lfull /
Efull, identity rows, right-hand-side points, FP64 and FP32+refine solves and
reaction-diffusion steps through the C ABI."""
import numpy as np
import pytest

import cpmgpu
import orc
from cpmgpu import workloads

pytestmark = pytest.mark.gpu

H = 0.05


def _mesh():
    return workloads.displaced_icosphere(3, seed=4)   # 1,280 triangles, radial displacement 0.1


def _sp(A, ncols):
    import scipy.sparse as sp
    return sp.csr_matrix((A.val, A.col, A.ptr), shape=(len(A.ptr) - 1, ncols))


def _parity(Ag, Ao, q, tol, what):
    """Entrywise to tol except rows moved by a closest point within ulps of a grid plane (libm
    tie), which must still give the smooth field q identically."""
    D = len(Ao.ptr) - 1
    diff = (_sp(Ag, D) - _sp(Ao, D)).tocsr()
    rowmax = np.zeros(D)
    if diff.nnz:
        r = np.repeat(np.arange(D), np.diff(diff.indptr))
        np.maximum.at(rowmax, r, np.abs(diff.data))
    moved = np.flatnonzero(rowmax > tol)
    print(f"{what}: rows={D} rows differing beyond {tol:g}: {len(moved)}", flush=True)
    assert len(moved) <= max(16, 2e-3 * D)
    if len(moved):
        assert np.abs(diff[moved] @ q).max() <= 1e-10 * max(1.0, np.abs(q).max())


def _build(gctx, tube, V, F, curves, order, btype):
    """The same problem from the oracle and from the GPU, compared part by part."""
    kw = dict(boundary_order=order, boundary_type=btype, on_ibc_tolerance=0.1 * H * H)
    orc.set_mesh(V, F)
    P = orc.IBCProblem("Triangulation", H, ibcs=[orc.IBC("Polyline", verts=c, **kw) for c in curves])
    Q = cpmgpu.Problem(tube, [cpmgpu.IBC("Polyline", verts=c, **kw) for c in curves])
    info = Q.info
    print(f"curves={len(curves)} D={P.D} N={P.N} G={P.G} K={P.K} build_ms={info['build_ms']:.1f}", flush=True)
    assert (info["rows"], info["tube_nodes"], info["ibc_dofs"], info["dirichlet_rows"]) == (P.D, P.N, P.G, P.K)
    dofs = []
    for c in range(len(curves)):
        for which in (0, 1):
            so, co, do = P.subset(c, which)
            sg, cg, dg = Q.subset(c, which)
            assert np.array_equal(sg, so), (c, which)      # same nodes, same breadth-first order
            assert np.abs(cg - co).max() <= 1e-14
            assert np.abs(dg - do).max() <= 1e-9 * max(1e-300, np.abs(do).max())
            if which == 1:
                dofs.append(so)
    Lg, Eg, Ig = Q.export()
    Lo, Eo, Io = P.L(), P.E(), P.ident()
    assert np.array_equal(Ig, Io)
    # a smooth field on every column: tube nodes, and for ghost columns their DOF node
    x = P.tube()["x"]
    xcol = np.zeros((P.D, 3))
    xcol[:P.N] = x
    xcol[P.N:P.N + P.G] = x[np.concatenate(dofs)]
    q = 1.0 + xcol[:, 0] - 0.5 * xcol[:, 1] ** 2 + xcol[:, 0] * xcol[:, 2] + 0.25 * xcol[:, 2] ** 3
    _parity(Lg, Lo, q, 1e-12 / H ** 2, "L")
    _parity(Eg, Eo, q, 1e-12, "E")
    pg, kg = Q.rhs_points()
    po, ko = P.rhs_points()
    assert np.array_equal(kg, ko)
    assert np.abs(pg - po).max() <= 1e-14
    return P, Q


@pytest.mark.parametrize("seed", [700, 703])
def test_serving_request_matches_oracle(gpu_ctx, seed):
    """A request of the serving workload on a 1,280-triangle displaced icosphere at h = 0.05:
    1-3 closed plane-section curves, 2nd-order Dirichlet, one colour per curve and channel
    (PoissonWithIBC::GetRHS, problems/PoissonWithIBC.cpp:50-133), Laplace Solver(0, -1)."""
    print(f"case serving request seed={seed}", flush=True)
    V, F = _mesh()
    curves, cols = workloads.serving_request(V, F, seed, max_curves=3, spacing=0.5)
    assert curves
    tube = cpmgpu.Tube(gpu_ctx, "Triangulation", H, verts=V, faces=F)
    P, Q = _build(gpu_ctx, tube, V, F, curves, 2, (0, 0))
    rhs = [np.zeros(P.D) for _ in range(3)]
    for c in range(len(curves)):
        _, _, r0, nr = Q.curve_rows(c)
        for ch in range(3):
            rhs[ch][r0:r0 + nr] = cols[c, ch]
    Lo, Eo, Io = P.L(), P.E(), P.ident()
    refs = [orc.solve(Lo, Eo, Io, b, 0.0, -1.0, tol=1e-13, max_it=20000)[0] for b in rhs]
    op = Q.operator(0.0, -1.0)
    x, inf = op.solve(rhs[0], cpmgpu.opts(tol=1e-13, max_iters=20000))
    rel = np.linalg.norm(x - refs[0]) / np.linalg.norm(refs[0])
    print(f"fp64 iters={inf['iters']} rel_diff={rel:.2e}", flush=True)
    assert inf["status"] == "converged" and rel <= 1e-10
    xs, infs = gpu_ctx.solve_batch([op] * 3, rhs, cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP32_REFINE))
    for xr, r, i in zip(xs, refs, infs):
        e = np.linalg.norm(xr - r) / np.linalg.norm(r)
        print(f"fp32+refine iters={i['iters']} updates={i['updates']} fallback={i['fallback_iters']} "
              f"rel_diff={e:.2e}", flush=True)
        assert i["status"] == "converged" and e <= 1e-8


def _fields(pts, kind, seed):
    """u = 1 - pert, v = pert / 2 on tube rows (0 elsewhere), a smooth deterministic stand-in for the
    fish example's random perturbation (examples/ReactionDiffusionFish/ReactionDiffusionFish.cpp:66-71)."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    pert = 0.5 * np.exp(-(10.0 * (z - 0.1)) ** 2) + 0.25 * (1 + np.sin(5 * x + seed) * np.cos(7 * y - 2 * z))
    tube = kind == 0
    return np.where(tube, 1.0 - pert, 0.0), np.where(tube, 0.5 * pert, 0.0)


@pytest.mark.parametrize("method,steps", [(0, 10), (1, 2)])
def test_neumann_curves_gray_scott_matches_oracle(gpu_ctx, method, steps):
    """The time-stepped batch's system at a small size: 2 Neumann plane-section curves on the
    displaced icosphere, Gray-Scott (F 0.037, k 0.06, D_u 2e-5, D_v 1e-5) with the reference's
    explicit step (ReactionDiffusionWithIBC::TimeStep, problems/ReactionDiffusionWithIBC.cpp:79-104)
    or implicit diffusion with dt = 1, two fields, against the oracle's steps on its own problem."""
    print(f"case neumann curves gray-scott method={method} steps={steps}", flush=True)
    V, F = _mesh()
    curves = workloads.plane_section_curves(V, F, 2, seed=5)
    tube = cpmgpu.Tube(gpu_ctx, "Triangulation", H, verts=V, faces=F)
    P, Q = _build(gpu_ctx, tube, V, F, curves, 2, (1, 1))
    p = cpmgpu.rd_params(H, method=method, F=0.037, k=0.06, nu_u=2e-5, nu_v=1e-5)
    if method == 1:
        p.dt = 1.0
    op = Q.operator(1.0, -1.0)
    pts, kind = Q.rhs_points()
    D = Q.D
    fields = [_fields(pts, kind, s) for s in (0.0, 2.1)]
    refs = [P.rd_run(method, u0, v0, p.F, p.k, p.nu_u, p.nu_v, p.dt, steps) for u0, v0 in fields]
    d_u = [gpu_ctx.alloc(8 * D) for _ in fields]
    d_v = [gpu_ctx.alloc(8 * D) for _ in fields]
    d_g = gpu_ctx.alloc(8 * D)
    gpu_ctx.h2d(d_g, np.zeros(D))
    for (u0, v0), a, b in zip(fields, d_u, d_v):
        gpu_ctx.h2d(a, u0)
        gpu_ctx.h2d(b, v0)
    o = cpmgpu.opts(tol=1e-10)
    for _ in range(steps):
        infos = op.rd_step_device(p, d_u, d_v, d_g, d_g, o)
        assert all(i["status"] == "converged" for i in infos)
    for (ur, vr), a, b in zip(refs, d_u, d_v):
        u = gpu_ctx.d2h(np.zeros(D), a)
        v = gpu_ctx.d2h(np.zeros(D), b)
        du = np.linalg.norm(u - ur) / np.linalg.norm(ur)
        dv = np.linalg.norm(v - vr) / np.linalg.norm(vr)
        print(f"gray-scott method {method}: D={D} rel_diff u={du:.2e} v={dv:.2e}", flush=True)
        tol = 1e-12 if method == 0 else 1e-8
        assert du <= tol and dv <= tol
    for a in d_u + d_v + [d_g]:
        gpu_ctx.free(a)
