"""GPU tube construction vs the oracle: node indices, neighbours and base
points bit-exact; closest points / E / L to rounding; operator and solve parity."""
import numpy as np
import pytest

import cpmgpu
from cpmgpu import workloads
import orc

pytestmark = pytest.mark.gpu

CASES = [("Sphere", 0.2, None, None), ("Sphere", 0.1, None, None), ("Sphere", 0.05, None, None),
         ("Torus", 0.2, None, None), ("Hemisphere", 0.1, None, None),
         ("Torus", 0.02, [1.0, 0.4], (-1.4, 1.4)),            # the torus of the IBC workload, h = 0.02
         ("Sphere", 0.1, [0.7, 0.1, -0.05, 0.02], (-0.7, 0.85)),
         ("Ellipsoid", 0.1, None, None), ("Ellipsoid", 0.05, [0.6, 1.3, 0.9], (-1.3, 1.3))]


@pytest.mark.parametrize("surface,dx,params,bbox", CASES)
def test_tube_matches_oracle(gpu_ctx, surface, dx, params, bbox):
    print(f"case tube {surface} dx={dx} params={params}", flush=True)
    P = orc.Problem(surface, dx, params=params, bbox=bbox)
    To = P.tube()
    tg = cpmgpu.Tube(gpu_ctx, surface, dx, params=params, bbox=bbox)
    Tg = tg.export()
    assert tg.N == P.N
    assert tg.info["n"] == P.n and tg.info["N1d"] == P.N1d and tg.info["xstart"] == P.xstart
    assert tg.info["tube_radius"] == P.r_t
    assert np.array_equal(Tg["rec"], To["rec"])                  # identical breadth-first order
    assert np.array_equal(Tg["nbr"], To["nbr"])
    assert np.array_equal(Tg["bdy"], To["bdy"])
    assert np.array_equal(Tg["x"], To["x"])
    # cp via atan2/acos/sin/cos: ROCm ocml vs glibc differ by a few ulps
    assert np.abs(Tg["cp"] - To["cp"]).max() <= 8 * 2.2e-16 * max(1.0, np.abs(To["cp"]).max())
    assert np.abs(Tg["dist"] - To["dist"]).max() <= 1e-15
    # Base points: floor((cp - xstart) / dx) - 1.  Where cp lies within a few
    # ulps of a grid plane the floor is decided by the last bit of the
    # platform's atan2/acos/sin/cos (glibc on the CPU, ROCm's ocml on the GPU);
    # every other node must agree exactly.
    tie = _tie_nodes(To["cp"], P.xstart, dx)
    diff = np.flatnonzero(Tg["bpt"] != To["bpt"])
    print(f"nodes={P.N} base-point ties={tie.sum()} differing={len(diff)}")
    assert tie[diff].all()
    assert len(diff) <= max(8, 0.1 * tie.sum())
    if P.N < 400000:
        Lg, Lo = tg.L(), P.L()
        assert np.array_equal(Lg.ptr, Lo.ptr) and np.array_equal(Lg.col, Lo.col) and np.array_equal(Lg.val, Lo.val)
        Eg, Eo = tg.E(), P.E()
        assert np.array_equal(Eg.ptr, Eo.ptr)
        W = np.diff(Eo.ptr)[0]
        same = np.ones(P.N, bool)
        same[diff] = False
        cg, co = Eg.col.reshape(-1, W), Eo.col.reshape(-1, W)
        vg, vo = Eg.val.reshape(-1, W), Eo.val.reshape(-1, W)
        assert np.array_equal(cg[same], co[same])
        assert np.abs(vg[same] - vo[same]).max() <= 2e-13   # weights from cp that differ by ulps
        # at tie nodes both stencils interpolate smooth data identically
        x = To["x"]
        q = x[:, 0] ** 3 - 2 * x[:, 1] * x[:, 2] ** 2 + 0.5 * x[:, 1] + x[:, 2]
        assert np.abs(Eg.matvec(q) - Eo.matvec(q)).max() <= 1e-12


def _tie_nodes(cp, xstart, dx):
    s = (cp - xstart) / dx
    return (np.abs(s - np.rint(s)) <= 1e-12 * np.maximum(1.0, np.abs(s))).any(axis=1)


def test_interp_matrix_plot_points(gpu_ctx):
    print("case plot interpolation", flush=True)
    P = orc.Problem("Sphere", 0.1)
    Eo = P.plot_E()
    tg = cpmgpu.Tube(gpu_ctx, "Sphere", 0.1)
    th = np.repeat(np.arange(200) * (2 * np.pi / 199), 200)
    ph = np.tile(np.arange(200) * (np.pi / 199), 200)
    xp = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    Eg = tg.interp(xp)
    assert np.array_equal(Eg.ptr, Eo.ptr)
    W = np.diff(Eo.ptr)[0]
    same = (Eg.col.reshape(-1, W) == Eo.col.reshape(-1, W)).all(axis=1)
    assert same.mean() > 0.999
    vg, vo = Eg.val.reshape(-1, W), Eo.val.reshape(-1, W)
    assert np.abs(vg[same] - vo[same]).max() <= 1e-14
    u = np.random.default_rng(0).standard_normal(P.N)
    # rows that differ are ties of a query point on a grid plane: the same interpolant up to rounding
    assert np.abs(Eg.matvec(u)[~same] - Eo.matvec(u)[~same]).max(initial=0) <= 1e-12 * np.abs(u).max() * W


@pytest.mark.parametrize("prec", [cpmgpu.CPM_FP64, cpmgpu.CPM_FP32_REFINE])
def test_solve_on_gpu_tube_matches_oracle(gpu_ctx, prec):
    """The unit-sphere system (h=0.05, (1 - Laplace) u = 3z + 21xy) built entirely on the GPU."""
    print(f"case unit sphere solve prec={prec}", flush=True)
    P = orc.Problem("Sphere", 0.05)
    cp = P.tube()["cp"]
    b = 3 * cp[:, 2] + 21 * cp[:, 0] * cp[:, 1]
    x_ref, _, _ = orc.solve(P.L(), P.E(), P.ident(), b, 1.0, -1.0, tol=1e-14, max_it=20000)
    tg = cpmgpu.Tube(gpu_ctx, "Sphere", 0.05)
    op = tg.operator(1.0, -1.0)
    x, info = op.solve(b, cpmgpu.opts(tol=1e-13 if prec == cpmgpu.CPM_FP64 else 1e-12, precision=prec))
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert info["status"] == "converged"
    assert rel <= (1e-10 if prec == cpmgpu.CPM_FP64 else 1e-8), rel


@pytest.mark.parametrize("surface,dx,params,bbox", [("Torus", 0.1, None, None), ("Sphere", 0.05, None, None)])
def test_tensor_product_format(gpu_ctx, monkeypatch, surface, dx, params, bbox):
    """Tube operators run in the tensor-product format; its result equals the ELL
    rows it was derived from and the oracle's operator."""
    print(f"case tp format {surface} dx={dx}", flush=True)
    tg = cpmgpu.Tube(gpu_ctx, surface, dx, params=params, bbox=bbox)
    op_tp = tg.operator(1.0, -1.0)
    monkeypatch.setenv("CPMGPU_OP_FORMAT", "ell")
    op_ell = tg.operator(1.0, -1.0)
    assert (op_tp.format, op_ell.format) == ("tp", "ell")
    P = orc.Problem(surface, dx, params=params, bbox=bbox)
    x = np.random.default_rng(14).standard_normal(op_tp.D)
    y_tp, y_ell = op_tp.multiply(x, 1.0, -1.0), op_ell.multiply(x, 1.0, -1.0)
    scale = np.abs(y_ell).max()
    assert np.abs(y_tp - y_ell).max() <= 1e-15 * scale
    assert np.abs(y_tp - orc.multiply(P.L(), P.E(), P.ident(), x, 1.0, -1.0)).max() <= 1e-12 * scale
    b = np.random.default_rng(15).standard_normal(op_tp.D)
    xs, infos = gpu_ctx.solve_batch([op_tp, op_ell, op_tp], [b, b, 2 * b], cpmgpu.opts(tol=1e-11))
    assert all(i["status"] == "converged" for i in infos)
    assert np.linalg.norm(xs[0] - xs[1]) <= 1e-9 * np.linalg.norm(xs[1])
    assert np.linalg.norm(xs[2] - 2 * xs[1]) <= 1e-9 * np.linalg.norm(xs[2])


@pytest.mark.parametrize("closed", [True, False])
def test_polyline_tube_matches_oracle(gpu_ctx, closed):
    """A 2000-segment torus-knot polyline as the surface (codimension two), closed and open: the GPU's
    segment hierarchy finds the same closest points, bit for bit, as the oracle's scan over every
    segment (same formula, lowest segment index among equal distances), so the tubes are identical."""
    print(f"case polyline tube closed={closed}", flush=True)
    K = workloads.torus_knot_polyline(1.0, 0.4, 2000)
    if not closed:
        K = K[:1500]
    dx = 0.05
    bb = (min(0.0, K.min()), max(0.0, K.max()))
    P = orc.Problem("Polyline", dx, verts=K, bbox=bb)
    To = P.tube()
    tg = cpmgpu.Tube(gpu_ctx, "Polyline", dx, verts=K, bbox=bb)
    Tg = tg.export()
    print(f"nodes={P.N}", flush=True)
    assert tg.N == P.N
    for k in ("rec", "nbr", "bdy", "x", "cp", "dist", "bpt"):
        assert np.array_equal(Tg[k], To[k]), k
    if not closed:
        assert (Tg["bdy"] == 1).any() and (Tg["bdy"] == 2).any()
