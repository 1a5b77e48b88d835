"""Tubes around triangle meshes (surface type "Triangulation", closest points
of Surface::cpTri, src/Surface.cpp:1751-1779) built on the GPU with its BVH,
against the oracle's brute force over all triangles."""
import numpy as np
import pytest

import cpmgpu
import orc
from cpmgpu import workloads

pytestmark = pytest.mark.gpu


def _mesh(name):
    if name == "icosphere2":
        return workloads.icosphere(2)
    if name == "displaced3":
        return workloads.displaced_icosphere(3, seed=11)
    if name == "box":
        return workloads.box_mesh(-0.8, 0.9)
    raise ValueError(name)


@pytest.mark.parametrize("builder", ["host", "gpu"])
@pytest.mark.parametrize("name,dx", [("icosphere2", 0.1), ("displaced3", 0.05), ("box", 0.1)])
def test_mesh_tube_matches_oracle(gpu_ctx, monkeypatch, name, dx, builder):
    """Both hierarchy builders (host SAH, GPU linear BVH): the traversal is exact either way."""
    print(f"case mesh tube {name} dx={dx} builder={builder}", flush=True)
    monkeypatch.setenv("CPMGPU_BVH", builder)
    V, F = _mesh(name)
    P = orc.mesh_problem(V, F, dx)
    To = P.tube()
    tg = cpmgpu.Tube(gpu_ctx, "Triangulation", dx, verts=V, faces=F)
    Tg = tg.export()
    print(f"nodes gpu={tg.N} oracle={P.N} build_ms={tg.info['build_ms']:.1f}", flush=True)
    assert tg.N == P.N and tg.info["xstart"] == P.xstart and tg.info["N1d"] == P.N1d
    assert np.array_equal(Tg["rec"], To["rec"])
    assert np.array_equal(Tg["nbr"], To["nbr"])
    assert np.array_equal(Tg["bdy"], To["bdy"])
    # same FP64 operations in the same order on both sides (no contraction)
    assert np.abs(Tg["dist"] - To["dist"]).max() <= 1e-15
    assert np.abs(Tg["cp"] - To["cp"]).max() <= 1e-15
    assert np.array_equal(Tg["bpt"], To["bpt"])
    Lg, Lo = tg.L(), P.L()
    assert np.array_equal(Lg.col, Lo.col) and np.array_equal(Lg.val, Lo.val)
    Eg, Eo = tg.E(), P.E()
    assert np.array_equal(Eg.col, Eo.col) and np.abs(Eg.val - Eo.val).max() <= 1e-15


def test_box_mesh_closed_form(gpu_ctx):
    """Cube [-1, 1]^3: nodes outside project to clip(x, -1, 1), inside nodes to the nearest face."""
    print("case mesh box closed form", flush=True)
    V, F = workloads.box_mesh()
    tg = cpmgpu.Tube(gpu_ctx, "Triangulation", 0.07, verts=V, faces=F)
    T = tg.export()
    x, cp, d = T["x"], T["cp"], T["dist"]
    out = (np.abs(x) > 1).any(1)
    c = np.clip(x, -1, 1)
    assert np.abs(cp[out] - c[out]).max() < 1e-15
    assert np.abs(d[out] - np.linalg.norm(x[out] - c[out], axis=1)).max() < 1e-15
    gap = 1 - np.abs(x[~out])
    assert np.abs(d[~out] - gap.min(1)).max() < 1e-15
    assert (d <= tg.info["tube_radius"]).all()


def test_mesh_serving_size(gpu_ctx, monkeypatch):
    """The serving workload's surface: displaced icosphere with 81,920 triangles at h = 0.01
    (about 1.1e6 tube nodes).  A sample of nodes is checked against the brute force, and the tubes
    of the GPU-built (the default at this size) and host-built hierarchies are identical."""
    print("case mesh serving size", flush=True)
    V, F = workloads.displaced_icosphere(6, seed=4)
    tg = cpmgpu.Tube(gpu_ctx, "Triangulation", 0.01, verts=V, faces=F)
    T = tg.export()
    monkeypatch.setenv("CPMGPU_BVH", "host")
    th = cpmgpu.Tube(gpu_ctx, "Triangulation", 0.01, verts=V, faces=F)
    Th = th.export()
    monkeypatch.delenv("CPMGPU_BVH")
    print(f"build_ms gpu bvh {tg.info['build_ms']:.1f} host bvh {th.info['build_ms']:.1f}", flush=True)
    for k in T:
        assert np.array_equal(T[k], Th[k]), k
    del th, Th
    print(f"nodes={tg.N} build_ms={tg.info['build_ms']:.1f}", flush=True)
    assert 0.8e6 < tg.N < 1.6e6
    assert (T["dist"] <= tg.info["tube_radius"]).all()
    idx = np.random.default_rng(1).choice(tg.N, 1500, replace=False)
    cp, d = orc.mesh_cp(V, F, T["x"][idx])
    assert np.abs(T["dist"][idx] - d).max() <= 1e-15
    assert np.abs(T["cp"][idx] - cp).max() <= 1e-15


def test_mesh_solve_matches_oracle(gpu_ctx):
    """(1 - Delta) u = f on an icosphere tube: GPU-built operator and solve vs the oracle's."""
    print("case mesh solve", flush=True)
    V, F = workloads.icosphere(3)
    dx = 0.1
    P = orc.mesh_problem(V, F, dx)
    To = P.tube()
    c = To["cp"]
    f = c[:, 2] + 3 * c[:, 0] * c[:, 1]
    uo = orc.solve(P.L(), P.E(), P.ident(), f, 1.0, -1.0, tol=1e-13)[0]
    tg = cpmgpu.Tube(gpu_ctx, "Triangulation", dx, verts=V, faces=F)
    op = tg.operator(1.0, -1.0)
    ug, info = op.solve(f, cpmgpu.opts(tol=1e-13))
    assert info["status"] == "converged"
    assert np.linalg.norm(ug - uo) / np.linalg.norm(uo) < 1e-10
