"""Gray-Scott reaction-diffusion stepping on the GPU (cpm_rd_step_device)
against the oracle's restatement of ReactionDiffusionWithIBC::TimeStep
(problems/ReactionDiffusionWithIBC.cpp:79-104) and of the implicit-diffusion
variant, on a sphere with a Circle3D curve (2nd-order Dirichlet, 1st-order
Dirichlet, Neumann), Dirichlet values 0 as in the fish example."""
import numpy as np
import pytest

import cpmgpu
import orc

pytestmark = pytest.mark.gpu


def _circle(btype=0):
    """The heat study's Circle3D curve (HeatProblemsWithIBC.cpp:94-111), which the oracle builds:
    tilted for Dirichlet curves, the great circle y = 0 for Neumann curves."""
    if btype == 1:
        return [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0]
    r, th, ph = 0.27, 0.132, 0.221
    pp = np.array([r * np.cos(ph) * np.sin(th), r * np.sin(ph) * np.sin(th), r * np.cos(th)])
    n = np.array([1.1, 1.234, 1.321])
    n /= np.linalg.norm(n)
    rho = -pp @ n
    return [np.sqrt(1 - rho * rho), *(rho * n), *n]


def _initial(pts, kind, seed):
    """u = 1 - pert, v = pert / 2 on tube rows (0 on ghost and Dirichlet rows), a smooth
    deterministic stand-in for the example's rand() perturbation (ReactionDiffusionFish.cpp:66-71)."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    pert = 0.5 * np.exp(-(10.0 * (z - 0.1)) ** 2) + 0.25 * (1 + np.sin(5 * x + seed) * np.cos(7 * y - 2 * z))
    tube = kind == 0
    return np.where(tube, 1.0 - pert, 0.0), np.where(tube, 0.5 * pert, 0.0)


@pytest.mark.parametrize("method,steps,order,btype", [(0, 20, 2, 0), (0, 20, 1, 0), (0, 20, 2, 1), (1, 3, 2, 0),
                                                      (1, 3, 2, 1)])
def test_rd_steps_match_oracle(gpu_ctx, method, steps, order, btype):
    print(f"case rd method={method} steps={steps} order={order} type={btype}", flush=True)
    dx = 0.1
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", dx)
    prob = cpmgpu.Problem(tube, [cpmgpu.IBC("Circle3D", params=_circle(btype), boundary_order=order,
                                            boundary_type=(btype, btype), on_ibc_tolerance=0.1 * dx * dx)])
    p = cpmgpu.rd_params(dx, method=method)
    if method == 1:
        p.dt = 5.0                                    # beyond the explicit limit
    op = prob.operator(1.0, -1.0)
    pts, kind = prob.rhs_points()
    D = prob.D
    fields = [_initial(pts, kind, s) for s in (0.0, 1.3)]
    refs = [orc.rd_run(method, dx, u0, v0, p.F, p.k, p.nu_u, p.nu_v, p.dt, steps, order, btype) for u0, v0 in fields]
    assert len(refs[0][0]) == D
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
        assert len(infos) == 2 * len(fields)
        assert all(i["status"] == "converged" for i in infos)
    for (ur, vr), a, b in zip(refs, d_u, d_v):
        u = gpu_ctx.d2h(np.zeros(D), a)
        v = gpu_ctx.d2h(np.zeros(D), b)
        du = np.linalg.norm(u - ur) / np.linalg.norm(ur)
        dv = np.linalg.norm(v - vr) / np.linalg.norm(vr)
        print(f"rd method {method}: D={D} rel_diff u={du:.2e} v={dv:.2e}", flush=True)
        # explicit: operator rounding only; implicit: per-step solves to 1e-10 on both sides
        tol = 1e-12 if method == 0 else 1e-8
        assert du <= tol and dv <= tol
    for a in d_u + d_v + [d_g]:
        gpu_ctx.free(a)
