"""HeatWithIBC time stepping on the GPU (cpm_heat_step_device) against the
oracle's restatement of HeatWithIBC::TimeStep on the heat study's system
(sphere with a Circle3D curve; HeatProblemsWithIBC.cpp:90-193)."""
import numpy as np
import pytest

import cpmgpu
import orc

pytestmark = pytest.mark.gpu


def _heat_ibc_params(btype=0):
    """The study's curve (HeatProblemsWithIBC.cpp:94-111): a tilted circle for Dirichlet curves,
    the great circle y = 0 for Neumann curves."""
    if btype == 1:
        return [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0]
    r, th, ph = 0.27, 0.132, 0.221
    pp = np.array([r * np.cos(ph) * np.sin(th), r * np.sin(ph) * np.sin(th), r * np.cos(th)])
    n = np.array([1.1, 1.234, 1.321])
    n /= np.linalg.norm(n)
    rho = -pp @ n
    return [np.sqrt(1 - rho * rho), *(rho * n), *n]


# (method, steps, boundary order, boundary type): the study's 2nd-order
# Dirichlet curve, a 1st-order Dirichlet curve (Dirichlet values on the ghost
# rows) and a Neumann curve.
CASES = [(0, 40, 2, 0), (1, 5, 2, 0), (2, 5, 2, 0), (0, 40, 1, 0), (1, 5, 1, 0), (2, 5, 1, 0), (0, 40, 2, 1),
         (2, 5, 2, 1)]


@pytest.mark.parametrize("method,steps,order,btype", CASES)
def test_heat_steps_match_oracle(gpu_ctx, method, steps, order, btype):
    print(f"case heat method={method} steps={steps} order={order} type={btype}", flush=True)
    dx = 0.1
    u_ref, dt = orc.heat_run(method, dx, steps, order, btype)
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", dx)
    prob = cpmgpu.Problem(tube, [cpmgpu.IBC("Circle3D", params=_heat_ibc_params(btype), boundary_order=order,
                                            boundary_type=(btype, btype), on_ibc_tolerance=0.1 * dx * dx)])
    assert prob.D == len(u_ref)
    r = 1.0 if method == 1 else 0.5
    op = prob.operator(1.0, -r * dt)
    pts, kind = prob.rhs_points()
    f = np.cos(np.arccos(pts[:, 2]))               # f = cos(phi), phi = acos(z / R)
    D = prob.D
    d_u, d_g = gpu_ctx.alloc(8 * D), gpu_ctx.alloc(8 * D)
    gpu_ctx.h2d(d_u, f)                            # the study's initial condition
    o = cpmgpu.opts(tol=1e-10)
    for k in range(steps):
        t = (k + 1) * dt
        gpu_ctx.h2d(d_g, np.where(kind == 2, np.exp(-2 * t) * f, 0.0))
        info = op.heat_step_device(method, dt, d_g, d_u, o)
        assert info["status"] == "converged"
    u = gpu_ctx.d2h(np.zeros(D), d_u)
    rel = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    print(f"heat method {method}: D={D} dt={dt:.4g} rel_diff={rel:.2e}", flush=True)
    # explicit: operator rounding only; implicit: per-step solves to 1e-10 on both sides
    assert rel <= (1e-12 if method == 0 else 1e-8)
    gpu_ctx.free(d_u)
    gpu_ctx.free(d_g)
