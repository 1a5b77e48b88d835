"""Geodesic distance by the heat method on the GPU (cpm_geodesic_distance,
problems/GeodesicDistance.cpp:11-67) against the oracle's restatement."""
import numpy as np
import pytest

import cpmgpu
import orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("surface,dx,ibc", [
    ("Sphere", 0.1, cpmgpu.IBC("Point", params=[0.0, 0.0, 1.0], boundary_order=1)),
    ("Torus", 0.1, cpmgpu.IBC("Circle3D", params=[3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0], boundary_order=1)),
])
def test_geodesic_matches_oracle(gpu_ctx, parallel_oracle, surface, dx, ibc):
    print(f"case geodesic {surface} dx={dx} {ibc.type}", flush=True)
    ibc.on_ibc_tolerance = 0.1 * dx * dx
    P = orc.IBCProblem(surface, dx, ibcs=[orc.IBC(ibc.type, params=ibc.params, boundary_order=1,
                                                  on_ibc_tolerance=ibc.on_ibc_tolerance)])
    phi_o, it_o = P.geodesic(dx * dx, 2000)
    tube = cpmgpu.Tube(gpu_ctx, surface, dx)
    prob = cpmgpu.Problem(tube, [ibc])
    assert prob.D == P.D
    phi_g, infos = prob.geodesic(dx * dx, 2000, cpmgpu.opts(tol=1e-10))
    assert all(i["status"] == "converged" for i in infos)
    rel = np.linalg.norm(phi_g - phi_o) / np.linalg.norm(phi_o)
    print(f"D={P.D} iters gpu={[i['iters'] for i in infos]} oracle={list(it_o)} rel_diff={rel:.2e}", flush=True)
    assert phi_g.min() == 0.0
    # four chained solves to 1e-10, with a normalisation of the gradient in between
    assert rel < 1e-7
