"""Many tubes in one batched solve: the 8 surfaces of the multi-tube batch
(spheres, tori, ellipsoids) at two grid spacings, 8 spherical-harmonic
right-hand sides each, (1 - Laplace_S) u = f solved with FP32 Krylov + FP64
refinement in one cpm_solve_batch call (Solver(1, -1), src/Solver.cpp:8-114,
once per instance in the reference)."""
import numpy as np
import pytest

import cpmgpu
import orc
from cpmgpu import workloads

pytestmark = pytest.mark.gpu

LMAX = 8


def test_multi_surface_batch(gpu_ctx):
    hs = (0.05, 0.03)
    surf = workloads.multi_surface_set()
    ops, rhs, meta = [], [], []
    keep = []
    for h in hs:
        for js, (label, typ, par, bb) in enumerate(surf):
            t = cpmgpu.Tube(gpu_ctx, typ, h, params=par, bbox=bb)
            op = t.operator(1.0, -1.0)
            cp = t.export()["cp"]
            # the same coefficients at both spacings, so errors compare across h
            A = np.stack([workloads.sh_coefficients(8 * js + k, LMAX) for k in range(8)], 1)
            F = workloads.sh_rhs(cp, A, LMAX)
            keep.append((t, op))
            for k in range(8):
                ops.append(op)
                rhs.append(np.ascontiguousarray(F[:, k]))
                meta.append((label, typ, par, bb, h, cp, A[:, k], t))
            print(f"tube {label} h={h}: {t.N} nodes", flush=True)
    print(f"case multi-surface batch: {len(ops)} solves, {sum(o.D for o in ops[::8])} rows", flush=True)
    x32, i32 = gpu_ctx.solve_batch(ops, rhs, cpmgpu.opts(tol=1e-10, precision=cpmgpu.CPM_FP32_REFINE))
    x64, i64 = gpu_ctx.solve_batch(ops, rhs, cpmgpu.opts(tol=1e-12, precision=cpmgpu.CPM_FP64))
    assert all(i["status"] == "converged" for i in i32 + i64)
    worst = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(x32, x64))
    print(f"worst FP32+refine vs FP64 relative difference {worst:.2e}", flush=True)
    assert worst <= 1e-8

    # the oracle on two of the small instances (same node order as the GPU tube)
    for q in (0, 5 * 8 + 3):
        label, typ, par, bb, h, cp, a, _ = meta[q]
        P = orc.Problem(typ, h, params=par, bbox=bb)
        cpo = P.tube()["cp"]
        # closest points agree to a few ulps (platform atan2/acos/sin/cos)
        assert P.D == len(rhs[q]) and np.abs(cpo - cp).max() <= 1e-14
        bo = workloads.sh_rhs(cpo, a, LMAX)[:, 0]
        xo, _, _ = orc.solve(P.L(), P.E(), P.ident(), bo, 1.0, -1.0, tol=1e-12)
        rel = np.linalg.norm(x32[q] - xo) / np.linalg.norm(xo)
        print(f"{label} h={h}: vs oracle {rel:.2e}", flush=True)
        assert rel <= 1e-8

    # spheres: u = sum a_lm Y_lm / (1 + l (l + 1) / r^2) exactly.  The
    # surface values of the closest-point solution, E u at the nodes' closest
    # points (rows at the tube's edge, whose Laplacian stencil is cut, hold no
    # surface values), converge to it at second order
    ls = np.concatenate([[l] * (2 * l + 1) for l in range(LMAX + 1)])
    err = {}
    for q, (label, typ, par, bb, h, cp, a, t) in enumerate(meta):
        if typ != "Sphere" or q % 8:
            continue
        r = par[0]
        u = workloads.sh_rhs(cp, a / (1 + ls * (ls + 1) / r ** 2), LMAX)[:, 0]
        err[(label, h)] = np.abs(t.E().matvec(x64[q]) - u).max() / np.abs(u).max()
    for label in {k[0] for k in err}:
        e5, e3 = err[(label, 0.05)], err[(label, 0.03)]
        print(f"{label}: max error h=0.05 {e5:.2e}, h=0.03 {e3:.2e}, ratio {e5 / e3:.2f}", flush=True)
        assert e5 < 3e-2 and e3 < e5 / 2.5     # second order: (0.05 / 0.03)^2 = 2.78
