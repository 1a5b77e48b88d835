"""cpm_solve_device_noalloc (workspaces in the library's fixed device scratch)
against cpm_solve_device, through the C ABI: identical results on two calls
in a row, the second on a smaller operator, and a clear rejection of an
operator too large for the scratch."""
import numpy as np
import pytest

import cpmgpu
import orc

pytestmark = pytest.mark.gpu


def _dev(ctx, a):
    p = ctx.alloc(a.nbytes)
    ctx.h2d(p, a)
    return p


def _solve_both(ctx, op, b, o):
    """x from the no-alloc entry point, then from cpm_solve_device, on the same device rhs."""
    d_b = _dev(ctx, b)
    d_x1 = _dev(ctx, np.full_like(b, np.nan))
    d_x2 = _dev(ctx, np.full_like(b, np.nan))
    try:
        i1 = op.solve_device_noalloc(d_b, d_x1, o)
        i2 = op.solve_device(d_b, d_x2, o)
        x1 = ctx.d2h(np.empty_like(b), d_x1)
        x2 = ctx.d2h(np.empty_like(b), d_x2)
    finally:
        for p in (d_b, d_x1, d_x2):
            ctx.free(p)
    return x1, i1, x2, i2


@pytest.mark.parametrize("precision", [cpmgpu.CPM_FP64, cpmgpu.CPM_FP32_REFINE])
def test_noalloc_matches_solve_device(gpu_ctx, precision):
    o = cpmgpu.opts(precision=precision)
    # first call: a tube-built (tensor-product) operator of ~2.6e5 rows
    print(f"case noalloc precision={precision} torus dx=0.05 tube operator", flush=True)
    tube = cpmgpu.Tube(gpu_ctx, "Torus", 0.05)
    big = tube.operator(1.0, -1.0)
    rng = np.random.default_rng(11)
    b1 = rng.standard_normal(big.D)
    x1, i1, x2, i2 = _solve_both(gpu_ctx, big, b1, o)
    assert i1["status"] == "converged" and i1 == i2, (i1, i2)
    assert np.array_equal(x1, x2)
    # second call, smaller: an oracle-built CSR operator of ~1.3e4 rows (the
    # scratch still holds the first call's larger vectors)
    print(f"case noalloc precision={precision} sphere dx=0.1 CSR operator", flush=True)
    P = orc.Problem("Sphere", 0.1)
    L, E, I = P.L(), P.E(), P.ident()
    small = cpmgpu.Operator(gpu_ctx, L, E, I, 1.0, -1.0)
    assert small.D < big.D
    b2 = rng.standard_normal(small.D)
    y1, j1, y2, j2 = _solve_both(gpu_ctx, small, b2, o)
    assert j1["status"] == "converged" and j1 == j2, (j1, j2)
    assert np.array_equal(y1, y2)
    # and the solution is the oracle's (1e-8 relative at the 1e-10 stop)
    x_ref, _, _ = orc.solve(L, E, I, b2, 1.0, -1.0, tol=1e-13, max_it=20000)
    assert np.linalg.norm(y1 - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    small.close()
    big.close()
    tube.close()


def test_noalloc_rejects_large_operator(gpu_ctx):
    print("case noalloc rejection sphere dx=0.01 (1.0e6 rows)", flush=True)
    tube = cpmgpu.Tube(gpu_ctx, "Sphere", 0.01)
    op = tube.operator(1.0, -1.0)
    assert op.D > 800_000
    d_b = _dev(gpu_ctx, np.ones(op.D))
    d_x = _dev(gpu_ctx, np.zeros(op.D))
    try:
        with pytest.raises(cpmgpu.CpmError) as e:
            op.solve_device_noalloc(d_b, d_x)
        assert e.value.code == -1 and "no-alloc scratch" in str(e.value)
        x = gpu_ctx.d2h(np.empty(op.D), d_x)
        assert not x.any()   # nothing ran
    finally:
        gpu_ctx.free(d_b)
        gpu_ctx.free(d_x)
    op.close()
    tube.close()
