"""The C++ facade (cpm-ibc_amd/host/cpm_solver.h) used with the reference's own
calling sequence -- Init(m, n, L, E, diagonal, identity_rows); Solve(rhs) --
on matrices produced by the oracle, compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import orc

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "solver_dropin")


def _write(d, P, b, m, n):
    L, E, I = P.L(), P.E(), P.ident()
    for name, A in (("L", L), ("E", E)):
        np.asarray(A.ptr, np.int32).tofile(os.path.join(d, f"{name}_ptr.bin"))
        np.asarray(A.col, np.int32).tofile(os.path.join(d, f"{name}_col.bin"))
        np.asarray(A.val, np.float64).tofile(os.path.join(d, f"{name}_val.bin"))
    np.asarray(I, np.uint8).tofile(os.path.join(d, "ident.bin"))
    np.asarray(b, np.float64).tofile(os.path.join(d, "b.bin"))
    with open(os.path.join(d, "mn.txt"), "w") as f:
        f.write(f"{P.D} {m!r} {n!r}\n")
    return L, E, I


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "sphere+circle"])
def test_cpp_solver_dropin(tmp_path, case):
    print(f"case cpp drop-in {case}", flush=True)
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/solver_dropin is not built (run __graft_entry__.build())")
    if case == "sphere":
        P, m, n = orc.Problem("Sphere", 0.1), 1.0, -1.0
    else:
        dx = 0.1
        P = orc.IBCProblem("Sphere", dx, ibcs=[orc.IBC("Circle3D", params=[0.6, 0.0, 0.0, 0.8, 0.0, 0.0, 1.0],
                                                      boundary_order=2, on_ibc_tolerance=0.1 * dx * dx)])
        m, n = 0.0, -1.0
    b = np.random.default_rng(21).standard_normal(P.D)
    L, E, I = _write(str(tmp_path), P, b, m, n)
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    rd = lambda name: np.fromfile(os.path.join(str(tmp_path), name), np.float64)
    A = P.explicit_operator(m, n)
    x64, x32, ax, xb = rd("x_fp64.bin"), rd("x_fp32.bin"), rd("Ax.bin"), rd("x_batch.bin")
    bn = np.linalg.norm(b)
    assert np.linalg.norm(b - A.matvec(x64)) / bn <= 2e-10
    assert np.linalg.norm(b - A.matvec(x32)) / bn <= 2e-10
    ref_ax = orc.multiply(L, E, I, x64, m, n)
    assert np.abs(ax - ref_ax).max() <= 1e-12 * np.abs(ref_ax).max()
    D = P.D
    assert np.linalg.norm(xb[:D] - x64) <= 1e-12 * np.linalg.norm(x64)
    assert np.linalg.norm(xb[D:2 * D] - 2 * x64) <= 1e-9 * np.linalg.norm(x64)
    assert np.linalg.norm(xb[2 * D:] + x64) <= 1e-9 * np.linalg.norm(x64)
    assert '"threw": 1' in out.stdout
