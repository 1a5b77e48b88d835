"""The C++ facade (cpm-ibc_amd/host/cpm_solver.h) used with the reference's own
calling sequence -- Init(m, n, L, E, diagonal, identity_rows); Solve(rhs) --
on matrices produced by the oracle, compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import cpmgpu
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


PROBLEMS_EXE = os.path.join(HERE, "cpp", "problems_dropin")


@pytest.mark.gpu
def test_cpp_problem_classes(tmp_path):
    """The reference's problem classes through the C++ facade (tests/cpp/problems_dropin.cpp) against
    the oracle: Poisson + Solver::Init(m, n, GetLfull(), GetEfull(), ...), PoissonWithIBC::GetRHS +
    Solve + GetPlotInterpMatrix, HeatWithIBC::TimeStep (backward Euler), ReactionDiffusionWithIBC::
    TimeStep (explicit), GeodesicDistance::ComputeDistance and IBCParameter."""
    import json
    print("case cpp problem classes", flush=True)
    if not os.path.exists(PROBLEMS_EXE):
        pytest.fail("tests/cpp/problems_dropin is not built (run __graft_entry__.build())")
    out = subprocess.run([PROBLEMS_EXE, str(tmp_path)], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    res = json.loads(out.stdout.strip().splitlines()[-1])
    rd = lambda name: np.fromfile(os.path.join(str(tmp_path), name), np.float64)
    dx = 0.1
    # Poisson: right-hand side at the closest points, solution residual under the oracle's operator
    P = orc.Problem("Sphere", dx)
    cp = P.tube()["cp"]
    b = 3 * cp[:, 2] + 21 * cp[:, 0] * cp[:, 1]
    assert np.abs(rd("poisson_rhs.bin") - b).max() <= 1e-13
    u = rd("poisson_u.bin")
    u_ref, _, _ = orc.solve(P.L(), P.E(), P.ident(), b, 1.0, -1.0, tol=1e-13, max_it=20000)
    assert res["poisson_status"] == 1 and np.linalg.norm(u - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
    assert res["poisson_err"] < 3e-3               # second-order CPM error at h = 0.1
    # PoissonWithIBC (sphere + circle, study problem 2 with g = 0)
    r1 = [0.0, 0.25, 0.55]
    n = np.array(r1) / np.linalg.norm(r1)
    pp = 0.2 * np.array([np.cos(0.2) * np.sin(0.1), np.sin(0.2) * np.sin(0.1), np.cos(0.1)])
    rho = -pp @ n
    circle = [np.sqrt(1 - rho * rho), *(rho * n), *n]
    Q = orc.IBCProblem("Sphere", dx, ibcs=[orc.IBC("Circle3D", params=circle, boundary_order=2,
                                                   on_ibc_tolerance=0.1 * dx * dx)])

    def harmonic(x):
        th, ph = np.arctan2(x[:, 1], x[:, 0]), np.arccos(x[:, 2])
        return np.cos(3 * th) * np.sin(ph) ** 3 * (9 * np.cos(ph) ** 2 - 1)
    pts, kind = Q.rhs_points()
    b2 = np.where(kind == 2, harmonic(pts), np.where(kind <= 1, 30 * harmonic(pts), 0.0))
    assert np.abs(rd("ibc_rhs.bin") - b2).max() <= 1e-12
    x_ref, _, _ = orc.solve(Q.L(), Q.E(), Q.ident(), b2, 0.0, -1.0, tol=1e-13, max_it=20000)
    u2 = rd("ibc_u.bin")
    assert res["ibc_status"] == 1 and np.linalg.norm(u2 - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    th = np.repeat(np.arange(200) * (2 * np.pi / 199), 200)
    ph = np.tile(np.arange(200) * (np.pi / 199), 200)
    xp = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    up_ref = Q.plot_matrix(xp).matvec(x_ref)
    assert np.abs(rd("ibc_uplot.bin") - up_ref).max() <= 1e-7
    assert abs(res["ibc_err"] - np.abs(up_ref - harmonic(xp)).max()) <= 1e-7
    # IBCParameter at the Dirichlet points
    par_ref = Q.ibc_parameter(0, pts[kind == 2])
    assert np.abs(rd("ibc_param.bin") - par_ref).max() <= 1e-12 and abs(res["param0"] - par_ref[0]) <= 1e-12
    # HeatWithIBC: 5 backward-Euler steps of the study's system
    uh_ref, _ = orc.heat_run(1, dx, 5, 2, 0)
    uh = rd("heat_u.bin")
    assert res["heat_status"] == 1 and np.linalg.norm(uh - uh_ref) <= 1e-8 * np.linalg.norm(uh_ref)
    # ReactionDiffusionWithIBC: 20 explicit steps with the reference's constants
    p = cpmgpu.rd_params(dx, method=0)
    R = orc.IBCProblem("Sphere", dx, ibcs=[orc.IBC("Circle3D", params=_heat_circle(), boundary_order=2,
                                                   on_ibc_tolerance=0.1 * dx * dx)])
    rp, rk = R.rhs_points()
    x, y, z = rp[:, 0], rp[:, 1], rp[:, 2]
    pert = 0.5 * np.exp(-(10.0 * (z - 0.1)) ** 2) + 0.25 * (1 + np.sin(5 * x) * np.cos(7 * y - 2 * z))
    u0, v0 = np.where(rk == 0, 1.0 - pert, 0.0), np.where(rk == 0, 0.5 * pert, 0.0)
    ur_ref, vr_ref = R.rd_run(0, u0, v0, p.F, p.k, p.nu_u, p.nu_v, p.dt, 20)
    assert np.linalg.norm(rd("rd_u.bin") - ur_ref) <= 1e-12 * np.linalg.norm(ur_ref)
    assert np.linalg.norm(rd("rd_v.bin") - vr_ref) <= 1e-12 * np.linalg.norm(vr_ref)
    # GeodesicDistance from the north pole
    G = orc.IBCProblem("Sphere", dx, ibcs=[orc.IBC("Point", params=[0.0, 0.0, 1.0], boundary_order=1,
                                                   on_ibc_tolerance=0.1 * dx * dx)])
    phi_ref, _ = G.geodesic(dx * dx, 2000)
    phi = rd("geo_phi.bin")
    assert res["geo_status"] == 1 and np.linalg.norm(phi - phi_ref) <= 1e-7 * np.linalg.norm(phi_ref)
    assert res["geo_err"] < 0.1                     # the heat method's error at h = 0.1 (acos(z) exact)


def _heat_circle():
    r, th, ph = 0.27, 0.132, 0.221
    pp = np.array([r * np.cos(ph) * np.sin(th), r * np.sin(ph) * np.sin(th), r * np.cos(th)])
    n = np.array([1.1, 1.234, 1.321])
    n /= np.linalg.norm(n)
    rho = -pp @ n
    return [np.sqrt(1 - rho * rho), *(rho * n), *n]
