"""Pins of the CPU oracle against what fixes it independently of its own code:
the reference's convergence-order known answers, closed forms, brute force and
exact polynomial reproduction.  (CPU only.)"""
import math
import os

import numpy as np
import pytest

import orc


def test_tube_radius_and_grid_closed_forms():
    # Tube.cpp:136 with m_order = 2: r_t = 1.0001 sqrt((dim-1)((p+1)/2)^2 + ((p+3)/2)^2) h
    P = orc.Problem("Sphere", 0.05)
    assert P.rfac == pytest.approx(1.0001 * math.sqrt(17.0), rel=0, abs=1e-15)
    # Tube.cpp:148-157: bbox [-1,1], h=0.05 -> 40 spacings, n = ceil(log2(50)) = 6,
    # N1d = 65, spacing_diff = ceil((64-40-10)/2) = 7, xstart = -1 - 12 h = -1.6
    assert (P.n, P.N1d) == (6, 65)
    assert P.xstart == pytest.approx(-1.6, abs=1e-15)
    P2 = orc.Problem("Circle", 0.05, params=[1.0], bbox=(-1.0, 1.0), dim=2)
    assert P2.rfac == pytest.approx(1.0001 * math.sqrt(13.0), abs=1e-15)


@pytest.mark.parametrize("surface,dx", [("Sphere", 0.2), ("Sphere", 0.1), ("Torus", 0.25)])
def test_tube_membership_brute_force(surface, dx):
    """The BFS tube equals {grid nodes with |dist| <= r_t} (closed surfaces: one component)."""
    P = orc.Problem(surface, dx)
    T = P.tube()
    N1d, x0 = P.N1d, P.xstart
    g = x0 + dx * np.arange(N1d)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    if surface == "Sphere":
        dist = np.sqrt(X * X + Y * Y + Z * Z) - 1.0
    else:
        R, r = 2.0, 1.0
        rho = np.sqrt(X * X + Y * Y)
        dist = np.sqrt((rho - R) ** 2 + Z * Z) - r
    inside = np.flatnonzero((np.abs(dist) <= P.r_t).ravel())
    assert np.array_equal(np.sort(T["rec"]), inside)
    # node 0 is the grid node nearest cp(xstart) (Tube.cpp:540-551)
    # BFS order: graph distance from node 0 is non-decreasing along the order
    nbr = T["nbr"]
    lev = np.full(P.N, -1)
    lev[0] = 0
    frontier = [0]
    while frontier:
        nxt = []
        for i in frontier:
            for j in nbr[i]:
                if j >= 0 and lev[j] < 0:
                    lev[j] = lev[i] + 1
                    nxt.append(j)
        frontier = nxt
    assert (np.diff(lev) >= 0).all()
    # neighbour symmetry: +d of i is j  <=>  -d of j is i
    for d in range(3):
        i = np.flatnonzero(nbr[:, 2 * d] >= 0)
        assert (nbr[nbr[i, 2 * d], 2 * d + 1] == i).all()


def test_weights_textbook_lagrange():
    xs = np.array([0.1, 0.2, 0.3, 0.4])
    for xq in (0.1234, 0.25, 0.3999):
        w = orc.weights_1d(xs, xq)
        ref = [np.prod([(xq - xs[m]) / (xs[j] - xs[m]) for m in range(4) if m != j]) for j in range(4)]
        assert np.allclose(w, ref, rtol=0, atol=1e-14)
    w = orc.weights_1d(xs, 0.2)           # exact hit (Interpolation.cpp:169-184)
    assert list(w) == [0.0, 1.0, 0.0, 0.0]


@pytest.mark.parametrize("surface,dx", [("Sphere", 0.1), ("Torus", 0.2)])
def test_E_reproduces_cubics_and_rows_sum_to_one(surface, dx):
    P = orc.Problem(surface, dx)
    T = P.tube()
    E = P.E()
    assert np.allclose(E.matvec(np.ones(P.N)), 1.0, atol=1e-13)
    assert (np.diff(E.ptr) == 64).all()
    x, cp = T["x"], T["cp"]
    for q in (lambda v: v[:, 0] ** 3 - 2 * v[:, 1] * v[:, 2] ** 2 + 0.5 * v[:, 1],
              lambda v: v[:, 0] * v[:, 1] * v[:, 2] + v[:, 2] ** 3):
        assert np.allclose(E.matvec(q(x)), q(cp), atol=1e-12)


def test_L_exact_on_quadratics():
    P = orc.Problem("Sphere", 0.1)
    T = P.tube()
    L = P.L()
    x = T["x"]
    u = x[:, 0] ** 2 + 3 * x[:, 1] * x[:, 2] - x[:, 2] ** 2 + x[:, 0]
    lap = 2.0 - 2.0   # d2/dx2 x^2 + d2/dz2 (-z^2)
    full = (T["nbr"] >= 0).all(axis=1)
    assert np.allclose(L.matvec(u)[full], lap, atol=1e-9)
    assert (np.diff(L.ptr)[full] == 7).all()


def test_multiply_equals_explicit_operator():
    P = orc.Problem("Sphere", 0.2)
    L, E, I = P.L(), P.E(), P.ident()
    x = np.random.default_rng(1).standard_normal(P.D)
    for m, n in ((1.0, -1.0), (0.0, 1.0), (1.0, -0.37)):
        A = P.explicit_operator(m, n)
        assert np.allclose(orc.multiply(L, E, I, x, m, n), A.matvec(x), rtol=0, atol=1e-9)


def test_bicgstab_residual_independent():
    P = orc.Problem("Sphere", 0.1)
    L, E, I = P.L(), P.E(), P.ident()
    b = np.random.default_rng(2).standard_normal(P.D)
    x, it, rr = orc.solve(L, E, I, b, 1.0, -1.0, tol=1e-10)
    A = P.explicit_operator(1.0, -1.0)
    true_rel = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert rr <= 1e-10 and true_rel < 1e-9 and it > 5


@pytest.mark.parametrize("problem,dx0,levels,expected", [
    (5, 0.2, 3, 2.18548),    # PoissonProblems.cpp:119-137,169 (1+Laplace) on the sphere
    (6, 0.2, 3, 2.11767),    # :140-154,169 hemisphere, first-order Dirichlet rim
    (4, 0.05, 5, 1.9876),    # :97-115,169 shifted Poisson on the circle (2-D)
])
def test_reference_convergence_orders(problem, dx0, levels, expected):
    errs, dxs = [], []
    for i in range(levels):
        dx = dx0 * 0.5 ** i
        e, _ = orc.poisson_problem_error(problem, dx)
        errs.append(e)
        dxs.append(dx)
    orders = [math.log(errs[i - 1] / errs[i]) / math.log(dxs[i - 1] / dxs[i]) for i in range(1, levels)]
    avg = sum(orders) / len(orders)
    assert abs(avg - expected) < 1e-2      # the reference's own pass criterion (:233)
    assert abs(avg - expected) < 1e-4      # and in fact to the printed digits


def test_config1_closed_form_second_order():
    """(1 - Laplace_S) u = f with u = z + 3xy, f = 3z + 21xy on the unit sphere (closed form)."""
    errs = []
    for dx in (0.2, 0.1):
        P = orc.Problem("Sphere", dx)
        T = P.tube()
        cp = T["cp"]
        b = 3 * cp[:, 2] + 21 * cp[:, 0] * cp[:, 1]
        x, it, rr = orc.solve(P.L(), P.E(), P.ident(), b, 1.0, -1.0)
        Ep = P.plot_E()
        th = np.repeat(np.arange(200) * (2 * np.pi / 199), 200)
        ph = np.tile(np.arange(200) * (np.pi / 199), 200)
        xp = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
        uex = xp[:, 2] + 3 * xp[:, 0] * xp[:, 1]
        errs.append(np.abs(Ep.matvec(x) - uex).max())
    order = math.log(errs[0] / errs[1]) / math.log(2)
    assert errs[1] < 5e-3 and order > 1.7


# PoissonProblemsWithIBC.cpp:360-381: average convergence orders of the
# reference's IBC studies, keyed (boundary order, problem), per seeded
# geometry g (:387-405): problem 2 sphere + circle, 5 sphere + point,
# 8 sphere + arc of the equator.
IBC_ORDERS = {
    (2, 2): [2.03624, 2.20571, 2.09173, 2.0828, 2.1572, 2.10696, 1.97858, 1.97003, 2.17008, 1.9772],
    (2, 5): [2.16283, 2.13198, 1.92074, 2.22093, 2.27176, 2.36021, 1.89006, 2.12619, 2.22663, 2.70325],
    (1, 2): [0.954594, 0.990541, 1.07685, 0.865972, 0.784669, 0.922806, 0.991011, 0.987983, 0.90721, 0.956464],
    (1, 5): [2.42046, 2.03873, 1.03422, 0.973755, 1.54908, 2.12488, 2.12239, 1.02281, 1.05949, 1.37816],
    (2, 8): [1.91054, 1.90547, 1.88648, 1.92365, 1.90556, 1.90494, 1.89661, 1.91465, 1.81442, 1.73931],
    (1, 8): [1.76665, 1.79881, 1.72063, 1.74305, 1.72427, 1.75405, 1.73974, 1.80824, 1.6929, 1.6667],
}
# The arc study reproduces its table for the arcs that end on the polar
# angle's branch cut (g = 0 at both orders, g = 9 at order 1); for g = 1..8 the
# restatement's orders differ from the table (parity unpinned for those arcs).
IBC_PINNED = [(o, p, g) for (o, p) in IBC_ORDERS for g in range(10) if p != 8] + [(2, 8, 0), (1, 8, 0), (1, 8, 9)]
IBC_FAST = [(2, 2, 0), (1, 5, 3), (2, 8, 0)]


def _ibc_study(args):
    o, p, g = args
    return orc.ibc_study_order(p, g, o)


def _spawn_pool(n):
    import multiprocessing
    return multiprocessing.get_context("spawn").Pool(n)


@pytest.mark.parametrize("order,problem,g", IBC_FAST)
def test_reference_ibc_convergence_orders(order, problem, g):
    """The oracle's PoissonWithIBC reproduces the reference's IBC known answers
    (the reference's own pass criterion is 1e-2, PoissonProblemsWithIBC.cpp:484-493)."""
    v = orc.ibc_study_order(problem, g, order)
    assert abs(v - IBC_ORDERS[(order, problem)][g]) < 1e-2, v


@pytest.mark.slow
def test_reference_ibc_convergence_orders_all():
    """Every pinned entry of the tables (1st- and 2nd-order Dirichlet circles and points, arcs),
    in 8 processes."""
    cases = [c for c in IBC_PINNED if c not in IBC_FAST]
    with _spawn_pool(min(8, os.cpu_count() or 1)) as pool:
        got = pool.map(_ibc_study, cases)
    bad = [(c, v, IBC_ORDERS[c[:2]][c[2]]) for c, v in zip(cases, got) if not abs(v - IBC_ORDERS[c[:2]][c[2]]) < 1e-2]
    assert not bad, bad


@pytest.mark.parametrize("params,bbox", [(None, None), ([0.6, 1.3, 0.9], (-1.3, 1.3)), ([1.0, 1.0, 0.5], (-1.0, 1.0))])
def test_ellipsoid_closest_point(params, bbox):
    """Ellipsoid closest points (the reference's LBFGS minimiser is not vendored: parity vs the reference is
    unpinned).  Pinned here by what fixes the exact closest point: cp on the surface, x - cp along the
    surface normal at cp, |x - cp| = dist, and no point of a dense surface sampling closer than cp."""
    P = orc.Problem("Ellipsoid", 0.1, params=params, bbox=bbox)
    T = P.tube()
    a, b, c = params if params is not None else (1.5, 1.0, 0.75)
    x, cp, dist = T["x"], T["cp"], T["dist"]
    phi = (cp[:, 0] / a) ** 2 + (cp[:, 1] / b) ** 2 + (cp[:, 2] / c) ** 2 - 1.0
    assert np.abs(phi).max() < 1e-14
    assert np.allclose(np.linalg.norm(x - cp, axis=1), dist, rtol=0, atol=1e-15)
    grad = np.stack([cp[:, 0] / a ** 2, cp[:, 1] / b ** 2, cp[:, 2] / c ** 2], 1)
    grad /= np.linalg.norm(grad, axis=1)[:, None]
    cr = np.cross(x - cp, grad)
    assert np.abs(cr).max() < 1e-13
    # global minimum: dense sampling, 600 x 1200 points, on 100 nodes
    th = np.linspace(0, np.pi, 600)[:, None]
    ph = np.linspace(0, 2 * np.pi, 1200)[None, :]
    S = np.stack([a * np.sin(th) * np.cos(ph), b * np.sin(th) * np.sin(ph), c * np.cos(th) + 0 * ph], -1).reshape(-1, 3)
    idx = np.random.default_rng(2).choice(len(x), 100, replace=False)
    for i in idx:
        dmin = np.sqrt(((S - x[i]) ** 2).sum(1).min())
        assert dist[i] <= dmin + 1e-12


def test_heat_first_order_dirichlet_converges():
    """HeatWithIBC with a 1st-order Dirichlet curve (g_boundary_order = 1,
    HeatProblemsWithIBC.cpp:113-115): backward Euler to t = 0.1 against the
    closed form exp(-2 t) cos(phi); the Dirichlet value sits on the ghost rows
    themselves, and placing it elsewhere stops the convergence."""
    e = [orc.heat_problem_error(1, dx, 1, 0) for dx in (0.2, 0.1)]
    rate = np.log2(e[0] / e[1])
    assert e[1] < 0.03 and 0.8 < rate < 1.5, (e, rate)


# HeatProblemsWithIBC.cpp:297-318: average convergence orders of problems 3-5
# (sphere with a Circle3D curve; explicit, backward-Euler and Crank-Nicolson
# steps), keyed (boundary order, boundary type): Dirichlet curves are the
# tilted circle, Neumann curves the great circle y = 0 (:94-111).
HEAT_ORDERS = {
    (1, 0): [0.969203, 0.96295, 0.970163],
    (2, 0): [1.98885, 1.15183, 2.01012],
    (1, 1): [0.946114, 0.93853, 0.944366],
    (2, 1): [2.07915, 1.15527, 2.04239],
}


def _heat_study(args):
    o, t, m = args
    return orc.heat_study_order(m, o, t)


@pytest.mark.slow
def test_reference_heat_ibc_convergence_orders():
    """All twelve entries, in 8 processes, within the reference's criterion (:372)."""
    cases = [(o, t, m) for (o, t) in HEAT_ORDERS for m in range(3)]
    with _spawn_pool(min(8, os.cpu_count() or 1)) as pool:
        got = pool.map(_heat_study, cases)
    bad = [(c, v, HEAT_ORDERS[c[:2]][c[2]]) for c, v in zip(cases, got) if not abs(v - HEAT_ORDERS[c[:2]][c[2]]) < 1e-2]
    assert not bad, bad


@pytest.mark.parametrize("method", [0, 1])
def test_rd_uniform_state_follows_reaction_ode(method):
    """ReactionDiffusionWithIBC::TimeStep (ReactionDiffusionWithIBC.cpp:79-104) with a
    Neumann curve from a state uniform on every row.  Lfull rows with all six
    neighbours sum to 0 and Efull rows to 1, and E reads only such rows, so on
    them each step is the forward-Euler map of the scalar Gray-Scott ODE (the
    implicit variant solves (1 - dt nu L) u' = rhs, i.e. u' = rhs there).  Rows
    whose stencil is cut by the tube edge keep other values, unread."""
    dx, F, k, nu = 0.2, 0.054, 0.063, (0.2 / 3) ** 2
    dt = 0.9 if method == 0 else 5.0
    # the heat study's Neumann curve, the great circle y = 0 (HeatProblemsWithIBC.cpp:102-106)
    P = orc.IBCProblem("Sphere", dx, ibcs=[orc.IBC("Circle3D", params=[1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0],
                                                   boundary_order=2, boundary_type=(1, 1),
                                                   on_ibc_tolerance=0.1 * dx * dx)])
    L, E = P.L(), P.E()
    rowsum = np.bincount(np.repeat(np.arange(P.D), np.diff(L.ptr)), weights=L.val, minlength=P.D)
    complete = np.abs(rowsum) < 1e-8
    assert complete[E.col[E.val != 0]].all()       # E never reads a cut row
    u, v = orc.rd_run(method, dx, np.full(P.D, 0.7), np.full(P.D, 0.2), F, k, nu, nu / 3, dt, 4, 2, 1)
    a, b = 0.7, 0.2
    for _ in range(4):
        a, b = a + dt * (-a * b * b + F * (1 - a)), b + dt * (a * b * b - (F + k) * b)
    du, dv = np.abs(u - a)[complete].max(), np.abs(v - b)[complete].max()
    assert du < 1e-12 and dv < 1e-12, (du, dv)
    if method == 0:                                 # the final extension rewrites every row
        assert np.abs(u - a).max() < 1e-12


def _cp_mesh_by_projection(V, F, x):
    """Closest point on a triangle mesh by an independent route: the nearest of
    the plane projection (when its barycentric coordinates are all >= 0) and the
    clamped projections onto the three edges, over all triangles.  Returns the
    closest point, the distance and the margin to the next distinct candidate."""
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    N = np.cross(B - A, C - A)
    cps, ds, margins = [], [], []
    for p in x:
        cand = []
        t = ((p - A) * N).sum(1) / (N * N).sum(1)
        q = p - t[:, None] * N
        # barycentric coordinates of q by sub-triangle areas along N
        l0 = (np.cross(B - q, C - q) * N).sum(1)
        l1 = (np.cross(C - q, A - q) * N).sum(1)
        l2 = (np.cross(A - q, B - q) * N).sum(1)
        inside = (l0 >= 0) & (l1 >= 0) & (l2 >= 0)
        cand.append(np.where(inside[:, None], q, np.nan))
        for P, Q in ((A, B), (B, C), (C, A)):
            s = np.clip(((p - P) * (Q - P)).sum(1) / ((Q - P) ** 2).sum(1), 0, 1)
            cand.append(P + s[:, None] * (Q - P))
        Cq = np.concatenate(cand)
        d = np.linalg.norm(Cq - p, axis=1)
        d[np.isnan(d)] = np.inf
        i = np.argmin(d)
        far = np.linalg.norm(Cq - Cq[i], axis=1) > 1e-9
        cps.append(Cq[i])
        ds.append(d[i])
        margins.append(d[far].min() - d[i] if far.any() else np.inf)
    return np.array(cps), np.array(ds), np.array(margins)


@pytest.mark.parametrize("mesh", ["icosphere", "displaced"])
def test_mesh_closest_point_independent_route(mesh):
    """Surface::cpTri (Surface.cpp:1751-1779; fcpw in the reference, parity with
    fcpw unpinned): the oracle's Ericson region walk equals the projection
    route on random points around the mesh."""
    from cpmgpu import workloads
    V, F = workloads.icosphere(2) if mesh == "icosphere" else workloads.displaced_icosphere(2, seed=7)
    x = np.random.default_rng(5).uniform(-1.4, 1.4, (400, 3))
    cp, d = orc.mesh_cp(V, F, x)
    cq, dq, margin = _cp_mesh_by_projection(V, F, x)
    assert np.abs(d - dq).max() < 1e-13
    assert np.abs(d - np.linalg.norm(cp - x, axis=1)).max() < 1e-14
    uniq = margin > 1e-9
    assert uniq.mean() > 0.95 and np.abs(cp - cq)[uniq].max() < 1e-12


def test_mesh_closest_point_box_closed_form():
    """On the cube [-1, 1]^3 made of 12 triangles: outside points project to
    clip(x, -1, 1); inside points to the nearest face (unique nearest face only)."""
    from cpmgpu import workloads
    V, F = workloads.box_mesh()
    x = np.random.default_rng(6).uniform(-2, 2, (3000, 3))
    cp, d = orc.mesh_cp(V, F, x)
    out = (np.abs(x) > 1).any(1)
    c = np.clip(x, -1, 1)
    assert np.abs(cp[out] - c[out]).max() < 1e-15
    assert np.abs(d[out] - np.linalg.norm(x[out] - c[out], axis=1)).max() < 1e-15
    gap = 1 - np.abs(x[~out])
    assert np.abs(d[~out] - gap.min(1)).max() < 1e-15
    srt = np.sort(gap, 1)
    uniq = srt[:, 1] - srt[:, 0] > 1e-12
    k = np.argmin(gap, 1)
    exp = x[~out].copy()
    exp[np.arange(len(exp)), k] = np.sign(exp[np.arange(len(exp)), k])
    assert np.abs(cp[~out][uniq] - exp[uniq]).max() < 1e-15


def _geodesic_sphere_point_error(dx):
    P = orc.IBCProblem("Sphere", dx, ibcs=[orc.IBC("Point", params=[0.0, 0.0, 1.0], boundary_order=1,
                                                   on_ibc_tolerance=0.1 * dx * dx)])
    phi, it = P.geodesic(dx * dx, 2000)
    th = np.repeat(np.arange(200) * (2 * np.pi / 199), 200)
    ph = np.tile(np.arange(200) * (np.pi / 199), 200)
    xp = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    up = P.plot_matrix(xp).matvec(phi)                      # plotEfull * phi (GeodesicDistance.cpp:62-66)
    exact = np.arccos(np.clip(np.cos(ph), -1, 1))          # distance from the north pole on the unit sphere
    e = np.abs(up - exact)
    return phi, it, e.max(), e.mean()


def test_geodesic_distance_sphere_point(parallel_oracle):
    """GeodesicDistance::ComputeDistance (GeodesicDistance.cpp:11-67, dt = dx^2, 2000 Jacobi
    sweeps as in examples/GeodesicDistanceDzuikPoint) from a point on the unit sphere against
    the exact distance acos(z) at the plot points: phi >= 0 with minimum 0, errors of the
    heat method's size that shrink with dx (a wrong sign, a missing normal projection or
    extension gives O(1) errors)."""
    phi2, it2, emax2, emean2 = _geodesic_sphere_point_error(0.2)
    phi1, it1, emax1, emean1 = _geodesic_sphere_point_error(0.1)
    assert phi1.min() == 0.0 and phi2.min() == 0.0
    assert it1[0] >= 1 and all(k > 0 for k in it1[1:])
    assert emax1 < 0.1 and emean1 < 0.035, (emax1, emean1)
    assert emean1 < 0.85 * emean2 and emax1 < 0.8 * emax2, (emean1, emean2, emax1, emax2)


def test_ibc_parameter_closed_forms():
    """Surface::IBCParameter (src/Surface.cpp:2731-2788) on a square polyline, against closed forms:
    normalised arc length along the vertex list, which for a closed polyline holds each vertex once,
    so the closing edge is not counted (the total is 3 edges) and a point on it takes the parameter
    of its nearer end's edge; and the nearest vertex's value for parameterised samples."""
    a = 0.7
    sq = np.array([[a, a, 0], [-a, a, 0], [-a, -a, 0], [a, -a, 0], [a, a, 0]], float)
    P = orc.IBCProblem("Sphere", 0.2, ibcs=[orc.IBC("Polyline", verts=sq, boundary_order=2,
                                                    on_ibc_tolerance=0.004)])
    e = 2 * a
    s = np.array([0.1, 0.3, 0.55, 0.9, 1.25])
    pts = np.concatenate([np.stack([a - s, np.full_like(s, a), 0 * s], 1),        # edge 0 -> 1
                          np.stack([np.full_like(s, -a), a - s, 0 * s], 1),       # edge 1 -> 2
                          np.stack([s - a, np.full_like(s, -a), 0 * s], 1)])      # edge 2 -> 3
    want = np.concatenate([s, e + s, 2 * e + s]) / (3 * e)
    got = P.ibc_parameter(0, pts)
    assert np.abs(got - want).max() < 1e-14, (got, want)
    # on the vertices: the shorter-edge rule picks the following edge (equal edges)
    got = P.ibc_parameter(0, sq[1:3])
    assert np.abs(got - np.array([1, 2]) / 3).max() < 1e-14
    # the closing edge (vertex 3 -> vertex 0): nearer vertex 3, the last, counts back along edge 2 -> 3
    # from the end; nearer vertex 0, the first, counts its distance along edge 0 -> 1
    c = np.array([[a, -a + 0.2, 0], [a, a - 0.3, 0]])
    got = P.ibc_parameter(0, c)
    assert np.abs(got - np.array([3 * e - 0.2, 0.3]) / (3 * e)).max() < 1e-14, got
    # parameterised samples: the nearest vertex's value
    got = P.ibc_parameter(0, [[a - 0.2, a, 0], [-a, -a + 0.65, 0], [a, -a + 0.2, 0]], vertex_param=[0.0, 1.5, 2.5, 3.5])
    assert np.array_equal(got, [0.0, 2.5, 3.5])
