// Convergence-study known answers restated from
// convergence_studies/PoissonProblems/PoissonProblems.cpp.
// TEST INFRASTRUCTURE ONLY (see orc.h).
#include <cmath>

#include "orc.h"

namespace orc {

// ErrorAndVisualize, PoissonProblems.cpp:15-54: max |E_plot u - exact| over the
// surface parameterisation points.
static double plot_error(const Tube& t, const Vec& u, const std::function<double(double, double)>& exact) {
    CSR P = plot_E(t);
    Vec up((size_t)P.rows);
    P.matvec(u.data(), up.data());
    double e = 0.0;
    for (size_t i = 0; i < up.size(); ++i) {
        double ue = t.dim == 2 ? exact(t.surf.thetap[i], 0) : exact(t.surf.thetap[i], t.surf.phip[i]);
        e = std::max(e, std::fabs(up[i] - ue));
    }
    return e;
}

// Poisson::GetRHS(f(theta, phi)[, exact]), problems/Poisson.cpp:61-148.
static Vec rhs_theta_phi(const PoissonProblem& P, const std::function<double(double, double)>& f,
                         const std::function<double(double, double)>* exact) {
    const Tube& t = P.tube;
    const i64 N = t.N();
    Vec rhs((size_t)N);
    const double R = t.surf.spec().params[0];
    auto eval = [&](const std::function<double(double, double)>& g, i64 i) {
        const double* c = &t.cpx[(size_t)(i * t.dim)];
        double th = std::atan2(c[1], c[0]);
        return t.dim == 2 ? g(th, 0) : g(th, std::acos(c[2] / R));
    };
    for (i64 i = 0; i < N; ++i) rhs[(size_t)i] = eval(f, i);
    if (exact && !P.which_bdy.empty() && P.bdy_order == 1)
        for (i64 i = 0; i < N; ++i)
            for (i64 b : P.which_bdy)
                if (t.bdy[(size_t)i] == b) rhs[(size_t)i] = eval(*exact, i);
    return rhs;
}

static double solve_and_error(const PoissonProblem& P, double m, double n, const Vec& rhs,
                              const std::function<double(double, double)>& exact, int* iters) {
    System S;
    S.m = m;
    S.n = n;
    S.L = &P.Lfull;
    S.E = &P.Efull;
    S.ident = P.ident;
    S.precond = PRECOND_DIAGONAL;
    Vec u(rhs.size());
    SolveInfo info = bicgstab(S, rhs.data(), u.data(), 1e-10, 100000);   // Solver.cpp:110
    if (iters) *iters = info.iters;
    return plot_error(P.tube, u, exact);
}

// problem 4: ShiftedPoissonCircle (:97-115); 5: ShiftedPoissonSphere (:119-137);
// 6: PoissonHemisphere with first-order Dirichlet rim (:140-154).
double poisson_problem_error(int problem, double dx, int* iters) {
    if (problem == 4) {
        auto ue = [](double th, double) { return std::sin(th) + std::sin(12 * th); };
        auto f = [](double th, double) { return 2 * std::sin(th) + 145 * std::sin(12 * th); };
        PoissonProblem P = make_poisson(SurfaceSpec::defaults("Circle"), dx);
        return solve_and_error(P, 1.0, -1.0, rhs_theta_phi(P, f, nullptr), ue, iters);
    }
    auto ue = [](double th, double ph) {
        return std::cos(3 * th) * std::pow(std::sin(ph), 3) * (9 * std::pow(std::cos(ph), 2) - 1);
    };
    if (problem == 5) {
        auto f = [&](double th, double ph) { return -29 * ue(th, ph); };
        PoissonProblem P = make_poisson(SurfaceSpec::defaults("Sphere"), dx);
        return solve_and_error(P, 1.0, 1.0, rhs_theta_phi(P, f, nullptr), ue, iters);
    }
    if (problem == 6) {
        auto f = [&](double th, double ph) { return -30 * ue(th, ph); };
        std::function<double(double, double)> exact = ue;
        PoissonProblem P = make_poisson(SurfaceSpec::defaults("Hemisphere"), dx, 1, {1});
        return solve_and_error(P, 0.0, 1.0, rhs_theta_phi(P, f, &exact), ue, iters);
    }
    throw std::runtime_error("poisson_problem_error: unsupported problem");
}

}  // namespace orc

// ---------------------------------------------------------------- IBC studies
// convergence_studies/PoissonProblemsWithIBC/PoissonProblemsWithIBC.cpp:
// problem 2 PoissonSphereCircleIBC (:230-256), problem 5 PoissonSpherePointIBC
// (:274-301), boundary order 2, ten seeded geometries (:398-424).
#include "orc_ibc.h"

namespace orc {

static const double kRand1[10] = {0.0, 0.170, 0.297, 0.331, 0.481, 0.511, 0.631, 0.735, 0.813, 0.973};
static const double kRand2[10] = {0.25, 0.874, 0.473, 0.973, 0.812, 0.693, 0.873, 0.912, 0.945, 1.0};
static const double kRand3[10] = {0.55, 0.123, -0.743, 0.547, -0.239, -0.693, 0.731, -0.175, 0.453, -0.364};

// ComputeCircle3DParams (:117-138); normalises `normal` in place like the reference.
static Vec circle3d_params(Vec& normal) {
    double nn = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    if (nn > 0.0) for (auto& c : normal) c /= nn;
    const double r = 0.2, theta = 0.1, phi = 0.2;
    const double pp[3] = {r * std::cos(phi) * std::sin(theta), r * std::sin(phi) * std::sin(theta), r * std::cos(theta)};
    const double rho = -(pp[0] * normal[0] + pp[1] * normal[1] + pp[2] * normal[2]);
    const double cen[3] = {rho * normal[0], rho * normal[1], rho * normal[2]};
    return {std::sqrt(1 - rho * rho), cen[0], cen[1], cen[2], normal[0], normal[1], normal[2]};
}

static double sphere_harmonic(const double* x) {
    const double th = std::atan2(x[1], x[0]), ph = std::acos(x[2] / 1.0);
    return std::cos(3 * th) * std::pow(std::sin(ph), 3) * (9 * std::pow(std::cos(ph), 2) - 1);
}

// The study's curve of problem 2 (PoissonSphereCircleIBC, :221-233: circle
// with the seeded normal), 5 (PoissonSpherePointIBC, :274-279: point at
// angles (endpoint1, endpoint2)) or 8 (PoissonSphereArcIBC, :318-326: arc of
// the equator between endpoint1 and endpoint2; endpoints :398-405).
SurfaceSpec ibc_study_spec(int problem, int g, Vec& normal_state) {
    SurfaceSpec s;
    const double e1 = 2.0 * M_PI * kRand1[g] - M_PI, e2 = 2.0 * M_PI * kRand2[g] - M_PI;
    if (problem == 2) {
        s.type = "Circle3D";
        s.params = circle3d_params(normal_state);
    } else if (problem == 5) {
        s.type = "Point";
        s.params = {std::cos(e2) * std::cos(e1), std::cos(e2) * std::sin(e1), std::sin(e2)};
    } else if (problem == 8) {
        s.type = "Arc";
        s.params = {1.0, e1, e2};
        s.is_open = true;   // SetIsOpen (SurfaceSpecifier.cpp:377)
    } else {
        throw std::runtime_error("unsupported IBC study problem");
    }
    s.bbox = {-1.0, 1.0};
    s.dim = 3;
    return s;
}

// One level of the study: error of the Dirichlet IBC solve of boundary order
// `order` (g_boundary_order, :14) at dx.
double ibc_problem_error(int problem, int g, double dx, Vec& normal_state, int* iters, int order) {
    SurfaceSpec ibc = ibc_study_spec(problem, g, normal_state);
    ibc.meta.on_ibc_tolerance = 0.1 * dx * dx;
    ibc.meta.boundary_order = order;
    PoissonIBCProblem P = make_poisson_ibc(SurfaceSpec::defaults("Sphere"), {ibc}, dx);
    auto f = [](const double* x) { return 30 * sphere_harmonic(x); };
    Vec rhs = poisson_ibc_rhs(P, f, sphere_harmonic);
    System S;
    S.m = 0.0;
    S.n = -1.0;
    S.L = &P.Lfull;
    S.E = &P.Efull;
    S.ident = P.ident;
    S.precond = PRECOND_DIAGONAL;
    Vec u(rhs.size());
    SolveInfo info = bicgstab(S, rhs.data(), u.data(), 1e-10, 100000);
    if (iters) *iters = info.iters;
    CSR Pe = ibc_plot_E(P);
    Vec up((size_t)Pe.rows);
    Pe.matvec(u.data(), up.data());
    double e = 0.0;
    for (size_t i = 0; i < up.size(); ++i) e = std::max(e, std::fabs(up[i] - sphere_harmonic(P.tube.surf.xp[i].data())));
    return e;
}

double ibc_study_order(int problem, int g, int* iters_last, int order) {
    Vec normal = {kRand1[g], kRand2[g], kRand3[g]};
    Vec err, dxs;
    for (int lv = 0; lv < 3; ++lv) {
        const double dx = std::pow(0.5, lv) * 0.2;
        err.push_back(ibc_problem_error(problem, g, dx, normal, iters_last, order));
        dxs.push_back(dx);
    }
    return average_order(err, dxs);
}

}  // namespace orc

// ---------------------------------------------------------------- heat with an IBC
// HeatWithIBC (problems/HeatWithIBC.cpp:6-170) and the study
// convergence_studies/HeatProblemsWithIBC/HeatProblemsWithIBC.cpp:
// SphereCircleIBC (:90-193), 2nd-order Dirichlet, methods 0 explicit,
// 1 backward Euler, 2 Crank-Nicolson (problems 3-5), final time 0.1.
namespace orc {

// ComputeCircle3DParams (:71-88) with the study's plane point; normalises `normal`.
static Vec circle3d_params_pp(Vec normal, const double* pp) {
    double nn = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    if (nn > 0.0) for (auto& c : normal) c /= nn;
    const double rho = -(pp[0] * normal[0] + pp[1] * normal[1] + pp[2] * normal[2]);
    return {std::sqrt(1 - rho * rho), rho * normal[0], rho * normal[1], rho * normal[2], normal[0], normal[1], normal[2]};
}

// The study's curve (:94-111): for Dirichlet curves (g_boundary_type 0) the
// circle cut by the plane of normal (1.1, 1.234, 1.321) through a point at
// r = 0.27, theta 0.132, phi 0.221; for Neumann curves the great circle y = 0
// (normal (0, 1, 0) through the origin), on which u = e^{-2t} cos(phi) has a
// zero normal derivative.
SurfaceSpec heat_study_ibc(double dx, int order = 2, int btype = 0) {
    SurfaceSpec s;
    s.type = "Circle3D";
    if (btype == 0) {
        const double r = 0.27, theta = 0.132, phi = 0.221;
        const double pp[3] = {r * std::cos(phi) * std::sin(theta), r * std::sin(phi) * std::sin(theta),
                              r * std::cos(theta)};
        s.params = circle3d_params_pp({1.1, 1.234, 1.321}, pp);
    } else {
        const double pp[3] = {0.0, 0.0, 0.0};
        s.params = circle3d_params_pp({0.0, 1.0, 0.0}, pp);
    }
    s.bbox = {-1.0, 1.0};
    s.dim = 3;
    s.meta.on_ibc_tolerance = 0.1 * dx * dx;
    s.meta.boundary_order = order;     // g_boundary_order, g_boundary_type (:113-115)
    s.meta.boundary_type[0] = btype;
    return s;
}

// f = cos(phi), phi = acos(z / R) (:133-134), R = 1
static double heat_f(const double* x) { return std::cos(std::acos(x[2] / 1.0)); }

// Time-step size of the study (:118-127).
void heat_study_dt(int method, double dx, double final_time, double* dt, int* steps) {
    double d = method > 0 ? 0.1 * dx : 0.1 * dx * dx;
    const int n = (int)std::ceil(final_time / d);
    *dt = final_time / n;
    *steps = n;
}

// Row holding the Dirichlet value of DOF i of curve c (unoriented curves):
// the 2nd-order identity row, or the ghost row itself at 1st order; -1 for
// Neumann curves (HeatWithIBC.cpp:63-72).
static i64 heat_dirichlet_row(const IBCData& I, size_t c, size_t i) {
    if (I.meta[c].boundary_type[0] != 0) return -1;
    return I.meta[c].boundary_order == 2 ? I.identity_rows_start + I.dof[c].dir_sub_of_set[i] + I.dir_start[c]
                                         : I.identity_rows_start + (i64)i + I.start[c];
}

// The study's initial condition (:139-181): f at cp on tube rows, f at the
// curve's closest point on the ghost rows of Neumann and 2nd-order Dirichlet
// curves, the Dirichlet value uexact(0) on the Dirichlet rows.
Vec heat_initial(const PoissonIBCProblem& P) {
    const Tube& t = P.tube;
    const i64 N = t.N();
    Vec u((size_t)P.Lfull.rows, 0.0);
    for (i64 i = 0; i < N; ++i) u[(size_t)i] = heat_f(&t.cpx[(size_t)(i * 3)]);
    const IBCData& I = P.ibc;
    for (size_t c = 0; c < I.dof.size(); ++c) {
        const int ty = I.meta[c].boundary_type[0];
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i) {
            const double v = heat_f(I.dof[c].cpc[i].data());
            if (ty == 1 || (ty == 0 && I.meta[c].boundary_order == 2)) u[(size_t)I.start[c] + i] = v;
            const i64 r = heat_dirichlet_row(I, c, i);
            if (r >= 0) u[(size_t)r] = v;
        }
    }
    return u;
}

// SetDirichletValue (problems/HeatWithIBC.cpp:50-78): u on the Dirichlet rows.
static void heat_set_dirichlet(const PoissonIBCProblem& P, Vec& u, double t) {
    const IBCData& I = P.ibc;
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i) {
            const i64 r = heat_dirichlet_row(I, c, i);
            if (r >= 0) u[(size_t)r] = std::exp(-2 * t) * heat_f(I.dof[c].cpc[i].data());
        }
}

// TimeStep (problems/HeatWithIBC.cpp:80-103), `steps` steps from u.
int heat_run(const PoissonIBCProblem& P, int method, double dt, int steps, Vec& u) {
    const double r = method == 1 ? 1.0 : 0.5;
    System S;                                     // Initialize :160-166: Solver(1, -r dt)
    S.m = 1.0;
    S.n = -r * dt;
    S.L = &P.Lfull;
    S.E = &P.Efull;
    S.ident = P.ident;
    S.precond = PRECOND_DIAGONAL;
    const size_t D = u.size();
    Vec tmp(D), res(D);
    int iters = 0;
    for (int k = 0; k < steps; ++k) {
        const double t = (k + 1) * dt;
        if (method > 0) {
            if (method == 2) {
                S.multiply(u.data(), tmp.data(), 1.0, 0.5 * dt);
                u = tmp;
            }
            heat_set_dirichlet(P, u, t);
            SolveInfo info = bicgstab(S, u.data(), res.data(), 1e-10, 100000);
            iters += info.iters;
            u = res;
        } else {
            P.Lfull.matvec(u.data(), tmp.data());
            for (size_t i = 0; i < D; ++i) tmp[i] = u[i] + dt * tmp[i];
            P.Efull.matvec(tmp.data(), u.data());   // closest-point extension
            heat_set_dirichlet(P, u, t);
        }
    }
    return iters;
}

double heat_problem_error(int method, double dx, int order, int btype) {
    PoissonIBCProblem P = make_poisson_ibc(SurfaceSpec::defaults("Sphere"), {heat_study_ibc(dx, order, btype)}, dx);
    double dt;
    int steps;
    heat_study_dt(method, dx, 0.1, &dt, &steps);
    Vec u = heat_initial(P);
    heat_run(P, method, dt, steps, u);
    CSR Pe = ibc_plot_E(P);
    Vec up((size_t)Pe.rows);
    Pe.matvec(u.data(), up.data());
    double e = 0.0;
    for (size_t i = 0; i < up.size(); ++i)
        e = std::max(e, std::fabs(up[i] - std::exp(-2 * 0.1) * heat_f(P.tube.surf.xp[i].data())));
    return e;
}

// ConvergenceStudy(0.1, 3, Heat) (:363-366) of problems 3-5 for boundary
// order `order` and type `btype` (g_boundary_order, g_boundary_type, :20-21).
double heat_study_order(int method, int order, int btype) {
    Vec err, dxs;
    for (int lv = 0; lv < 3; ++lv) {
        const double dx = std::pow(0.5, lv) * 0.1;
        err.push_back(heat_problem_error(method, dx, order, btype));
        dxs.push_back(dx);
    }
    return average_order(err, dxs);
}

}  // namespace orc

namespace orc {
// The heat study's system at dx with `steps` steps of `method` from the
// study's initial condition; returns D and (when u != null) the final state.
int64_t heat_run_study(int method, double dx, int steps, int order, int btype, double* u_out, double* dt_out) {
    PoissonIBCProblem P = make_poisson_ibc(SurfaceSpec::defaults("Sphere"), {heat_study_ibc(dx, order, btype)}, dx);
    double dt;
    int n;
    heat_study_dt(method, dx, 0.1, &dt, &n);
    if (dt_out) *dt_out = dt;
    Vec u = heat_initial(P);
    if (u_out) {
        heat_run(P, method, dt, steps, u);
        std::copy(u.begin(), u.end(), u_out);
    }
    return (int64_t)u.size();
}
}  // namespace orc

// ---------------------------------------------------------------- reaction-diffusion
// ReactionDiffusionWithIBC (problems/ReactionDiffusionWithIBC.cpp:79-104) on the
// heat study's sphere and Circle3D curve (any order / boundary type), from a
// given initial state; the Dirichlet values are 0, as in
// examples/ReactionDiffusionFish/ReactionDiffusionFish.cpp:73-77.
//   method 0: the reference's explicit step
//   method 1: implicit diffusion, (1 - dt nu Lfull-with-E) u' = u + dt Ru(u, v)
//             with the identity rows of the right-hand side set (no counterpart
//             in the reference; the same Solver(1, -dt nu) as HeatWithIBC.cpp:160-166)
namespace orc {

static void rd_set_dirichlet(const IBCData& I, Vec& u, double value) {
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i) {
            const i64 r = heat_dirichlet_row(I, c, i);
            if (r >= 0) u[(size_t)r] = value;
        }
}

// `steps` steps on any IBC problem (Lfull, Efull, identity rows, curves I).
void rd_run_problem(const CSR& Lfull, const CSR& Efull, const std::vector<uint8_t>& ident, const IBCData& I,
                    int method, const double* prm, int steps, double* u_io, double* v_io) {
    const size_t D = (size_t)Lfull.rows;
    const double F = prm[0], k = prm[1], nu_u = prm[2], nu_v = prm[3], dt = prm[4];
    Vec u(u_io, u_io + D), v(v_io, v_io + D), un(D), vn(D), tu(D), tv(D);
    System Su, Sv;
    for (System* S : {&Su, &Sv}) {
        S->m = 1.0;
        S->L = &Lfull;
        S->E = &Efull;
        S->ident = ident;
        S->precond = PRECOND_DIAGONAL;
    }
    Su.n = -dt * nu_u;
    Sv.n = -dt * nu_v;
    for (int s = 0; s < steps; ++s) {
        if (method == 0) {
            Lfull.matvec(u.data(), tu.data());
            Lfull.matvec(v.data(), tv.data());
            for (size_t i = 0; i < D; ++i) {
                tu[i] = u[i] + nu_u * dt * tu[i];          // :86-87
                tv[i] = v[i] + nu_v * dt * tv[i];
                tu[i] += dt * (-u[i] * v[i] * v[i] + F * (1.0 - u[i]));   // :89-93
                tv[i] += dt * (u[i] * v[i] * v[i] - (F + k) * v[i]);
            }
            Efull.matvec(tu.data(), u.data());            // :96-97
            Efull.matvec(tv.data(), v.data());
        } else {
            for (size_t i = 0; i < D; ++i) {
                tu[i] = u[i] + dt * (-u[i] * v[i] * v[i] + F * (1.0 - u[i]));
                tv[i] = v[i] + dt * (u[i] * v[i] * v[i] - (F + k) * v[i]);
            }
            rd_set_dirichlet(I, tu, 0.0);
            rd_set_dirichlet(I, tv, 0.0);
            bicgstab(Su, tu.data(), u.data(), 1e-10, 100000);
            bicgstab(Sv, tv.data(), v.data(), 1e-10, 100000);
        }
        rd_set_dirichlet(I, u, 0.0);                         // :99-100
        rd_set_dirichlet(I, v, 0.0);
    }
    std::copy(u.begin(), u.end(), u_io);
    std::copy(v.begin(), v.end(), v_io);
}

int64_t rd_run_study(int method, double dx, int order, int btype, const double* prm, int steps, double* u_io,
                     double* v_io) {
    PoissonIBCProblem P = make_poisson_ibc(SurfaceSpec::defaults("Sphere"), {heat_study_ibc(dx, order, btype)}, dx);
    const size_t D = (size_t)P.Lfull.rows;
    if (!u_io) return (int64_t)D;
    rd_run_problem(P.Lfull, P.Efull, P.ident, P.ibc, method, prm, steps, u_io, v_io);
    return (int64_t)D;
}
}  // namespace orc

// ---------------------------------------------------------------- geodesic distance
// GeodesicDistance::ComputeDistance (problems/GeodesicDistance.cpp:11-67) up
// to the plotting interpolation: phi on the D rows.  Rows other than tube
// rows of the Poisson right-hand sides are 0 (the reference sizes that vector
// N + G and fills only its Dirichlet rows, :268-290).
namespace orc {

static double smooth_h(double scale, double t) { return 0.5 * std::tanh(-scale * t) + 0.5; }   // :73-78

// Dc[d] x on the tube rows, FDMatrices::BuildGradientMatrices (FDMatrices.cpp:64-92):
// +0.5/dx at the +d neighbour, -0.5/dx at the -d neighbour.
static double dc_row(const Tube& t, i64 i, int d, double c, const double* x) {
    double s = 0.0;
    const i64 p = t.nbr[(size_t)(6 * i + 2 * d)], m = t.nbr[(size_t)(6 * i + 2 * d + 1)];
    if (p >= 0) s += c * x[(size_t)p];
    if (m >= 0) s += (-c) * x[(size_t)m];
    return s;
}

// Step2 + ComputeDivergence (:171-262): div = sum_d Dc[d] P(E Dc[d] u) / |.|
static void geodesic_div(const Tube& t, const CSR& E, const std::vector<std::array<double, 3>>& nrm, const Vec& u,
                         Vec& div) {
    const i64 N = t.N();
    const double c = 0.5 / t.dx;
    std::vector<Vec> g(3, Vec((size_t)N)), eg(3, Vec((size_t)N));
    for (int d = 0; d < 3; ++d) {
        for (i64 i = 0; i < N; ++i) g[(size_t)d][(size_t)i] = dc_row(t, i, d, c, u.data());
        E.matvec(g[(size_t)d].data(), eg[(size_t)d].data());
    }
    for (i64 i = 0; i < N; ++i) {
        double v[3] = {eg[0][(size_t)i], eg[1][(size_t)i], eg[2][(size_t)i]};
        const double* n = nrm[(size_t)i].data();
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += v[d] * n[d];
        for (int d = 0; d < 3; ++d) v[d] = v[d] - s * n[d];
        double q = 0.0;
        for (int d = 0; d < 3; ++d) q += v[d] * v[d];
        q = std::sqrt(q);
        if (q != 0.0)
            for (int d = 0; d < 3; ++d) v[d] /= q;
        for (int d = 0; d < 3; ++d) eg[(size_t)d][(size_t)i] = v[d];
    }
    std::fill(div.begin(), div.end(), 0.0);
    for (i64 i = 0; i < N; ++i) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += dc_row(t, i, d, c, eg[(size_t)d].data());
        div[(size_t)i] = s;
    }
}

std::array<int, 4> geodesic_run(const Tube& t, const IBCData& I, const CSR& E, const CSR& Lfull, const CSR& Efull,
                                const std::vector<uint8_t>& ident, double dt, int sweeps, Vec& phi) {
    const i64 N = t.N(), D = Lfull.rows;
    std::array<int, 4> iters{};
    const double scale = std::atanh(1.0 - t.dx) / t.r_t;       // :70-72
    Vec rhs((size_t)D, 0.0), uf((size_t)D), u((size_t)N), u2((size_t)N), div((size_t)D);
    auto cpd = [&](size_t c, size_t i) {
        const auto& v = I.dof[c].cp_diff[i];
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += v[d] * v[d];
        return std::sqrt(s);
    };
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i) {
            const i64 r = heat_dirichlet_row(I, c, i);
            if (r >= 0) rhs[(size_t)r] = smooth_h(scale, -cpd(c, i));
        }
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i)
            rhs[(size_t)I.dof[c].set_of_sub[i]] = smooth_h(scale, cpd(c, i));
    System S1, S3;
    for (System* S : {&S1, &S3}) {
        S->L = &Lfull;
        S->E = &Efull;
        S->ident = ident;
    }
    S1.m = 1.0;
    S1.n = -dt;
    S1.precond = PRECOND_JACOBI;
    S1.precond_iters = sweeps;
    S3.m = 0.0;
    S3.n = 1.0;
    S3.precond = PRECOND_DIAGONAL;
    iters[0] = bicgstab(S1, rhs.data(), uf.data(), 1e-10, 100000).iters;
    for (i64 i = 0; i < N; ++i) u[(size_t)i] = -uf[(size_t)i];
    const auto nrm = tube_jcp_normals(t);
    phi.assign((size_t)D, 0.0);
    for (int pass = 0; pass < 3; ++pass) {
        if (pass > 0) {
            for (i64 i = 0; i < N; ++i) u2[(size_t)i] = phi[(size_t)i];
            E.matvec(u2.data(), u.data());
        }
        geodesic_div(t, E, nrm, u, div);
        iters[1 + pass] = bicgstab(S3, div.data(), phi.data(), 1e-10, 100000).iters;
        double mn = std::numeric_limits<double>::max();
        for (double v : phi) mn = std::min(mn, v);
        for (double& v : phi) v -= mn;
    }
    return iters;
}
}  // namespace orc
