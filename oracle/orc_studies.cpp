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
    SolveInfo info = bicgstab(S, rhs.data(), u.data(), 1e-10, 100000);   // Solver.cpp:164
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
