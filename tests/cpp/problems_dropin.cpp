// The reference's problem classes, called as its drivers call them, on the
// MI355X engine through the C++ facade (cpm-ibc_amd/host/cpm_solver.h):
//   Poisson + Solver::Init(m, n, GetLfull(), GetEfull(), diagonal, GetIdentityRows())
//     (convergence_studies/PoissonProblems/PoissonProblems.cpp),
//   PoissonWithIBC + GetRHS(f, exact) + Solver (PoissonProblemsWithIBC.cpp:230-256),
//   HeatWithIBC::TimeStep (HeatProblemsWithIBC.cpp:90-193),
//   ReactionDiffusionWithIBC::TimeStep (examples/ReactionDiffusionFish),
//   GeodesicDistance::ComputeDistance (problems/GeodesicDistance.cpp:11-67).
// Writes each result as raw doubles into <dir> for the test to compare with
// the oracle, and one JSON line of statuses.
//
// usage: problems_dropin <dir>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "cpm_solver.h"

using cpm::Scalar;
using cpm::VectorX;

static void write_bin(const std::string& path, const std::vector<double>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

// The 200 x 200 (theta, phi) points of the unit sphere's parameterisation (Surface.cpp:1973-1999).
static std::vector<std::vector<Scalar>> sphere_points() {
    std::vector<std::vector<Scalar>> xp;
    const int n = 200;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double th = i * (2 * M_PI / (n - 1)), ph = j * (M_PI / (n - 1));
            xp.push_back({std::sin(ph) * std::cos(th), std::sin(ph) * std::sin(th), std::cos(ph)});
        }
    return xp;
}

// The circle of PoissonProblemsWithIBC.cpp:117-138 for the seeded normal (0, 0.25, 0.55).
static std::vector<Scalar> study_circle() {
    double nr[3] = {0.0, 0.25, 0.55};
    const double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    for (double& c : nr) c /= nn;
    const double r = 0.2, th = 0.1, ph = 0.2;
    const double pp[3] = {r * std::cos(ph) * std::sin(th), r * std::sin(ph) * std::sin(th), r * std::cos(th)};
    const double rho = -(pp[0] * nr[0] + pp[1] * nr[1] + pp[2] * nr[2]);
    return {std::sqrt(1 - rho * rho), rho * nr[0], rho * nr[1], rho * nr[2], nr[0], nr[1], nr[2]};
}

// The heat study's Dirichlet circle (HeatProblemsWithIBC.cpp:94-111).
static std::vector<Scalar> heat_circle() {
    double nr[3] = {1.1, 1.234, 1.321};
    const double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    for (double& c : nr) c /= nn;
    const double r = 0.27, th = 0.132, ph = 0.221;
    const double pp[3] = {r * std::cos(ph) * std::sin(th), r * std::sin(ph) * std::sin(th), r * std::cos(th)};
    const double rho = -(pp[0] * nr[0] + pp[1] * nr[1] + pp[2] * nr[2]);
    return {std::sqrt(1 - rho * rho), rho * nr[0], rho * nr[1], rho * nr[2], nr[0], nr[1], nr[2]};
}

static double harmonic(const std::vector<Scalar>& x) {   // PoissonProblemsWithIBC.cpp:243-244
    const double theta = std::atan2(x[1], x[0]), phi = std::acos(x[2] / 1.0);
    return std::cos(3 * theta) * std::pow(std::sin(phi), 3) * (9 * std::pow(std::cos(phi), 2) - 1);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string d = argv[1];
    try {
        const auto xp = sphere_points();
        const double dx = 0.1;

        // ---- Poisson on the unit sphere: (1 - Laplace) u = 3z + 21xy, u = z + 3xy
        cpm::SurfaceSpecifier sphere("Sphere");
        cpm::Poisson poisson(sphere, dx);
        auto f1 = [](const std::vector<Scalar>& x) { return 3 * x[2] + 21 * x[0] * x[1]; };
        VectorX rhs1 = poisson.GetRHS(f1);
        cpm::Solver sv1;
        sv1.Init(1.0, -1.0, poisson.GetLfull(), poisson.GetEfull(), cpm::diagonal, poisson.GetIdentityRows());
        VectorX u1 = sv1.Solve(rhs1);
        const int st1 = sv1.LastInfo().status;
        VectorX up1 = cpm::Apply(poisson.GetInterpMatrix(xp), u1);
        double err1 = 0.0;
        for (size_t i = 0; i < xp.size(); ++i)
            err1 = std::max(err1, std::fabs(up1[i] - (xp[i][2] + 3 * xp[i][0] * xp[i][1])));
        write_bin(d + "/poisson_rhs.bin", rhs1);
        write_bin(d + "/poisson_u.bin", u1);

        // ---- PoissonWithIBC: sphere + Circle3D, 2nd-order Dirichlet, -Laplace u = 30 Y
        std::vector<cpm::SurfaceSpecifier> ibc{cpm::SurfaceSpecifier("Circle3D", {-1.0, 1.0}, study_circle(), 3)};
        cpm::IBCMetaData meta;
        meta.boundary_order = 2;
        meta.on_ibc_tolerance = 0.1 * dx * dx;
        ibc[0].SetIBCMetaData(meta);
        cpm::PoissonWithIBC pibc(sphere, ibc, dx, 3);
        auto f2 = [](const std::vector<Scalar>& x) { return 30 * harmonic(x); };
        VectorX rhs2 = pibc.GetRHS(f2, cpm::ScalarFunction(harmonic));
        cpm::Solver sv2;
        sv2.Init(pibc.Operator(0.0, -1.0));
        VectorX u2 = sv2.Solve(rhs2);
        const int st2 = sv2.LastInfo().status;
        VectorX up2 = cpm::Apply(pibc.GetPlotInterpMatrix(xp), u2);
        double err2 = 0.0;
        for (size_t i = 0; i < xp.size(); ++i) err2 = std::max(err2, std::fabs(up2[i] - harmonic(xp[i])));
        write_bin(d + "/ibc_rhs.bin", rhs2);
        write_bin(d + "/ibc_u.bin", u2);
        write_bin(d + "/ibc_uplot.bin", up2);

        // ---- HeatWithIBC: backward Euler, 5 steps of the study's system (dt = 0.1 dx)
        std::vector<cpm::SurfaceSpecifier> hibc{cpm::SurfaceSpecifier("Circle3D", {-1.0, 1.0}, heat_circle(), 3)};
        hibc[0].SetIBCMetaData(meta);
        const double dt = 0.1 / std::ceil(0.1 / (0.1 * dx));
        cpm::HeatWithIBC heat(sphere, hibc, dx, dt, 1, 3);
        auto fh = [](const std::vector<Scalar>& x) { return std::cos(std::acos(x[2] / 1.0)); };
        std::vector<double> pts;
        std::vector<int32_t> kind;
        heat.problem().RHSPoints(pts, kind);
        VectorX uh(kind.size());
        for (size_t r = 0; r < kind.size(); ++r) uh[r] = fh({pts[3 * r], pts[3 * r + 1], pts[3 * r + 2]});
        int sth = 1;
        for (int k = 0; k < 5; ++k) {
            const double t = (k + 1) * dt;
            uh = heat.TimeStep(uh, [&](size_t&, const size_t&, const std::vector<Scalar>& x) {
                return std::exp(-2 * t) * fh(x);
            });
            if (heat.LastInfo().status != CPM_SOLVE_CONVERGED) sth = heat.LastInfo().status;
        }
        write_bin(d + "/heat_u.bin", uh);

        // ---- ReactionDiffusionWithIBC: 20 explicit steps, Dirichlet values 0
        cpm::ReactionDiffusionWithIBC rd(sphere, hibc, dx, 0.0, 3);
        cpm_rd_params def;
        cpm_rd_params_default(dx, &def);
        rd.Params().dt = def.dt;
        rd.problem().RHSPoints(pts, kind);
        VectorX ur(kind.size(), 0.0), vr(kind.size(), 0.0);
        for (size_t r = 0; r < kind.size(); ++r) {
            if (kind[r] != 0) continue;
            const double x = pts[3 * r], y = pts[3 * r + 1], z = pts[3 * r + 2];
            const double pert = 0.5 * std::exp(-std::pow(10.0 * (z - 0.1), 2)) +
                                0.25 * (1 + std::sin(5 * x) * std::cos(7 * y - 2 * z));
            ur[r] = 1.0 - pert;
            vr[r] = 0.5 * pert;
        }
        auto zero = [](size_t&, const size_t&, const std::vector<Scalar>&) { return 0.0; };
        for (int k = 0; k < 20; ++k) rd.TimeStep(ur, vr, zero, zero);
        write_bin(d + "/rd_u.bin", ur);
        write_bin(d + "/rd_v.bin", vr);

        // ---- GeodesicDistance from the north pole (a Point curve, 1st-order Dirichlet)
        std::vector<cpm::SurfaceSpecifier> gibc{cpm::SurfaceSpecifier("Point", {-1.0, 1.0}, {0.0, 0.0, 1.0}, 3)};
        cpm::IBCMetaData m1;
        m1.boundary_order = 1;
        m1.on_ibc_tolerance = 0.1 * dx * dx;
        gibc[0].SetIBCMetaData(m1);
        cpm::PoissonWithIBC gp(sphere, gibc, dx, 3);
        cpm::GeodesicDistance geo;
        VectorX dist = geo.ComputeDistance(gp, dx * dx, xp);
        double gerr = 0.0;
        for (size_t i = 0; i < xp.size(); ++i) gerr = std::max(gerr, std::fabs(dist[i] - std::acos(xp[i][2])));
        int stg = 1;
        for (const auto& i : geo.Infos())
            if (i.status != CPM_SOLVE_CONVERGED) stg = i.status;
        write_bin(d + "/geo_phi.bin", geo.Phi());

        // ---- IBCParameter of the circle problem's Dirichlet points (normalised arc length)
        std::vector<double> p2;
        std::vector<int32_t> k2;
        pibc.RHSPoints(p2, k2);
        std::vector<std::vector<Scalar>> on_curve;
        for (size_t r = 0; r < k2.size(); ++r)
            if (k2[r] == 2) on_curve.push_back({p2[3 * r], p2[3 * r + 1], p2[3 * r + 2]});
        const std::vector<Scalar> par = pibc.IBCParameters(0, on_curve);
        write_bin(d + "/ibc_param.bin", par);
        const Scalar par0 = pibc.IBCParameter(0, on_curve[0]);

        std::printf("{\"poisson_status\": %d, \"poisson_err\": %.6e, \"ibc_status\": %d, \"ibc_err\": %.6e, "
                    "\"heat_status\": %d, \"geo_status\": %d, \"geo_err\": %.6e, \"param0\": %.17g}\n",
                    st1, err1, st2, err2, sth, stg, gerr, par0);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
