// CPU oracle for the closest-point-method screened-Poisson path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++17 restatement of the
// reference algorithm (nathandking/cpm-ibc), written from its sources, used
// solely as the checker for the GPU engine: only tests/, the smoke check and
// the CPU-baseline leg of bench.py load it.  Nothing in the product library
// links, calls or falls back to it.
//
// Each routine cites the reference file:line it restates.  Arithmetic is kept
// in the reference's order and built with -ffp-contract=off so that FP64
// values follow the same rounding sequence as the reference build (which has
// no -march flag, hence no FMA: CMakeLists.txt:1-43).
//
// Parity pins (see tests/): tube radius / grid closed forms; brute-force tube
// membership; E row sums and exact cubic reproduction; L exact on quadratics;
// operator vs explicit CSR product; BiCGStab residual; the reference's own
// convergence-order known answers (convergence_studies/*).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>
#include <functional>

namespace orc {

using i64 = int64_t;
using Vec = std::vector<double>;

struct Trip { i64 r, c; double v; };

// Row-major CSR with ascending columns inside each row (what Eigen's
// setFromTriplets + RowMajor produces; duplicates are summed).
struct CSR {
    i64 rows = 0, cols = 0;
    std::vector<i64> ptr{0};
    std::vector<i64> col;
    Vec val;
    static CSR from_triplets(i64 rows, i64 cols, std::vector<Trip> t);
    i64 nnz() const { return (i64)col.size(); }
    double coeff(i64 r, i64 c) const;           // 0 if absent
    void matvec(const double* x, double* y) const;   // y = A x (row order, column order)
    void prune_zeros();
};

// IBCMetaData, src/IBCMetaData.h:9-24.  on_ibc_tolerance is uninitialised in
// the reference; every caller sets it, we default it to 0.
struct IBCMeta {
    bool is_oriented = false;
    Vec point_on_side1;
    std::vector<int> boundary_type{0, 0};   // 0 Dirichlet, 1 Neumann
    int boundary_order = 1;
    double on_ibc_tolerance = 0.0;
    bool use_cp_diff_directions = false;
    bool use_Jcp_surface_normals = true;
    double ibc_DOFSubset_radius_scale = 1.0;
};

// SurfaceSpecifier, src/SurfaceSpecifier.cpp:65-256 (analytic defaults) and
// :298-343 (polyline / triangulation from vertices + faces).
struct SurfaceSpec {
    std::string type;
    Vec params;
    Vec bbox;           // {lo, hi}, same for every axis
    int dim = 3;
    std::vector<Vec> verts;                 // Polyline / Triangulation
    std::vector<std::vector<i64>> faces;
    bool is_open = false;
    IBCMeta meta;
    static SurfaceSpec defaults(const std::string& type);           // :65-242
    static SurfaceSpec polyline(const std::vector<Vec>& v, const Vec& bbox);  // :298-343
};

// Surface, src/Surface.cpp.  Closest point + parameterisation points used for
// plotting / error measurement.
// Exact closest point on the ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1 (abc = a, b, c).
void cp_ellipsoid(const double* abc, const double* x, double* cp);

class Surface {
public:
    Surface() = default;
    explicit Surface(const SurfaceSpec& s);
    void cp(const double* x, double* cpx, double& dist, i64& bdy) const;   // :61-190
    void cpbar(const double* x, double* cpx, double& dist) const;          // :399-409
    // :2731-2788; with vparam (one value per parameterisation point), the
    // nearest point's value (the rule for parameterised samples, :2737-2751)
    double ibc_parameter(const double* x, const Vec* vparam = nullptr) const;
    const SurfaceSpec& spec() const { return s_; }
    int dim() const { return s_.dim; }
    std::vector<Vec> xp;        // parameterisation points (m_mesh.vertices)
    Vec thetap, phip;
private:
    SurfaceSpec s_;
    void parameterise();        // :193-366
    void cp_polyline(const double* x, double* cpx, double& dist, i64& bdy) const;
    void cp_arc(const double* x, double* cpx, double& dist, i64& bdy) const;
    void cp_triangles(const double* x, double* cpx, double& dist) const;
    std::vector<i64> nearest_two_xp(const double* x, double& min_dist, size_t np) const;  // :2661-2728
};

// Tube + Grid + Node, src/Tube.cpp, src/Grid.cpp, src/Node.h.
struct Tube {
    int dim = 3, p = 3;
    double dx = 0;
    i64 n = 0, N1d = 0;
    double xstart = 0;          // same on every axis (Tube.cpp:154)
    double rfac = 0, r_t = 0;
    Surface surf;
    std::vector<int32_t> grid;  // dense rec -> tube index; -1 unvisited, -2 visited outside
    std::vector<i64> rec;       // tube -> rec (linear grid index)
    Vec x, cpx, dist;           // N*dim, N*dim, N
    std::vector<i64> bdy;
    std::vector<i64> nbr;       // N*6, order +x,-x,+y,-y,+z,-z; -1 = none (Node.h:23)
    std::vector<i64> bpt;       // interpolation base point (tube index)
    i64 N() const { return (i64)rec.size(); }
    i64 lin(const i64* t) const;                        // Grid.cpp:123-144
    void trip(i64 l, i64* t) const;                     // Grid.cpp:96-120
    i64 tube_index(i64 l) const { return grid[(size_t)l]; }
    void base_triplet(const double* xq, i64* t) const;  // Tube.cpp:502-531
    std::vector<i64> base_points(const std::vector<Vec>& xq) const;  // :464-483
    i64 nearest_tube_index(const double* xq) const;     // Tube.h:105-110
};

Tube build_tube(const SurfaceSpec& s, int p, double dx);    // Tube.cpp:115-461

// Interpolation, src/Interpolation.cpp.  E rows at query points xq; columns
// mapped through subset_of (tube index -> column, -1 drops the entry).
CSR build_E(const Tube& t, const std::vector<Vec>& xq, const std::vector<i64>* subset_of = nullptr,
            i64 ncols = -1);
CSR build_E_tube(const Tube& t);                            // Helpers.cpp:395-399
CSR build_L(const Tube& t);                                 // FDMatrices.cpp:11-49
CSR build_L_subset(const Tube& t, const std::vector<i64>& set_of_sub, const std::vector<i64>& sub_of_set);
std::vector<CSR> build_grad_subset(const Tube& t, const std::vector<i64>& set_of_sub,
                                   const std::vector<i64>& sub_of_set);    // FDMatrices.cpp:66-92
void weights_1d(const double* xs, double xq, int S, double* w);            // Interpolation.cpp:162-200

// GeneralizedBiCGStabSystem + BiCGStab + Solver (src/*.h).
enum Precond { PRECOND_IDENTITY = 0, PRECOND_DIAGONAL = 1, PRECOND_JACOBI = 2 };   // Defines.h:13-18

struct System {
    double m = 0, n = 0;
    const CSR* L = nullptr;
    const CSR* E = nullptr;
    std::vector<uint8_t> ident;
    int nD = 0;
    int precond = PRECOND_DIAGONAL;
    int precond_iters = 0;
    i64 dofs() const { return L->rows - nD; }
    void multiply(const double* x, double* y, double mm, double nn) const;   // GeneralizedBiCGStabSystem.h:213-233
    void multiply(const double* x, double* y) const { multiply(x, y, m, n); }
    void diag_inverse(const double* in, double* out) const;                  // :281-340
    void precondition(const double* in, double* out) const;                  // :350-385
};

struct SolveInfo { int iters = 0; int restarts = 0; double rel_res = 0; };

// Host-thread parallelism of the restated solver, for timing the CPU baseline
// only (results do not depend on it: every row and element is computed by one
// thread in the same order; inner products stay serial, as Eigen's dot is).
//   vec:   BiCGStab vector updates in OpenMP loops (BiCGStabVector.h:29-71, as shipped)
//   apply: operator apply and preconditioner in OpenMP loops (what -DOPENMP
//          enables in GeneralizedBiCGStabSystem.h:71-72, off as shipped)
struct Parallel { bool vec = false, apply = false; };
extern Parallel g_parallel;
// BiCGStab.h:30-108 with Solver.cpp:69-114 (x0 = 0, min 1, max_it).
SolveInfo bicgstab(const System& sys, const double* b, double* x, double tol, int max_it);

// Explicit operator m I + n (diag L + (L - diag L) E) with identity rows,
// Helpers.cpp:107-115 + Solver.cpp:122-138.  Independent of System::multiply.
CSR explicit_operator(const CSR& L, const CSR& E, double m, double n,
                      const std::vector<uint8_t>& ident, int nD);

// Helpers.cpp:403-442
void set_identity_rows(CSR& A, const std::vector<uint8_t>& ident);

// Poisson problem, problems/Poisson.cpp:5-238 (p = 3).
struct PoissonProblem {
    Tube tube;
    CSR L, E, Lfull, Efull;
    std::vector<uint8_t> ident;
    int bdy_order = 0;
    std::vector<i64> which_bdy;
};
PoissonProblem make_poisson(const SurfaceSpec& s, double dx, int bdy_order = 0,
                            std::vector<i64> which_bdy = {}, int p = 3);

// Helpers.cpp:446-463
double average_order(const Vec& err, const Vec& dx);

// Plotting interpolation matrix at the surface parameterisation points,
// Helpers.cpp:97-103.
CSR plot_E(const Tube& t);

}  // namespace orc
