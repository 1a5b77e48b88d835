// C entry points of the CPU oracle for the test-suite (ctypes).
// TEST INFRASTRUCTURE ONLY (see orc.h).  CSR arrays cross this boundary as
// int32 row pointers / columns, the layout of Eigen::SparseMatrix<double,
// RowMajor> and of the GPU engine's cpm_op_create_csr.
#include <omp.h>
#include <cstring>
#include <memory>
#include <string>

#include "orc.h"
#include "orc_ibc.h"

namespace orc {
double poisson_problem_error(int problem, double dx, int* iters);
double ibc_study_order(int problem, int g, int* iters_last, int order);
double heat_study_order(int method, int order, int btype);
double heat_problem_error(int method, double dx, int order, int btype);
int64_t heat_run_study(int method, double dx, int steps, int order, int btype, double* u_out, double* dt_out);
std::array<int, 4> geodesic_run(const Tube& t, const IBCData& I, const CSR& E, const CSR& Lfull, const CSR& Efull,
                                const std::vector<uint8_t>& ident, double dt, int sweeps, Vec& phi);
int64_t rd_run_study(int method, double dx, int order, int btype, const double* prm, int steps, double* u_io,
                     double* v_io);
void rd_run_problem(const CSR& Lfull, const CSR& Efull, const std::vector<uint8_t>& ident, const IBCData& I,
                    int method, const double* prm, int steps, double* u_io, double* v_io);
}

using namespace orc;

static thread_local std::string g_err;

namespace {
struct Handle {
    PoissonProblem P;              // tube, Lfull, Efull, ident (IBC problems fill the same fields)
    IBCData ibc;
    bool has_ibc = false;
    CSR extra;   // explicit operator or other matrix results
    Vec rhs_pts;                   // evaluation point of every row (3 per row)
    std::vector<int32_t> rhs_kind; // 0 tube, 1 ghost (f at cp_c), 2 identity (g at cp_c)
};

// Triangle mesh handed in by orc_set_mesh for the next "Triangulation" surface.
thread_local std::vector<Vec> g_mesh_v;
thread_local std::vector<std::vector<i64>> g_mesh_f;

SurfaceSpec make_spec(const char* type, int use_defaults, const double* params, int np, double bb0, double bb1,
                      int dim, const double* verts, int nv) {
    std::string ty(type);
    if (ty == "Triangulation") {   // SurfaceSpecifier.cpp:298-363
        if (g_mesh_f.empty()) throw std::runtime_error("orc_set_mesh first");
        SurfaceSpec s;
        s.type = ty;
        s.dim = 3;
        s.verts = g_mesh_v;
        s.faces = g_mesh_f;
        s.is_open = false;
        if (use_defaults) {        // SetSurface(): [min(0, coordinates), max(0, coordinates)]
            bb0 = 0.0;
            bb1 = 0.0;
            for (const auto& v : s.verts)
                for (double x : v) {
                    bb0 = std::min(bb0, x);
                    bb1 = std::max(bb1, x);
                }
        }
        s.bbox = {bb0, bb1};
        return s;
    }
    if (ty == "Polyline") {
        std::vector<Vec> v((size_t)nv, Vec((size_t)dim));
        for (int i = 0; i < nv; ++i)
            for (int d = 0; d < dim; ++d) v[(size_t)i][(size_t)d] = verts[i * dim + d];
        return SurfaceSpec::polyline(v, {bb0, bb1});
    }
    SurfaceSpec s = SurfaceSpec::defaults(ty);
    if (!use_defaults) {
        s.params.assign(params, params + np);
        s.bbox = {bb0, bb1};
        s.dim = dim;
    }
    return s;
}

void export_csr(const CSR& A, int32_t* ptr, int32_t* col, double* val) {
    for (size_t i = 0; i < A.ptr.size(); ++i) ptr[i] = (int32_t)A.ptr[i];
    for (size_t i = 0; i < A.col.size(); ++i) col[i] = (int32_t)A.col[i];
    std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
}

CSR import_csr(int64_t rows, const int32_t* ptr, const int32_t* col, const double* val) {
    CSR A;
    A.rows = A.cols = rows;
    A.ptr.assign(ptr, ptr + rows + 1);
    A.col.assign(col, col + ptr[rows]);
    A.val.assign(val, val + ptr[rows]);
    return A;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Closest points on the mesh of orc_set_mesh at n query points (Surface::cpTri,
// Surface.cpp:1751-1779: |signed distance|).
int orc_mesh_cp(const double* x, int64_t n, double* cp, double* dist) {
    try {
        SurfaceSpec s = make_spec("Triangulation", 1, nullptr, 0, 0, 0, 3, nullptr, 0);
        Surface S(s);
        for (int64_t i = 0; i < n; ++i) {
            i64 b;
            S.cp(x + 3 * i, cp + 3 * i, dist[i], b);
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_set_mesh(const double* v, int64_t nv, const int32_t* f, int64_t nf) {
    g_mesh_v.assign((size_t)nv, Vec(3));
    for (int64_t i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) g_mesh_v[(size_t)i][(size_t)d] = v[3 * i + d];
    g_mesh_f.assign((size_t)nf, std::vector<i64>(3));
    for (int64_t i = 0; i < nf; ++i)
        for (int k = 0; k < 3; ++k) g_mesh_f[(size_t)i][(size_t)k] = f[3 * i + k];
}

void* orc_problem_poisson(const char* type, int use_defaults, const double* params, int np, double bb0, double bb1,
                          int dim, const double* verts, int nv, double dx, int bdy_order, const int64_t* which_bdy,
                          int nwb, int p) {
    try {
        SurfaceSpec s = make_spec(type, use_defaults, params, np, bb0, bb1, dim, verts, nv);
        std::vector<i64> wb(which_bdy, which_bdy + nwb);
        auto h = std::make_unique<Handle>();
        h->P = make_poisson(s, dx, bdy_order, wb, p);
        return h.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// PoissonWithIBC(surface, ibc_specs, dx, p).  IBC i has type itypes[i],
// niparams[i] parameters (concatenated in iparams) or, for "Polyline",
// nverts[i] vertices (concatenated in iverts), boundary order/type and
// on_ibc_tolerance.
void* orc_problem_poisson_ibc(const char* type, int use_defaults, const double* params, int np, double bb0,
                              double bb1, int nibc, const char* const* itypes, const double* iparams,
                              const int* niparams, const double* iverts, const int64_t* nverts, const int* border,
                              const int* btype, const double* tol, double dx, int p) {
    try {
        SurfaceSpec s = make_spec(type, use_defaults, params, np, bb0, bb1, 3, nullptr, 0);
        std::vector<SurfaceSpec> specs;
        size_t po = 0, vo = 0;
        for (int c = 0; c < nibc; ++c) {
            SurfaceSpec is;
            std::string ty(itypes[c]);
            if (ty == "Polyline") {
                std::vector<Vec> v((size_t)nverts[c], Vec(3));
                for (int64_t i = 0; i < nverts[c]; ++i)
                    for (int d = 0; d < 3; ++d) v[(size_t)i][(size_t)d] = iverts[3 * (vo + i) + d];
                vo += (size_t)nverts[c];
                is = SurfaceSpec::polyline(v, {bb0, bb1});
            } else {
                is.type = ty;
                is.params.assign(iparams + po, iparams + po + niparams[c]);
                is.bbox = {bb0, bb1};
                is.dim = 3;
                is.is_open = ty == "Arc";   // SetIsOpen (SurfaceSpecifier.cpp:370-381)
            }
            po += (size_t)niparams[c];
            is.meta.boundary_order = border[c];
            is.meta.boundary_type = {btype[2 * c], btype[2 * c + 1]};
            is.meta.on_ibc_tolerance = tol[c];
            specs.push_back(is);
        }
        auto h = std::make_unique<Handle>();
        PoissonIBCProblem Q = make_poisson_ibc(s, specs, dx, p);
        h->P.tube = std::move(Q.tube);
        h->P.L = std::move(Q.L);
        h->P.E = std::move(Q.E);
        h->P.Lfull = std::move(Q.Lfull);
        h->P.Efull = std::move(Q.Efull);
        h->P.ident = std::move(Q.ident);
        h->ibc = std::move(Q.ibc);
        h->has_ibc = true;
        // evaluation points of the right-hand side, PoissonWithIBC::GetRHS (problems/PoissonWithIBC.cpp:50-90)
        const i64 D = h->P.Lfull.rows, N = h->P.tube.N();
        h->rhs_pts.assign((size_t)D * 3, 0.0);
        h->rhs_kind.assign((size_t)D, -1);
        for (i64 i = 0; i < N; ++i) {
            for (int d = 0; d < 3; ++d) h->rhs_pts[(size_t)(3 * i + d)] = h->P.tube.cpx[(size_t)(3 * i + d)];
            h->rhs_kind[(size_t)i] = 0;
        }
        const IBCData& I = h->ibc;
        for (size_t c = 0; c < I.dof.size(); ++c)
            for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i) {
                const int ty = I.meta[c].boundary_type[0];
                if (ty == 1 || (ty == 0 && I.meta[c].boundary_order == 2)) {
                    const i64 r = (i64)i + I.start[c];
                    for (int d = 0; d < 3; ++d) h->rhs_pts[(size_t)(3 * r + d)] = I.dof[c].cpc[i][d];
                    h->rhs_kind[(size_t)r] = 1;
                }
                if (ty == 0) {
                    const i64 r = I.meta[c].boundary_order == 2
                                      ? I.identity_rows_start + I.dof[c].dir_sub_of_set[i] + I.dir_start[c]
                                      : I.identity_rows_start + (i64)i + I.start[c];
                    for (int d = 0; d < 3; ++d) h->rhs_pts[(size_t)(3 * r + d)] = I.dof[c].cpc[i][d];
                    h->rhs_kind[(size_t)r] = 2;
                }
            }
        return h.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// ints: total IBC DOFs G, 2nd-order rows K, identity_rows_start, #IBCs, tagging size of IBC 0, DOF size of IBC 0
void orc_ibc_info(void* hp, int64_t* ints) {
    Handle* h = static_cast<Handle*>(hp);
    const IBCData& I = h->ibc;
    ints[0] = I.total;
    ints[1] = I.num2;
    ints[2] = I.identity_rows_start;
    ints[3] = (int64_t)I.dof.size();
    ints[4] = I.tag.empty() ? 0 : (int64_t)I.tag[0].set_of_sub.size();
    ints[5] = I.dof.empty() ? 0 : (int64_t)I.dof[0].set_of_sub.size();
}

// Surface::IBCParameter of n points on IBC c (vparam: one value per
// parameterisation point, or null for the arc-length rule).
int orc_ibc_parameter(void* hp, int c, const double* pts, int64_t n, const double* vparam, int64_t nv, double* out) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        const IBCData& I = h->ibc;
        if (c < 0 || c >= (int)I.surf.size()) throw std::runtime_error("no such IBC");
        Vec vp;
        if (vparam) vp.assign(vparam, vparam + nv);
        for (int64_t i = 0; i < n; ++i) out[i] = I.surf[(size_t)c].ibc_parameter(pts + 3 * i, vparam ? &vp : nullptr);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int64_t orc_ibc_subset_size(void* hp, int c, int which) {
    Handle* h = static_cast<Handle*>(hp);
    const IBCData& I = h->ibc;
    if (c < 0 || c >= (int)I.dof.size()) return -1;
    return which == 0 ? (int64_t)I.tag[(size_t)c].set_of_sub.size() : (int64_t)I.dof[(size_t)c].set_of_sub.size();
}

// Tagging / DOF subsets of IBC c: tube index of each member (in order), cp_c,
// directions (3 each).
void orc_ibc_subset(void* hp, int c, int which, int64_t* set_of_sub, double* cpc, double* dirs) {
    Handle* h = static_cast<Handle*>(hp);
    const IBCData& I = h->ibc;
    if (which == 0) {
        const TagSubset& T = I.tag[(size_t)c];
        for (size_t i = 0; i < T.set_of_sub.size(); ++i) {
            set_of_sub[i] = T.set_of_sub[i];
            for (int d = 0; d < 3; ++d) { cpc[3 * i + d] = T.cpc[i][d]; dirs[3 * i + d] = T.directions[i][d]; }
        }
    } else {
        const DofSubset& T = I.dof[(size_t)c];
        for (size_t i = 0; i < T.set_of_sub.size(); ++i) {
            set_of_sub[i] = T.set_of_sub[i];
            for (int d = 0; d < 3; ++d) { cpc[3 * i + d] = T.cpc[i][d]; dirs[3 * i + d] = T.directions[i][d]; }
        }
    }
}

void orc_rhs_points(void* hp, double* pts, int32_t* kind) {
    Handle* h = static_cast<Handle*>(hp);
    std::memcpy(pts, h->rhs_pts.data(), h->rhs_pts.size() * 8);
    std::memcpy(kind, h->rhs_kind.data(), h->rhs_kind.size() * 4);
}

double orc_ibc_study_order(int problem, int g, int order, int* iters) {
    try {
        return ibc_study_order(problem, g, iters, order);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1.0;
    }
}

double orc_heat_study_order(int method, int order, int btype) {
    try {
        return heat_study_order(method, order, btype);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1.0;
    }
}

double orc_heat_problem_error(int method, double dx, int order, int btype) {
    try {
        return heat_problem_error(method, dx, order, btype);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1.0;
    }
}

int64_t orc_heat_run(int method, double dx, int steps, int order, int btype, double* u_out, double* dt_out) {
    try {
        return heat_run_study(method, dx, steps, order, btype, u_out, dt_out);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int64_t orc_rd_run(int method, double dx, int order, int btype, const double* prm, int steps, double* u_io,
                   double* v_io) {
    try {
        return rd_run_study(method, dx, order, btype, prm, steps, u_io, v_io);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Gray-Scott steps (method 0 the reference's explicit step, 1 implicit
// diffusion) on an IBC problem built by orc_problem_poisson_ibc, Dirichlet
// values 0; u_io / v_io hold D values each and are updated in place.
int orc_rd_run_problem(void* hp, int method, const double* prm, int steps, double* u_io, double* v_io) {
    try {
        auto* h = static_cast<Handle*>(hp);
        if (!h->has_ibc) throw std::runtime_error("not an IBC problem");
        rd_run_problem(h->P.Lfull, h->P.Efull, h->P.ident, h->ibc, method, prm, steps, u_io, v_io);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// GeodesicDistance::ComputeDistance on an IBC problem: phi (D rows), iterations of the four solves.
int orc_geodesic(void* hp, double dt, int sweeps, double* phi, int32_t* iters) {
    try {
        auto* h = static_cast<Handle*>(hp);
        if (!h->has_ibc) throw std::runtime_error("the geodesic distance needs an IBC problem");
        Vec p;
        const auto it = geodesic_run(h->P.tube, h->ibc, h->P.E, h->P.Lfull, h->P.Efull, h->P.ident, dt, sweeps, p);
        std::copy(p.begin(), p.end(), phi);
        for (int k = 0; k < 4; ++k) iters[k] = it[(size_t)k];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_free(void* h) { delete static_cast<Handle*>(h); }

// ints: N, D, nnz(Lfull), nnz(Efull), n, N1d, dim, p, nnz(extra); dbl: xstart, r_t, rfac, dx
void orc_info(void* hp, int64_t* ints, double* dbl) {
    Handle* h = static_cast<Handle*>(hp);
    const Tube& t = h->P.tube;
    ints[0] = t.N();
    ints[1] = h->P.Lfull.rows;
    ints[2] = h->P.Lfull.nnz();
    ints[3] = h->P.Efull.nnz();
    ints[4] = t.n;
    ints[5] = t.N1d;
    ints[6] = t.dim;
    ints[7] = t.p;
    ints[8] = h->extra.nnz();
    dbl[0] = t.xstart;
    dbl[1] = t.r_t;
    dbl[2] = t.rfac;
    dbl[3] = t.dx;
}

void orc_tube(void* hp, int64_t* rec, double* x, double* cp, double* dist, int64_t* bdy, int64_t* nbr, int64_t* bpt) {
    const Tube& t = static_cast<Handle*>(hp)->P.tube;
    const size_t N = (size_t)t.N();
    std::memcpy(rec, t.rec.data(), N * 8);
    std::memcpy(x, t.x.data(), N * t.dim * 8);
    std::memcpy(cp, t.cpx.data(), N * t.dim * 8);
    std::memcpy(dist, t.dist.data(), N * 8);
    std::memcpy(bdy, t.bdy.data(), N * 8);
    std::memcpy(nbr, t.nbr.data(), N * 6 * 8);
    std::memcpy(bpt, t.bpt.data(), N * 8);
}

// which: 0 Lfull, 1 Efull, 2 extra
void orc_csr(void* hp, int which, int32_t* ptr, int32_t* col, double* val) {
    Handle* h = static_cast<Handle*>(hp);
    export_csr(which == 0 ? h->P.Lfull : which == 1 ? h->P.Efull : h->extra, ptr, col, val);
}

void orc_ident(void* hp, uint8_t* ident) {
    Handle* h = static_cast<Handle*>(hp);
    std::memcpy(ident, h->P.ident.data(), h->P.ident.size());
}

// Explicit m I + n M with identity rows into the handle's "extra" matrix.
int orc_build_explicit(void* hp, double m, double n, int nD) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        h->extra = explicit_operator(h->P.Lfull, h->P.Efull, m, n, h->P.ident, nD);
        return (int)h->extra.nnz();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Plot interpolation matrix at the surface parameterisation points into "extra".
// Plotting rows at nq points with the curves' ghost columns (AddIBCPlotInterpolation).
int orc_build_ibc_plot_E(void* hp, const double* xq, int64_t nq) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        if (!h->has_ibc) throw std::runtime_error("not an IBC problem");
        std::vector<Vec> xp((size_t)nq, Vec(3));
        for (int64_t i = 0; i < nq; ++i)
            for (int d = 0; d < 3; ++d) xp[(size_t)i][(size_t)d] = xq[3 * i + d];
        h->extra = ibc_plot_E_at(h->P.tube, h->ibc, xp);
        return (int)h->extra.rows;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int orc_build_plot_E(void* hp) {
    try {
        Handle* h = static_cast<Handle*>(hp);
        h->extra = plot_E(h->P.tube);
        return (int)h->extra.rows;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int orc_multiply(int64_t D, int nD, const int32_t* Lp, const int32_t* Lc, const double* Lv, const int32_t* Ep,
                 const int32_t* Ec, const double* Ev, const uint8_t* ident, double m, double n, const double* x,
                 double* y) {
    try {
        CSR L = import_csr(D, Lp, Lc, Lv), E = import_csr(D, Ep, Ec, Ev);
        System S;
        S.m = m;
        S.n = n;
        S.L = &L;
        S.E = &E;
        S.ident.assign(ident, ident + D);
        S.nD = nD;
        S.multiply(x, y);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int orc_solve(int64_t D, int nD, const int32_t* Lp, const int32_t* Lc, const double* Lv, const int32_t* Ep,
              const int32_t* Ec, const double* Ev, const uint8_t* ident, double m, double n, int precond,
              int precond_iters, const double* b, double* x, double tol, int max_it, int* iters, double* rel_res) {
    try {
        CSR L = import_csr(D, Lp, Lc, Lv), E = import_csr(D, Ep, Ec, Ev);
        System S;
        S.m = m;
        S.n = n;
        S.L = &L;
        S.E = &E;
        S.ident.assign(ident, ident + D);
        S.nD = nD;
        S.precond = precond;
        S.precond_iters = precond_iters;
        SolveInfo info = bicgstab(S, b, x, tol, max_it);
        if (iters) *iters = info.iters;
        if (rel_res) *rel_res = info.rel_res;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_set_parallel(int vec, int apply, int threads) {
    g_parallel.vec = vec != 0;
    g_parallel.apply = apply != 0;
    if (threads > 0) omp_set_num_threads(threads);
}

double orc_poisson_problem_error(int problem, double dx, int* iters) {
    try {
        return poisson_problem_error(problem, dx, iters);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1.0;
    }
}

void orc_weights_1d(const double* xs, double xq, int S, double* w) { weights_1d(xs, xq, S, w); }

}  // extern "C"
