// C entry points of the CPU oracle for the test-suite (ctypes).
// TEST INFRASTRUCTURE ONLY (see orc.h).  CSR arrays cross this boundary as
// int32 row pointers / columns, the layout of Eigen::SparseMatrix<double,
// RowMajor> and of the GPU engine's cpm_op_create_csr.
#include <cstring>
#include <memory>
#include <string>

#include "orc.h"

namespace orc {
double poisson_problem_error(int problem, double dx, int* iters);
}

using namespace orc;

static thread_local std::string g_err;

namespace {
struct Handle {
    PoissonProblem P;
    CSR extra;   // explicit operator or other matrix results
};

SurfaceSpec make_spec(const char* type, int use_defaults, const double* params, int np, double bb0, double bb1,
                      int dim, const double* verts, int nv) {
    std::string ty(type);
    if (ty == "Polyline") {
        std::vector<Vec> v((size_t)nv, Vec((size_t)dim));
        for (int i = 0; i < nv; ++i)
            for (int d = 0; d < dim; ++d) v[(size_t)i][(size_t)d] = verts[i * dim + d];
        return SurfaceSpec::polyline(v, {bb0, bb1});
    }
    if (ty == "Triangulation") throw std::runtime_error("use orc_tube_mesh for triangulations");
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
