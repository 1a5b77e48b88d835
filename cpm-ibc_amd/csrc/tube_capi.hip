// C ABI of the GPU tube (cpm_tube_*) and of operators built from it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cpm_gpu.h"
#include "ops.h"
#include "tube.h"
#include "ibc.h"

using namespace cpmgpu;

namespace cpmgpu {
void set_last_error(const std::string& s);
int status_of_current_exception();
}

namespace {
inline Ctx* C(cpm_ctx* c) { return reinterpret_cast<Ctx*>(c); }
inline Tube* T(cpm_tube* t) { return reinterpret_cast<Tube*>(t); }
inline const Tube* T(const cpm_tube* t) { return reinterpret_cast<const Tube*>(t); }
inline Problem* PR(cpm_problem* p) { return reinterpret_cast<Problem*>(p); }
inline const Problem* PR(const cpm_problem* p) { return reinterpret_cast<const Problem*>(p); }

// ELL (slot-major) -> CSR with ascending columns, duplicates summed, zeros dropped.
void ell_to_csr(int64_t D, int w, const std::vector<int>& col, const std::vector<double>& val, const double* diag,
                std::vector<int32_t>& ptr, std::vector<int32_t>& c, std::vector<double>& v) {
    ptr.assign(D + 1, 0);
    c.clear();
    v.clear();
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t r = 0; r < D; ++r) {
        row.clear();
        for (int s = 0; s < w; ++s) row.push_back({col[(size_t)s * D + r], val[(size_t)s * D + r]});
        if (diag) row.push_back({(int32_t)r, diag[r]});
        std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size();) {
            size_t e = k;
            double acc = row[k].second;
            while (e + 1 < row.size() && row[e + 1].first == row[k].first) acc += row[++e].second;
            if (acc != 0.0) {
                c.push_back(row[k].first);
                v.push_back(acc);
            }
            k = e + 1;
        }
        ptr[r + 1] = (int32_t)c.size();
    }
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return CPM_OK;
    } catch (...) {
        return status_of_current_exception();
    }
}

// Interpolation rows at device points -> host CSR with ascending columns
// (Eigen's setFromTriplets order).
void interp_csr(const Tube& t, const double* d_xq, int64_t nq, int32_t* rowptr, int32_t* col, double* val) {
    const int S = t.p + 1, W = S * S * S;
    const hipStream_t st = t.ctx->stream;
    int* ecol = reinterpret_cast<int*>(t.ctx->workspace("exp_ecol", (size_t)W * nq * 4));
    double* eval = t.ctx->workspace("exp_eval", (size_t)W * nq * 8);
    build_E_ell(t, d_xq, nq, ecol, eval, nullptr, nq, st);
    std::vector<int> hc((size_t)W * nq);
    std::vector<double> hv((size_t)W * nq);
    HIPCHK(hipMemcpyAsync(hc.data(), ecol, hc.size() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hv.data(), eval, hv.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int> ord(W);
    for (int64_t q = 0; q < nq; ++q) {
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return hc[(size_t)a * nq + q] < hc[(size_t)b * nq + q]; });
        rowptr[q] = (int32_t)(q * W);
        for (int s = 0; s < W; ++s) {
            col[q * W + s] = hc[(size_t)ord[s] * nq + q];
            val[q * W + s] = hv[(size_t)ord[s] * nq + q];
        }
    }
    rowptr[nq] = (int32_t)(nq * W);
}
}  // namespace

extern "C" {

int cpm_tube_create(cpm_ctx* ctx, const char* surface_type, const double* params, int nparams, double bbox_lo,
                    double bbox_hi, int dim, const double* verts, int64_t nverts, int p, double dx, cpm_tube** out) {
    return guard([&] {
        if (!ctx || !out || !surface_type) throw ArgError("null argument");
        HIPCHK(hipSetDevice(C(ctx)->device));
        SurfaceHost s = surface_from_desc(surface_type, params, nparams, bbox_lo, bbox_hi, dim, verts, nverts);
        *out = reinterpret_cast<cpm_tube*>(tube_build(C(ctx), s, p, dx));
    });
}

int cpm_tube_create_mesh(cpm_ctx* ctx, const double* verts, int64_t nverts, const int32_t* faces, int64_t nfaces,
                         double bbox_lo, double bbox_hi, int p, double dx, cpm_tube** out) {
    return guard([&] {
        if (!ctx || !out) throw ArgError("null argument");
        HIPCHK(hipSetDevice(C(ctx)->device));
        SurfaceHost s = surface_mesh(verts, nverts, faces, nfaces, bbox_lo, bbox_hi);
        *out = reinterpret_cast<cpm_tube*>(tube_build(C(ctx), s, p, dx));
    });
}

int cpm_tube_destroy(cpm_tube* t) {
    return guard([&] { delete T(t); });
}

int cpm_tube_get_info(const cpm_tube* tp, cpm_tube_info* info) {
    return guard([&] {
        if (!tp || !info) throw ArgError("null argument");
        const Tube& t = *T(tp);
        info->nodes = t.N;
        info->n = t.n;
        info->N1d = t.N1d;
        std::vector<int32_t> nb((size_t)t.N * 6);
        HIPCHK(hipMemcpy(nb.data(), t.nbr.p, nb.size() * 4, hipMemcpyDeviceToHost));
        int64_t c = t.N;
        for (int32_t v : nb) c += v >= 0;
        info->nnz_L = c;
        info->levels = t.levels;
        info->xstart = t.xstart;
        info->dx = t.dx;
        info->tube_radius = t.r_t;
        info->tube_radius_factor = t.rfac;
        info->build_ms = t.build_ms;
        info->dim = t.dim;
        info->p = t.p;
    });
}

int cpm_tube_export(const cpm_tube* tp, int64_t* rec, double* x, double* cpx, double* dist, int32_t* bdy,
                    int32_t* nbr, int32_t* bpt) {
    return guard([&] {
        if (!tp) throw ArgError("null tube");
        const Tube& t = *T(tp);
        const size_t N = (size_t)t.N;
        std::vector<int64_t> r(N);
        HIPCHK(hipMemcpy(r.data(), t.rec.p, N * 8, hipMemcpyDeviceToHost));
        if (rec) std::copy(r.begin(), r.end(), rec);
        if (x)   // Grid::CoordinateFromTriplet (Grid.cpp:28-38)
            for (size_t i = 0; i < N; ++i) {
                const int64_t l = r[i], k = l % t.N1d, j = (l / t.N1d) % t.N1d, ii = l / (t.N1d * t.N1d);
                x[3 * i + 0] = t.xstart + ii * t.dx;
                x[3 * i + 1] = t.xstart + j * t.dx;
                x[3 * i + 2] = t.xstart + k * t.dx;
            }
        if (cpx) HIPCHK(hipMemcpy(cpx, t.cpx.p, N * 24, hipMemcpyDeviceToHost));
        if (dist) HIPCHK(hipMemcpy(dist, t.dist.p, N * 8, hipMemcpyDeviceToHost));
        if (bdy) HIPCHK(hipMemcpy(bdy, t.bdy.p, N * 4, hipMemcpyDeviceToHost));
        if (nbr) HIPCHK(hipMemcpy(nbr, t.nbr.p, N * 24, hipMemcpyDeviceToHost));
        if (bpt) HIPCHK(hipMemcpy(bpt, t.bpt.p, N * 4, hipMemcpyDeviceToHost));
    });
}

int cpm_tube_export_E(const cpm_tube* tp, int32_t* rowptr, int32_t* col, double* val) {
    return guard([&] {
        if (!tp || !rowptr || !col || !val) throw ArgError("null argument");
        const Tube& t = *T(tp);
        interp_csr(t, t.cpx.as<double>(), t.N, rowptr, col, val);
    });
}

int cpm_tube_interp_matrix(const cpm_tube* tp, const double* xq, int64_t nq, int32_t* rowptr, int32_t* col,
                           double* val) {
    return guard([&] {
        if (!tp || !xq || !rowptr || !col || !val || nq < 0) throw ArgError("null argument");
        const Tube& t = *T(tp);
        double* d = t.ctx->workspace("exp_xq", (size_t)std::max<int64_t>(nq, 1) * 24);
        HIPCHK(hipMemcpy(d, xq, (size_t)nq * 24, hipMemcpyHostToDevice));
        interp_csr(t, d, nq, rowptr, col, val);
    });
}

int cpm_tube_export_L(const cpm_tube* tp, int32_t* rowptr, int32_t* col, double* val) {
    return guard([&] {
        if (!tp || !rowptr || !col || !val) throw ArgError("null argument");
        const Tube& t = *T(tp);
        const size_t N = (size_t)t.N;
        std::vector<int32_t> nb(N * 6);
        HIPCHK(hipMemcpy(nb.data(), t.nbr.p, nb.size() * 4, hipMemcpyDeviceToHost));
        const double hh = t.dx * t.dx, dg = -6.0 / hh, off = 1.0 / hh;
        int64_t k = 0;
        for (size_t i = 0; i < N; ++i) {
            rowptr[i] = (int32_t)k;
            std::vector<std::pair<int32_t, double>> e{{(int32_t)i, dg}};
            for (int s = 0; s < 6; ++s)
                if (nb[6 * i + s] >= 0) e.push_back({nb[6 * i + s], off});
            std::sort(e.begin(), e.end());
            for (auto& pr : e) {
                col[k] = pr.first;
                val[k] = pr.second;
                ++k;
            }
        }
        rowptr[N] = (int32_t)k;
    });
}

int cpm_problem_create(cpm_tube* tp, int n_ibc, const cpm_ibc_desc* ibcs, cpm_problem** out) {
    return cpm_problem_create_on(nullptr, tp, n_ibc, ibcs, out);
}

int cpm_problem_create_on(cpm_ctx* bctx, cpm_tube* tp, int n_ibc, const cpm_ibc_desc* ibcs, cpm_problem** out) {
    return guard([&] {
        if (!tp || !out || n_ibc < 0 || (n_ibc > 0 && !ibcs)) throw ArgError("null argument");
        Tube* t = T(tp);
        HIPCHK(hipSetDevice(t->ctx->device));
        std::vector<IbcSpec> specs((size_t)n_ibc);
        for (int c = 0; c < n_ibc; ++c) {
            const cpm_ibc_desc& d = ibcs[c];
            if (!d.type) throw ArgError("IBC type is null");
            specs[c].surf = surface_from_desc(d.type, d.params, d.nparams, 0.0, 0.0, 3, d.verts, d.nverts);
            specs[c].order = d.boundary_order;
            specs[c].btype[0] = d.boundary_type[0];
            specs[c].btype[1] = d.boundary_type[1];
            specs[c].tol = d.on_ibc_tolerance;
            specs[c].dof_scale = d.dof_radius_scale > 0 ? d.dof_radius_scale : 1.0;
        }
        *out = reinterpret_cast<cpm_problem*>(problem_build(t, specs, C(bctx)));
    });
}

int cpm_problem_destroy(cpm_problem* p) {
    return guard([&] { delete PR(p); });
}

int cpm_problem_get_info(const cpm_problem* pp, cpm_problem_info* info) {
    return guard([&] {
        if (!pp || !info) throw ArgError("null argument");
        const Problem& P = *PR(pp);
        info->rows = P.D;
        info->tube_nodes = P.N;
        info->ibc_dofs = P.G;
        info->dirichlet_rows = P.K;
        info->identity_rows_start = P.identity_rows_start;
        info->tagging_size0 = P.ibc.empty() ? 0 : P.ibc[0].nt;
        info->dof_size0 = P.ibc.empty() ? 0 : P.ibc[0].nd;
        info->n_ibc = (int)P.ibc.size();
        info->build_ms = P.build_ms;
    });
}

int cpm_problem_rhs_points(const cpm_problem* pp, double* pts, int32_t* kind) {
    return guard([&] {
        if (!pp) throw ArgError("null problem");
        const Problem& P = *PR(pp);
        if (pts) HIPCHK(hipMemcpy(pts, P.rhs_pts.p, P.D * 24, hipMemcpyDeviceToHost));
        if (kind) HIPCHK(hipMemcpy(kind, P.rhs_kind.p, P.D * 4, hipMemcpyDeviceToHost));
    });
}

int cpm_problem_subset(const cpm_problem* pp, int c, int which, int32_t* set_of_sub, double* cpc, double* dirs) {
    return guard([&] {
        if (!pp) throw ArgError("null problem");
        const Problem& P = *PR(pp);
        if (c < 0 || c >= (int)P.ibc.size()) throw ArgError("no such IBC");
        const IbcData& I = P.ibc[c];
        const int64_t n = which == 0 ? I.nt : I.nd;
        if (set_of_sub) HIPCHK(hipMemcpy(set_of_sub, (which == 0 ? I.t_set : I.d_set).p, n * 4, hipMemcpyDeviceToHost));
        if (cpc) HIPCHK(hipMemcpy(cpc, (which == 0 ? I.t_cpc : I.d_cpc).p, n * 24, hipMemcpyDeviceToHost));
        if (dirs) HIPCHK(hipMemcpy(dirs, (which == 0 ? I.t_dir : I.d_dir).p, n * 24, hipMemcpyDeviceToHost));
    });
}

int cpm_problem_ibc_parameter(cpm_problem* pp, int c, const double* pts, int64_t n, const double* vertex_param,
                              double* out) {
    return guard([&] {
        if (!pp) throw ArgError("null problem");
        HIPCHK(hipSetDevice(PR(pp)->ctx->device));
        problem_ibc_parameter(PR(pp), c, pts, n, vertex_param, out);
    });
}

int cpm_problem_subset_size(const cpm_problem* pp, int c, int which, int64_t* n) {
    return guard([&] {
        if (!pp || !n) throw ArgError("null argument");
        const Problem& P = *PR(pp);
        if (c < 0 || c >= (int)P.ibc.size()) throw ArgError("no such IBC");
        *n = which == 0 ? P.ibc[c].nt : P.ibc[c].nd;
    });
}

int cpm_problem_curve_rows(const cpm_problem* pp, int c, int64_t* rows) {
    return guard([&] {
        if (!pp || !rows) throw ArgError("null argument");
        const Problem& P = *PR(pp);
        if (c < 0 || c >= (int)P.ibc.size()) throw ArgError("no such IBC");
        const IbcData& I = P.ibc[c];
        rows[0] = I.start;
        rows[1] = I.nd;
        if (I.second_dirichlet()) {
            rows[2] = P.identity_rows_start + I.dir_start;
            rows[3] = I.ndir;
        } else if (I.spec.btype[0] == 0) {   // 1st-order Dirichlet: the ghost rows carry the values
            rows[2] = I.start;
            rows[3] = I.nd;
        } else {
            rows[2] = -1;
            rows[3] = 0;
        }
    });
}

int cpm_problem_export(const cpm_problem* pp, int32_t* Lp, int32_t* Lc, double* Lv, int32_t* Ep, int32_t* Ec,
                       double* Ev, uint8_t* ident, int64_t* nnzL, int64_t* nnzE) {
    return guard([&] {
        if (!pp) throw ArgError("null problem");
        const Problem& P = *PR(pp);
        const int64_t D = P.D;
        std::vector<int> ec((size_t)P.ew * D), lc((size_t)6 * D);
        std::vector<double> ev((size_t)P.ew * D), lv((size_t)6 * D), ld(D);
        std::vector<uint8_t> fl(D);
        HIPCHK(hipMemcpy(ec.data(), P.ecol.p, ec.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ev.data(), P.eval64.p, ev.size() * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lc.data(), P.lcol.p, lc.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lv.data(), P.lval64.p, lv.size() * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ld.data(), P.ldiag64.p, ld.size() * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(fl.data(), P.flag.p, fl.size(), hipMemcpyDeviceToHost));
        // identity rows carry e_r in E (Helpers::SetIdentityRows semantics)
        for (int64_t r = 0; r < D; ++r)
            if (fl[r] & ROW_IDENTITY) {
                for (int s = 0; s < P.ew; ++s) { ec[(size_t)s * D + r] = (int)r; ev[(size_t)s * D + r] = 0.0; }
                ev[r] = 1.0;
            }
        std::vector<int32_t> p1, c1, p2, c2;
        std::vector<double> v1, v2;
        ell_to_csr(D, 6, lc, lv, ld.data(), p1, c1, v1);
        ell_to_csr(D, P.ew, ec, ev, nullptr, p2, c2, v2);
        if (nnzL) *nnzL = (int64_t)c1.size();
        if (nnzE) *nnzE = (int64_t)c2.size();
        if (Lp && Lc && Lv) {
            std::copy(p1.begin(), p1.end(), Lp);
            std::copy(c1.begin(), c1.end(), Lc);
            std::copy(v1.begin(), v1.end(), Lv);
        }
        if (Ep && Ec && Ev) {
            std::copy(p2.begin(), p2.end(), Ep);
            std::copy(c2.begin(), c2.end(), Ec);
            std::copy(v2.begin(), v2.end(), Ev);
        }
        if (ident)
            for (int64_t r = 0; r < D; ++r) ident[r] = (fl[r] & ROW_IDENTITY) ? 1 : 0;
    });
}

int cpm_op_create_from_problem(cpm_problem* pp, double m, double n, int precond, int precond_iters, cpm_op** out) {
    return guard([&] {
        if (!pp || !out) throw ArgError("null argument");
        if (precond < 0 || precond > 2) throw ArgError("unknown preconditioner");
        Problem* P = PR(pp);
        HIPCHK(hipSetDevice(P->ctx->device));
        Op* op = op_from_problem(P, m, n, precond, precond_iters);
        P->ctx->op_bytes += (int64_t)op->device_bytes();
        *out = reinterpret_cast<cpm_op*>(op);
    });
}

int cpm_problem_plot_matrix(cpm_problem* prob, const double* xq, int64_t nq, int32_t* rowptr, int32_t* col,
                            double* val) {
    return guard([&] {
        if (!prob || (nq > 0 && (!xq || !rowptr || !col || !val))) throw ArgError("null argument");
        if (nq < 0) throw ArgError("nq must be >= 0");
        Problem* P = reinterpret_cast<Problem*>(prob);
        HIPCHK(hipSetDevice(P->ctx->device));
        if (nq == 0) {
            if (rowptr) rowptr[0] = 0;
            return;
        }
        problem_plot_matrix(P, xq, nq, rowptr, col, val);
    });
}

int cpm_geodesic_distance(cpm_problem* prob, double dt, int jacobi_sweeps, const cpm_solve_opts* opts, double* d_phi,
                          cpm_solve_info* infos) {
    return guard([&] {
        if (!prob || !d_phi) throw ArgError("null argument");
        if (!(dt > 0)) throw ArgError("dt must be positive");
        if (jacobi_sweeps < 1 || jacobi_sweeps > 100000) throw ArgError("jacobi_sweeps must be in 1..100000");
        Problem* P = reinterpret_cast<Problem*>(prob);
        HIPCHK(hipSetDevice(P->ctx->device));
        geodesic_distance(P, dt, jacobi_sweeps, resolve_opts(opts), d_phi, infos);
    });
}

int cpm_op_create_from_tube(cpm_tube* tp, double m, double n, int precond, int precond_iters, cpm_op** out) {
    return guard([&] {
        if (!tp || !out) throw ArgError("null argument");
        if (precond < 0 || precond > 2) throw ArgError("unknown preconditioner");
        Tube* t = T(tp);
        HIPCHK(hipSetDevice(t->ctx->device));
        Op* op = op_from_tube(t, m, n, precond, precond_iters);
        t->ctx->op_bytes += (int64_t)op->device_bytes();
        *out = reinterpret_cast<cpm_op*>(op);
    });
}

}  // extern "C"
