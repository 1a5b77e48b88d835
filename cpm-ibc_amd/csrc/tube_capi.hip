// C ABI of the GPU tube (cpm_tube_*) and of operators built from it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cpm_gpu.h"
#include "ops.h"
#include "tube.h"

using namespace cpmgpu;

namespace cpmgpu {
void set_last_error(const std::string& s);
int status_of_current_exception();
}

namespace {
inline Ctx* C(cpm_ctx* c) { return reinterpret_cast<Ctx*>(c); }
inline Tube* T(cpm_tube* t) { return reinterpret_cast<Tube*>(t); }
inline const Tube* T(const cpm_tube* t) { return reinterpret_cast<const Tube*>(t); }

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
        if (x)   // Grid::CoordinateFromTriplet (Grid.cpp:292-303)
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
