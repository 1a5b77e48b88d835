// Geodesic distance from interior boundary curves by the heat method on the
// tube: GeodesicDistance::ComputeDistance (problems/GeodesicDistance.cpp:11-67).
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>

#include "ibc.h"
#include "ops.h"
#include "tube.h"

namespace cpmgpu {

// SmoothH of Step1 (GeodesicDistance.cpp:73-78).
__device__ __forceinline__ double smooth_h(double scale, double t) { return 0.5 * tanh(-scale * t) + 0.5; }

__device__ __forceinline__ double norm3(const double* v) {
    double s = 0.0;
    s += v[0] * v[0];
    s += v[1] * v[1];
    s += v[2] * v[2];
    return sqrt(s);
}

// Step1's right-hand side on the Dirichlet rows of curve c (:80-100): SmoothH(-|cp_diff|).
__global__ void k_geo_rhs_dirichlet(int64_t nd, int order, int64_t start, int64_t ident0, const int32_t* d_set,
                                    const int32_t* t_sub, const double* t_cpdiff, double scale, double* rhs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const double t = -norm3(t_cpdiff + 3 * (int64_t)t_sub[d_set[i]]);
    rhs[order == 2 ? ident0 + i : start + i] = smooth_h(scale, t);
}

// ... and on the curve's DOF nodes of the tube (:102-110): SmoothH(|cp_diff|).
__global__ void k_geo_rhs_tube(int64_t nd, const int32_t* d_set, const int32_t* t_sub, const double* t_cpdiff,
                               double scale, double* rhs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const double t = norm3(t_cpdiff + 3 * (int64_t)t_sub[d_set[i]]);
    rhs[d_set[i]] = smooth_h(scale, t);
}

__global__ void k_geo_negate(int64_t n, const double* a, double* b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = -a[i];
}

#pragma clang fp contract(off)

// Row of the centred-difference matrix Dc[d] (FDMatrices.cpp:64-92): entries
// +0.5/dx at the +d neighbour and -0.5/dx at the -d neighbour when present.
__device__ __forceinline__ double dc_row(const int32_t* nb, int d, double c, const double* u) {
    double s = 0.0;
    const int p = nb[2 * d], m = nb[2 * d + 1];
    if (p >= 0) s += c * u[p];
    if (m >= 0) s += (-c) * u[m];
    return s;
}

// g[d] = Dc[d] u (Step2, GeodesicDistance.cpp:178-181), g is 3 x N.
__global__ void k_geo_grad(int64_t N, const int32_t* nbr, double c, const double* u, double* g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int32_t* nb = nbr + 6 * i;
    for (int d = 0; d < 3; ++d) g[d * N + i] = dc_row(nb, d, c, u);
}

// Tangential projection with the tube's unoriented normals and normalisation (:188-219).
__global__ void k_geo_project(int64_t N, const double* normals, double* g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double* n = normals + 3 * i;
    double v[3] = {g[i], g[N + i], g[2 * N + i]};
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s += v[d] * n[d];
    for (int d = 0; d < 3; ++d) v[d] = v[d] - s * n[d];
    double q = 0.0;
    for (int d = 0; d < 3; ++d) q += v[d] * v[d];
    q = sqrt(q);
    if (q != 0.0)
        for (int d = 0; d < 3; ++d) v[d] /= q;
    for (int d = 0; d < 3; ++d) g[d * N + i] = v[d];
}

// div = sum_d Dc[d] g[d] over the tube rows, 0 on the other rows (ComputeDivergence
// :253-262 and Step3's right-hand side :268-290).
__global__ void k_geo_div(int64_t N, int64_t D, const int32_t* nbr, double c, const double* g, double* div) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    if (i >= N) { div[i] = 0.0; return; }
    const int32_t* nb = nbr + 6 * i;
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s += dc_row(nb, d, c, g + d * N);
    div[i] = s;
}

#pragma clang fp contract(on)

// Minimum of x over n entries: block minima, then one block over them.
__global__ void k_min_blocks(int64_t n, const double* x, double* part) {
    __shared__ double sm[256];
    double m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = fmin(m, x[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void k_min_final(int nb, double* part) {
    __shared__ double sm[256];
    double m = INFINITY;
    for (int i = threadIdx.x; i < nb; i += 256) m = fmin(m, part[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[0] = sm[0];
}

// x -= min(x) (Step3 :301-313).
__global__ void k_sub_min(int64_t n, const double* mn, double* x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] -= mn[0];
}

static unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void geodesic_distance(Problem* P, double dt, int sweeps, const cpm_solve_opts& o, double* d_phi,
                       cpm_solve_info* infos) {
    Ctx* ctx = P->ctx;
    Tube* t = P->tube;
    const hipStream_t st = ctx->stream;
    const int64_t N = t->N, D = P->D;
    if (P->ibc.empty()) throw ArgError("the geodesic distance needs at least one curve");
    // operators: Step1 Solver(1, -dt) with the Jacobi preconditioner (:112-115),
    // Step3 Solver(0, 1) with the diagonal one (:292-295), and the plain E (:183-186)
    std::unique_ptr<Op> op1(op_from_problem(P, 1.0, -dt, CPM_PRECOND_JACOBI, sweeps));
    std::unique_ptr<Op> op3(op_from_problem(P, 0.0, 1.0, CPM_PRECOND_DIAGONAL, 0));
    std::unique_ptr<Op> opE(op_from_tube(t, 1.0, 0.0, CPM_PRECOND_DIAGONAL, 0));
    double* w = ctx->workspace("geo_vecs", sizeof(double) * (3 * D + 8 * N + 1024));
    double *rhs = w, *uf = w + D, *div = w + 2 * D, *u = w + 3 * D, *u2 = u + N, *g = u2 + N, *eg = g + 3 * N;
    double* part = eg + 3 * N;
    const double scale = std::atanh(1.0 - t->dx) / t->r_t;   // :70-72
    const double c = 0.5 / t->dx;
    HIPCHK(hipMemsetAsync(rhs, 0, sizeof(double) * D, st));
    for (auto& I : P->ibc)
        if (I.spec.btype[0] == 0)
            hipLaunchKernelGGL(k_geo_rhs_dirichlet, dim3(nblk(I.nd)), dim3(256), 0, st, I.nd, I.spec.order, I.start,
                               P->identity_rows_start + I.dir_start, I.d_set.as<int32_t>(), I.t_sub.as<int32_t>(),
                               I.t_cpdiff.as<double>(), scale, rhs);
    for (auto& I : P->ibc)
        hipLaunchKernelGGL(k_geo_rhs_tube, dim3(nblk(I.nd)), dim3(256), 0, st, I.nd, I.d_set.as<int32_t>(),
                           I.t_sub.as<int32_t>(), I.t_cpdiff.as<double>(), scale, rhs);
    HIPCHK(hipGetLastError());
    std::vector<cpm_solve_info> inf;
    solve_batch(ctx, {op1.get()}, {rhs}, {uf}, o, inf);
    if (infos) infos[0] = inf[0];
    hipLaunchKernelGGL(k_geo_negate, dim3(nblk(N)), dim3(256), 0, st, N, uf, u);   // u = -Step1 (:38)
    const unsigned nbm = 256;
    for (int pass = 0; pass < 3; ++pass) {
        if (pass > 0) {   // u = E phi on the tube rows (:49-56)
            HIPCHK(hipMemcpyAsync(u2, d_phi, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            apply_E(opE.get(), u2, u);
        }
        hipLaunchKernelGGL(k_geo_grad, dim3(nblk(N)), dim3(256), 0, st, N, t->nbr.as<int32_t>(), c, u, g);
        for (int d = 0; d < 3; ++d) apply_E(opE.get(), g + d * N, eg + d * N);
        hipLaunchKernelGGL(k_geo_project, dim3(nblk(N)), dim3(256), 0, st, N, P->normals.as<double>(), eg);
        hipLaunchKernelGGL(k_geo_div, dim3(nblk(D)), dim3(256), 0, st, N, D, t->nbr.as<int32_t>(), c, eg, div);
        HIPCHK(hipGetLastError());
        solve_batch(ctx, {op3.get()}, {div}, {d_phi}, o, inf);
        if (infos) infos[1 + pass] = inf[0];
        hipLaunchKernelGGL(k_min_blocks, dim3(nbm), dim3(256), 0, st, D, d_phi, part);
        hipLaunchKernelGGL(k_min_final, dim3(1), dim3(256), 0, st, (int)nbm, part);
        hipLaunchKernelGGL(k_sub_min, dim3(nblk(D)), dim3(256), 0, st, D, part, d_phi);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
}

}  // namespace cpmgpu
