// GPU assembly of the IBC-modified closest-point operator.
//
// Reference: src/IBCSubset.cpp (tagging subset, cp_diff, frame-based
// directions), src/Geometry.cpp (Jacobian of cp and its eigenvectors),
// src/IBCDOFSubset.cpp (DOF subset, cp_c_bar, extension rows),
// src/IBCSubsets.cpp (column offsets), src/IBCMatrixManipulations.cpp
// (stencil tagging, ghost columns and rows) and problems/PoissonWithIBC.cpp.
//
// Layout of the assembled system (rows = columns = D = N + G + K):
//   [0, N)        tube rows: E / L rows whose stencil entries that cross the
//                 curve are redirected to ghost columns,
//   [N, N+G)      ghost rows: extension rows at cp_c_bar (2nd order) scaled by
//                 -1 (Dirichlet) or +1 (Neumann), plus 2.0 on the Dirichlet
//                 identity column; Laplacian rows copied from the DOF nodes,
//   [N+G, D)      2nd-order Dirichlet identity rows.
// Everything is computed on the device; the host only sizes buffers.
#include <hip/hip_runtime.h>

#include <cfloat>

#include <chrono>
#include <cmath>

#include "ibc.h"

#pragma clang fp contract(off)

namespace cpmgpu {

namespace {

constexpr int kMaxIbc = 32;

__device__ __forceinline__ void node_coords(int64_t l, int64_t N1d, double xs, double dx, double* x) {
    const int64_t k = l % N1d, j = (l / N1d) % N1d, i = l / (N1d * N1d);
    x[0] = xs + i * dx;
    x[1] = xs + j * dx;
    x[2] = xs + k * dx;
}

__device__ __forceinline__ void cross3(const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Unit eigenvector of a symmetric 3x3 matrix for its smallest (which = 0) or
// largest (which = 2) eigenvalue.  Eigenvalues from the trigonometric
// solution of the characteristic cubic; the vector is the largest cross
// product of two rows of (A - lambda I).  The sign is arbitrary: every use
// (unoriented interpolation, b = t x n, (b . d) b) is sign invariant.
// Also returns the eigenvalue.
__device__ void sym3_extreme_eigvec(const double A[3][3], int which, double* v, double* lambda) {
    const double p1 = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    double lam;
    if (p1 == 0.0) {
        int k = 0;
        for (int i = 1; i < 3; ++i)
            if (which == 0 ? A[i][i] < A[k][k] : A[i][i] > A[k][k]) k = i;
        v[0] = v[1] = v[2] = 0.0;
        v[k] = 1.0;
        *lambda = A[k][k];
        return;
    }
    const double q = (A[0][0] + A[1][1] + A[2][2]) / 3.0;
    const double p2 = (A[0][0] - q) * (A[0][0] - q) + (A[1][1] - q) * (A[1][1] - q) + (A[2][2] - q) * (A[2][2] - q) + 2.0 * p1;
    const double p = sqrt(p2 / 6.0);
    double B[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i][j] = (A[i][j] - (i == j ? q : 0.0)) / p;
    const double detB = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                        B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
    const double r = fmin(1.0, fmax(-1.0, detB / 2.0));
    const double phi = acos(r) / 3.0;
    if (which == 2) lam = q + 2.0 * p * cos(phi);
    else lam = q + 2.0 * p * cos(phi + 2.0 * M_PI / 3.0);
    double M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = A[i][j] - (i == j ? lam : 0.0);
    double c[3][3];
    cross3(M[0], M[1], c[0]);
    cross3(M[0], M[2], c[1]);
    cross3(M[1], M[2], c[2]);
    int best = 0;
    double bn = dot3(c[0], c[0]);
    for (int k = 1; k < 3; ++k) {
        const double nk = dot3(c[k], c[k]);
        if (nk > bn) { bn = nk; best = k; }
    }
    if (bn == 0.0) {   // degenerate (repeated extreme eigenvalue): any vector of the eigenspace
        v[0] = 1.0; v[1] = 0.0; v[2] = 0.0;
    } else {
        const double nn = sqrt(bn);
        for (int k = 0; k < 3; ++k) v[k] = c[best][k] / nn;
    }
    *lambda = lam;
}

// Central-difference gradient row of FDMatrices::BuildGradientMatrices
// (FDMatrices.cpp:66-92) applied to one coordinate array: entries +0.5/dx at
// the +c neighbour and -0.5/dx at the -c neighbour, summed in ascending
// column order like the CSR product.
__device__ __forceinline__ double grad_entry(int32_t cp_, int32_t cm, double vp, double vm, double dx) {
    const double tp = 0.5 / dx, tm = -0.5 / dx;
    if (cp_ >= 0 && cm >= 0) {
        double s;
        if (cp_ < cm) { s = tp * vp; s += tm * vm; }
        else { s = tm * vm; s += tp * vp; }
        return s;
    }
    if (cp_ >= 0) return tp * vp;
    if (cm >= 0) return tm * vm;
    return 0.0;
}

// ---------------------------------------------------------------- whole-tube normals
// Gradients of the closest point coordinates: G[(r*3+c)][i] = (Dc_c cp_r)_i.
__global__ void k_fill_f64(double* a, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_grad_tube(int64_t N, const int32_t* nbr, const double* cpx, double dx, double* G) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int c = 0; c < 3; ++c) {
        const int32_t np = nbr[6 * i + 2 * c], nm = nbr[6 * i + 2 * c + 1];
        for (int r = 0; r < 3; ++r) {
            const double vp = np >= 0 ? cpx[3 * (int64_t)np + r] : 0.0, vm = nm >= 0 ? cpx[3 * (int64_t)nm + r] : 0.0;
            G[(int64_t)(r * 3 + c) * N + i] = grad_entry(np, nm, vp, vm, dx);
        }
    }
}

// J = E G (extension of the Jacobian), symmetrised; eigenvector of the
// smallest eigenvalue = unoriented surface normal (IBCSubset.cpp:192-226).
__global__ void k_tube_normals(int64_t N, const int* ecol, const double* eval, int ew, int64_t ld, const double* G,
                               double* normals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < ew; ++s) {
        const int col = ecol[s * ld + i];
        const double w = eval[s * ld + i];
        for (int k = 0; k < 9; ++k) J[k] += w * G[(int64_t)k * N + col];
    }
    double A[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A[r][c] = 0.5 * (J[r * 3 + c] + J[c * 3 + r]);
    double v[3], lam;
    sym3_extreme_eigvec(A, 0, v, &lam);
    normals[3 * i + 0] = v[0];
    normals[3 * i + 1] = v[1];
    normals[3 * i + 2] = v[2];
}

// ---------------------------------------------------------------- tagging BFS
// Breadth-first traversal over tube nodes (IBCSubset.cpp:39-109), level
// synchronous with the FIFO position reproduced by (parent rank, neighbour
// slot) claims.  The start node is not pre-marked, as in the reference.
__global__ void k_tbfs_expand(const int32_t* frontier, int64_t nf, const int32_t* nbr, int32_t* vis, int32_t* cand,
                              int* ncand) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * 6) return;
    const int64_t fi = t / 6;
    const int slot = (int)(t - fi * 6);
    const int32_t nb = nbr[6 * (int64_t)frontier[fi] + slot];
    if (nb < 0) return;
    const int32_t v = __hip_atomic_load(vis + nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v < kCand || v == kOutside) return;
    const int32_t old = atomicMin(vis + nb, kCand + (int32_t)t);
    if (old == kUnvisited) cand[atomicAdd(ncand, 1)] = nb;
}

__global__ void k_tbfs_eval(const int32_t* cand, const int* ncand, const int64_t* rec, int64_t N1d, double xs, double dx,
                            SurfDev sd, double radius, int32_t* vis, int32_t* slot, double* ccp, double* cdist,
                            int32_t* cbdy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *ncand) return;
    const int32_t nb = cand[i];
    double x[3], c[3], d;
    int b;
    node_coords(rec[nb], N1d, xs, dx, x);
    cp_surface(sd, x, c, d, b);
    if (d < radius) {
        slot[vis[nb] - kCand] = (int32_t)i;
        for (int k = 0; k < 3; ++k) ccp[3 * i + k] = c[k];
        cdist[i] = d;
        cbdy[i] = b;
    } else {
        vis[nb] = kOutside;
    }
}

__global__ void k_tbfs_emit(const int32_t* slot, int64_t ns, const int* boff, const int32_t* cand, const double* ccp,
                            const double* cdist, const int32_t* cbdy, int64_t base, int32_t* vis, int32_t* t_set,
                            double* t_cpc, double* t_dist, int32_t* t_bdy, int32_t* next) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int32_t ci = i < ns ? slot[i] : -1;
    const int occ = ci >= 0;
    const unsigned long long m = __ballot(occ);
    __shared__ int ws[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = __popcll(m);
    __syncthreads();
    if (!occ) return;
    int before = boff[blockIdx.x];
    for (int k = 0; k < w; ++k) before += ws[k];
    before += __popcll(m & ((1ull << lane) - 1ull));
    const int64_t t = base + before;
    const int32_t nb = cand[ci];
    vis[nb] = (int32_t)t;
    t_set[t] = nb;
    for (int k = 0; k < 3; ++k) t_cpc[3 * t + k] = ccp[3 * ci + k];
    t_dist[t] = cdist[ci];
    t_bdy[t] = cbdy[ci];
    next[before] = nb;   // the frontier holds tube indices
}

// ---------------------------------------------------------------- ordered selection
// flags -> compacted indices in order (generic helper: count, scan, emit)
__global__ void k_sel_count(const uint8_t* flag, int64_t n, int* bcount) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const unsigned long long m = __ballot(i < n && flag[i]);
    __shared__ int ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}
__global__ void k_sel_emit(const uint8_t* flag, int64_t n, const int* boff, int32_t* out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int f = i < n && flag[i];
    const unsigned long long m = __ballot(f);
    __shared__ int ws[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = __popcll(m);
    __syncthreads();
    if (!f) return;
    int before = boff[blockIdx.x];
    for (int k = 0; k < w; ++k) before += ws[k];
    before += __popcll(m & ((1ull << lane) - 1ull));
    out[before] = (int32_t)i;
}

// ---------------------------------------------------------------- cp_diff
// cp_diff = cp_S(x) - cp_C(x) (IBCSubset.cpp:156-187), flagging |cp_diff| < tol.
__global__ void k_cpdiff(int64_t nt, const int32_t* t_set, const double* cpx, const double* t_cpc, double tol,
                         double* cpdiff, uint8_t* onibc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const int64_t s = t_set[i];
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = cpx[3 * s + k] - t_cpc[3 * i + k];
    for (int k = 0; k < 3; ++k) cpdiff[3 * i + k] = d[k];
    onibc[i] = sqrt(dot3(d, d)) < tol;
}

// Nodes on the curve take the averaged, sign-aligned cp_diff of their subset
// neighbours (Geometry.cpp:144-188), in subset order: a later node sees the
// already-updated value of an earlier one, so one thread walks the list.
__global__ void k_onibc_fix(const int32_t* list, int64_t nl, const int32_t* t_set, const int32_t* t_sub,
                            const int32_t* nbr, double tol, double* cpdiff) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t q = 0; q < nl; ++q) {
        const int32_t i = list[q];
        const int64_t s = t_set[i];
        double acc[3] = {0, 0, 0}, first[3] = {0, 0, 0};
        int num = 0;
        for (int nb = 0; nb < 6; ++nb) {
            const int32_t ni = nbr[6 * s + nb];
            if (ni < 0) continue;
            const int32_t sn = t_sub[ni];
            if (sn < 0) continue;
            const double* w = cpdiff + 3 * (int64_t)sn;
            if (sqrt(dot3(w, w)) >= tol) {
                if (num == 0) { first[0] = w[0]; first[1] = w[1]; first[2] = w[2]; }
                const double sg = dot3(first, w) > 0 ? 1.0 : -1.0;
                for (int k = 0; k < 3; ++k) acc[k] += sg * w[k];
                ++num;
            }
        }
        if (num > 0)
            for (int k = 0; k < 3; ++k) cpdiff[3 * (int64_t)i + k] = acc[k] / (double)num;
    }
}

// ---------------------------------------------------------------- frame directions
// Per tagging node: unoriented normal interpolated at cp_c
// (Geometry.cpp:205-266, normalised), tangent of the curve from the
// subset-restricted Jacobian of cp_C (Geometry.cpp:27-69 with the subset
// maps), binormal b = t x n and direction (b . cp_diff) b; a point curve
// (largest eigenvalue ~ 0) uses cp_diff minus its normal component
// (IBCSubset.cpp:242-265).  ecol/eval: E rows at cp_c (tube columns).
__global__ void k_frame(int64_t nt, const int* ecol, const double* eval, int ew, int64_t ld, const double* tube_normals,
                        const int32_t* t_set, const int32_t* t_sub, const int32_t* nbr, const double* t_cpc,
                        const double* cpdiff, double dx, double* gsub, int pass, double* dir, double* bn) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    if (pass == 0) {   // subset-restricted gradients of cp_C
        const int64_t s = t_set[i];
        for (int c = 0; c < 3; ++c) {
            int32_t np = nbr[6 * s + 2 * c], nm = nbr[6 * s + 2 * c + 1];
            np = np >= 0 ? t_sub[np] : -1;
            nm = nm >= 0 ? t_sub[nm] : -1;
            for (int r = 0; r < 3; ++r) {
                const double vp = np >= 0 ? t_cpc[3 * (int64_t)np + r] : 0.0, vm = nm >= 0 ? t_cpc[3 * (int64_t)nm + r] : 0.0;
                gsub[(int64_t)(r * 3 + c) * nt + i] = grad_entry(np, nm, vp, vm, dx);
            }
        }
        return;
    }
    // normal: sign-aligned interpolation; the reference's first entry is the
    // smallest column of the row
    int first = -1;
    for (int s = 0; s < ew; ++s) {
        const int col = ecol[s * ld + i];
        if (first < 0 || col < ecol[first * ld + i]) first = s;
    }
    double f0[3], n[3] = {0, 0, 0};
    const int fc = ecol[first * ld + i];
    for (int k = 0; k < 3; ++k) f0[k] = tube_normals[3 * (int64_t)fc + k];
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < ew; ++s) {
        const int col = ecol[s * ld + i];
        const double w = eval[s * ld + i];
        const double* v = tube_normals + 3 * (int64_t)col;
        const double sg = dot3(f0, v) > 0.0 ? 1.0 : -1.0;
        for (int k = 0; k < 3; ++k) n[k] += sg * (w * v[k]);
        const int32_t sc = t_sub[col];
        if (sc >= 0)
            for (int k = 0; k < 9; ++k) J[k] += w * gsub[(int64_t)k * nt + sc];
    }
    const double nn = sqrt(dot3(n, n));
    if (nn > 0.0) for (int k = 0; k < 3; ++k) n[k] /= nn;
    else n[0] = n[1] = n[2] = 0.0;
    double A[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A[r][c] = 0.5 * (J[r * 3 + c] + J[c * 3 + r]);
    double t[3], lam;
    sym3_extreme_eigvec(A, 2, t, &lam);
    const double* d = cpdiff + 3 * i;
    if (fabs(lam) < 1e-8) {
        const double sdn = dot3(n, d);
        for (int k = 0; k < 3; ++k) {
            dir[3 * i + k] = d[k] - sdn * n[k];
            bn[3 * i + k] = 0.0;
        }
    } else {
        double b[3];
        cross3(t, n, b);
        const double sb = dot3(b, d);
        for (int k = 0; k < 3; ++k) {
            dir[3 * i + k] = sb * b[k];
            bn[3 * i + k] = b[k];
        }
    }
}

// DOF subset members: tagging nodes with dist_c < scale r_t (IBCDOFSubset.cpp:18-29).
__global__ void k_dof_flag(int64_t nt, const double* t_dist, double lim, uint8_t* f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nt) f[i] = t_dist[i] < lim;
}

__global__ void k_dof_fill(int64_t nd, const int32_t* sel, const int32_t* t_set, const double* t_cpc, const double* t_dir,
                           int32_t* d_set, double* d_cpc, double* d_dir, int32_t* d_sub) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const int32_t ti = sel[i];
    d_set[i] = t_set[ti];
    d_sub[t_set[ti]] = (int32_t)i;
    for (int k = 0; k < 3; ++k) {
        d_cpc[3 * i + k] = t_cpc[3 * (int64_t)ti + k];
        d_dir[3 * i + k] = t_dir[3 * (int64_t)ti + k];
    }
}

// Extension query points of the ghost rows: cp_c (first order) or
// cp_c_bar = cp_S(2 cp_c - x) (second order, IBCDOFSubset.cpp:81-90).
__global__ void k_ghost_query(int64_t nd, int order, const int32_t* d_set, const double* d_cpc, const int64_t* rec,
                              int64_t N1d, double xs, double dx, SurfDev surf, double* q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    if (order != 2) {
        for (int k = 0; k < 3; ++k) q[3 * i + k] = d_cpc[3 * i + k];
        return;
    }
    double x[3], pr[3], c[3], d;
    int b;
    node_coords(rec[d_set[i]], N1d, xs, dx, x);
    for (int k = 0; k < 3; ++k) pr[k] = 2.0 * d_cpc[3 * i + k] - x[k];
    cp_surface(surf, pr, c, d, b);
    for (int k = 0; k < 3; ++k) q[3 * i + k] = c[k];
}

// ---------------------------------------------------------------- assembly
// TagStencils (IBCMatrixManipulations.cpp:423-463) + MoveToIBCColumns
// (:189-216) for one entry (row r of the director set, column col): the
// column moves to the ghost column of IBC c when the directions of the row's
// director and of the stencil point disagree (dot < 0) and the stencil point
// belongs to the DOF subset.  Returns the new column (col if not moved).
__device__ __forceinline__ int64_t moved_col(const IbcDev* ib, int nibc, const double* const* rowdir, int64_t col,
                                             double flip, int* nmoved) {
    int64_t out = col;
    *nmoved = 0;
    for (int c = 0; c < nibc; ++c) {
        if (!rowdir[c]) continue;
        const int32_t sp = ib[c].t_sub[col];
        if (sp < 0 || ib[c].t_bdy[sp] != 0) continue;
        const double dp = flip * dot3(rowdir[c], ib[c].t_dir + 3 * (int64_t)sp);
        if (!(dp < 0.0)) continue;
        const int32_t ds = ib[c].d_sub[col];
        if (ds < 0) continue;
        if (*nmoved == 0) out = ib[c].start + ds;
        ++*nmoved;
    }
    return out;
}

struct AsmArgs {
    IbcDev ib[kMaxIbc];
    int nibc;
    int64_t N, D, ld;        // ld = D (slot stride of the output ELL)
    int ew;
    double h;
};

// Tube rows of Efull and Lfull.
__global__ void k_asm_tube(AsmArgs a, const int* ecol_t, const double* eval_t, int ew_t, int64_t ld_t,
                           const int32_t* nbr, int* ecol, double* eval, float* eval32, int* lcol, double* lval,
                           float* lval32, double* ldiag, float* ldiag32, uint8_t* flag, int* err) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.N) return;
    const double* rowdir[kMaxIbc];
    for (int c = 0; c < a.nibc; ++c) {
        const int32_t ts = a.ib[c].t_sub[r];
        rowdir[c] = ts >= 0 ? a.ib[c].t_dir + 3 * (int64_t)ts : nullptr;
    }
    int s = 0;
    for (; s < ew_t; ++s) {
        const int64_t col = ecol_t[s * ld_t + r];
        int nm;
        const int64_t nc = moved_col(a.ib, a.nibc, rowdir, col, 1.0, &nm);
        if (nm > 1) atomicOr(err, 1);
        const double v = eval_t[s * ld_t + r];
        ecol[s * a.ld + r] = (int)nc;
        eval[s * a.ld + r] = v;
        eval32[s * a.ld + r] = (float)v;
    }
    for (; s < a.ew; ++s) {
        ecol[s * a.ld + r] = (int)r;
        eval[s * a.ld + r] = 0.0;
        eval32[s * a.ld + r] = 0.0f;
    }
    const double hh = a.h * a.h, dg = -6.0 / hh, off = 1.0 / hh;
    for (int k = 0; k < 6; ++k) {
        const int32_t nb = nbr[6 * r + k];
        int64_t nc = r;
        double v = 0.0;
        if (nb >= 0) {
            int nm;
            nc = moved_col(a.ib, a.nibc, rowdir, nb, 1.0, &nm);
            if (nm > 1) atomicOr(err, 1);
            v = off;
        }
        lcol[k * a.ld + r] = (int)nc;
        lval[k * a.ld + r] = v;
        lval32[k * a.ld + r] = (float)v;
    }
    ldiag[r] = dg;
    ldiag32[r] = (float)dg;
    flag[r] = 0;
}

// Ghost rows of IBC c (AddExtensionIBCRows :118-186, AddFiniteDifferenceIBCRows
// :219-251) and its 2nd-order identity rows.
__global__ void k_asm_ghost(AsmArgs a, int c, int64_t nd, int order, int btype, int64_t ident0, const int* ecol_g,
                            const double* eval_g, int ew_g, const int32_t* d_set, const int32_t* nbr, int* ecol,
                            double* eval, float* eval32, int* lcol, double* lval, float* lval32, double* ldiag,
                            float* ldiag32, uint8_t* flag, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const IbcDev& ib = a.ib[c];
    const int64_t r = ib.start + i;
    const bool dirichlet = btype == 0;
    const double scale = dirichlet ? -1.0 : 1.0;
    const bool rows = btype == 1 || (dirichlet && order == 2);
    // E: extension row at the ghost query point, tagged with flipped directions
    const double* rowdir[kMaxIbc];
    for (int k = 0; k < a.nibc; ++k) rowdir[k] = nullptr;
    rowdir[c] = ib.d_dir + 3 * i;
    int s = 0;
    if (rows)
        for (; s < ew_g; ++s) {
            const int64_t col = ecol_g[s * nd + i];
            const int32_t sp = ib.t_sub[col];
            int64_t nc = col;
            if (sp >= 0 && ib.t_bdy[sp] == 0 && -dot3(rowdir[c], ib.t_dir + 3 * (int64_t)sp) < 0.0 && ib.d_sub[col] >= 0)
                nc = ib.start + ib.d_sub[col];
            const double v = scale * eval_g[s * nd + i];
            ecol[s * a.ld + r] = (int)nc;
            eval[s * a.ld + r] = v;
            eval32[s * a.ld + r] = (float)v;
        }
    if (rows && dirichlet && order == 2) {   // 2.0 on the Dirichlet identity column
        const int64_t idc = ident0 + i;
        ecol[s * a.ld + r] = (int)idc;
        eval[s * a.ld + r] = 2.0;
        eval32[s * a.ld + r] = 2.0f;
        ++s;
    }
    for (; s < a.ew; ++s) {
        ecol[s * a.ld + r] = (int)r;
        eval[s * a.ld + r] = 0.0;
        eval32[s * a.ld + r] = 0.0f;
    }
    // L: the DOF node's Laplacian row; crossing entries keep their column,
    // others move to ghost columns when the column is a DOF node.
    const int64_t sn = d_set[i];
    const double hh = a.h * a.h, dg = -6.0 / hh, off = 1.0 / hh;
    int k = 0;
    for (int q = 0; q < 6; ++q) {
        const int32_t nb = nbr[6 * sn + q];
        if (nb < 0) continue;
        const int32_t ts_r = ib.t_sub[sn], sp = ib.t_sub[nb];
        const bool crossing = ts_r >= 0 && sp >= 0 && ib.t_bdy[sp] == 0 &&
                              dot3(ib.t_dir + 3 * (int64_t)ts_r, ib.t_dir + 3 * (int64_t)sp) < 0.0;
        int64_t nc = -1;
        if (crossing) nc = nb;
        else if (ib.d_sub[nb] >= 0) nc = ib.start + ib.d_sub[nb];
        if (nc < 0) continue;
        lcol[k * a.ld + r] = (int)nc;
        lval[k * a.ld + r] = off;
        lval32[k * a.ld + r] = (float)off;
        ++k;
    }
    for (; k < 6; ++k) {
        lcol[k * a.ld + r] = (int)r;
        lval[k * a.ld + r] = 0.0;
        lval32[k * a.ld + r] = 0.0f;
    }
    ldiag[r] = dg;
    ldiag32[r] = (float)dg;
    flag[r] = (dirichlet && order != 2) ? ROW_IDENTITY : 0;   // first-order Dirichlet ghost rows are identity rows
    if (dirichlet && order == 2) {
        const int64_t ir = ident0 + i;
        for (int q = 0; q < a.ew; ++q) {
            ecol[q * a.ld + ir] = (int)ir;
            eval[q * a.ld + ir] = 0.0;
            eval32[q * a.ld + ir] = 0.0f;
        }
        for (int q = 0; q < 6; ++q) {
            lcol[q * a.ld + ir] = (int)ir;
            lval[q * a.ld + ir] = 0.0;
            lval32[q * a.ld + ir] = 0.0f;
        }
        ldiag[ir] = 1.0;
        ldiag32[ir] = 1.0f;
        flag[ir] = ROW_IDENTITY;
    }
}

// Evaluation points of the right-hand side, PoissonWithIBC::GetRHS
// (problems/PoissonWithIBC.cpp:50-90): kind 0 tube row (f at cp), 1 ghost row
// (f at cp_c), 2 Dirichlet identity row (boundary value at cp_c).
__global__ void k_rhs_points(int64_t nd, int64_t start, int64_t ident0, int order, int btype, const double* d_cpc,
                             double* pts, int32_t* kind) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const bool dirichlet = btype == 0;
    if (btype == 1 || (dirichlet && order == 2)) {
        for (int k = 0; k < 3; ++k) pts[3 * (start + i) + k] = d_cpc[3 * i + k];
        kind[start + i] = 1;
    }
    if (dirichlet) {
        const int64_t r = order == 2 ? ident0 + i : start + i;
        for (int k = 0; k < 3; ++k) pts[3 * r + k] = d_cpc[3 * i + k];
        kind[r] = 2;
    }
}

__global__ void k_rhs_tube(int64_t N, const double* cpx, double* pts, int32_t* kind) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int k = 0; k < 3; ++k) pts[3 * i + k] = cpx[3 * i + k];
    kind[i] = 0;
}

template <typename T>
void realloc(Buffer& b, size_t n) {
    b.~Buffer();
    new (&b) Buffer(n * sizeof(T));
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// Ordered selection of flagged indices; returns the count.
static int64_t select_flags(Ctx* ctx, const uint8_t* flag, int64_t n, int32_t* out) {
    const hipStream_t st = ctx->stream;
    const int64_t nb = (n + 255) / 256;
    if (nb == 0) return 0;
    int* bcount = reinterpret_cast<int*>(ctx->workspace("sel_bcount", nb * 4));
    int* total = reinterpret_cast<int*>(ctx->workspace("sel_total", 64));
    hipLaunchKernelGGL(k_sel_count, dim3((unsigned)nb), dim3(256), 0, st, flag, n, bcount);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, bcount, nb, total);
    hipLaunchKernelGGL(k_sel_emit, dim3((unsigned)nb), dim3(256), 0, st, flag, n, bcount, out);
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, total, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
}

// Tagging subset of one curve (IBCSubset::ComputeSubset, IBCSubset.cpp:12-36).
static void tagging_subset(Tube* t, Ctx* ctx, IbcData& I, const double* tube_normals) {
    const hipStream_t st = ctx->stream;
    const int64_t N = t->N;
    const double radius = 2.0 * t->r_t;
    // FindStartingPoint (IBCSubset.cpp:112-116): nearest tube node of the curve's first parameterisation point
    const double* p0 = I.spec.surf.type == "Polyline" ? I.spec.surf.verts.data() : nullptr;
    double x0[3];
    if (p0) { x0[0] = p0[0]; x0[1] = p0[1]; x0[2] = p0[2]; }
    else if (I.spec.surf.type == "Point") { for (int k = 0; k < 3; ++k) x0[k] = I.spec.surf.params[k]; }
    else if (I.spec.surf.type == "Arc") {   // paramArc, first point at angle1 (Surface.cpp:1907-1924)
        const auto& P = I.spec.surf.params;
        x0[0] = P[0] * std::cos(P[1]);
        x0[1] = P[0] * std::sin(P[1]);
        x0[2] = 0.0;
    }
    else if (I.spec.surf.type == "Circle3D") {   // paramCircle3D, theta = 0 (Surface.cpp:1883-1904)
        const auto& P = I.spec.surf.params;
        double nr[3] = {P[4], P[5], P[6]}, t1[3], t2[3];
        if (std::fabs(nr[0]) > std::fabs(nr[1])) { t1[0] = -nr[2]; t1[1] = 0; t1[2] = nr[0]; }
        else { t1[0] = 0; t1[1] = nr[2]; t1[2] = -nr[1]; }
        t2[0] = nr[1] * t1[2] - nr[2] * t1[1];
        t2[1] = nr[2] * t1[0] - nr[0] * t1[2];
        t2[2] = nr[0] * t1[1] - nr[1] * t1[0];
        const double n1 = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
        const double n2 = std::sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
        for (int k = 0; k < 3; ++k) x0[k] = P[1 + k] + P[0] * (std::cos(0.0) * (t1[k] / n1) + std::sin(0.0) * (t2[k] / n2));
    } else throw ArgError("IBC curve type '" + I.spec.surf.type + "' is not supported on the GPU");
    int64_t tr[3];
    for (int k = 0; k < 3; ++k) tr[k] = (int64_t)std::round((x0[k] - t->xstart) / t->dx);
    for (int k = 0; k < 3; ++k)
        if (tr[k] < 0 || tr[k] >= t->N1d) throw GeomError("IBC start point outside the grid");
    const int64_t l0 = (tr[0] * t->N1d + tr[1]) * t->N1d + tr[2];
    int32_t start = -1;
    HIPCHK(hipMemcpy(&start, t->gidx.as<int32_t>() + l0, 4, hipMemcpyDeviceToHost));
    if (start < 0 || start >= kCand) throw GeomError("IBC start point is not in the tube");

    int32_t* vis = reinterpret_cast<int32_t*>(ctx->workspace("ibc_vis", N * 4));
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min<unsigned>(nblk(N), 4096)), dim3(256), 0, st, vis, N, kUnvisited);
    int64_t cap = 1 << 14;
    realloc<int32_t>(I.t_set, cap);
    realloc<double>(I.t_cpc, cap * 3);
    realloc<double>(I.t_dist, cap);
    realloc<int32_t>(I.t_bdy, cap);
    std::unique_ptr<Buffer> fr(new Buffer(cap * 4)), fr2(new Buffer(cap * 4));
    HIPCHK(hipMemcpyAsync(fr->p, &start, 4, hipMemcpyHostToDevice, st));
    int* dint = reinterpret_cast<int*>(ctx->workspace("ibc_int", 64));
    int* hbuf = reinterpret_cast<int*>(ctx->host_states);
    int64_t nt = 0, nf = 1;
    while (nf > 0) {
        HIPCHK(hipStreamSynchronize(st));
        const int64_t ns = nf * 6, nb = (ns + 255) / 256;
        int32_t* cand = reinterpret_cast<int32_t*>(ctx->workspace("ibc_cand", ns * 4));
        int32_t* slot = reinterpret_cast<int32_t*>(ctx->workspace("ibc_slot", ns * 4));
        double* ccp = ctx->workspace("ibc_ccp", ns * 24);
        double* cdist = ctx->workspace("ibc_cdist", ns * 8);
        int32_t* cbdy = reinterpret_cast<int32_t*>(ctx->workspace("ibc_cbdy", ns * 4));
        int* bcount = reinterpret_cast<int*>(ctx->workspace("ibc_bcount", nb * 4));
        HIPCHK(hipMemsetAsync(dint, 0, 16, st));
        hipLaunchKernelGGL(k_fill_i32, dim3(std::min<unsigned>((unsigned)nb, 4096)), dim3(256), 0, st, slot, ns, -1);
        hipLaunchKernelGGL(k_tbfs_expand, dim3((unsigned)nb), dim3(256), 0, st, fr->as<int32_t>(), nf,
                           t->nbr.as<int32_t>(), vis, cand, dint);
        hipLaunchKernelGGL(k_tbfs_eval, dim3((unsigned)nb), dim3(256), 0, st, cand, dint, t->rec.as<int64_t>(), t->N1d,
                           t->xstart, t->dx, I.sd, radius, vis, slot, ccp, cdist, cbdy);
        hipLaunchKernelGGL(k_count, dim3((unsigned)nb), dim3(256), 0, st, slot, ns, bcount);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, bcount, nb, dint + 1);
        HIPCHK(hipMemcpyAsync(hbuf, dint, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const int64_t cnt = hbuf[1];
        if (nt + cnt > cap || cnt > cap) {
            int64_t nc = cap;
            while (nc < nt + cnt || nc < cnt) nc *= 2;
            auto grow = [&](Buffer& b, size_t esz) {
                Buffer nbuf(nc * esz);
                if (nt) HIPCHK(hipMemcpy(nbuf.p, b.p, nt * esz, hipMemcpyDeviceToDevice));
                std::swap(b.p, nbuf.p);
                std::swap(b.bytes, nbuf.bytes);
            };
            grow(I.t_set, 4);
            grow(I.t_cpc, 24);
            grow(I.t_dist, 8);
            grow(I.t_bdy, 4);
            std::unique_ptr<Buffer> a(new Buffer(nc * 4)), b(new Buffer(nc * 4));
            HIPCHK(hipMemcpy(a->p, fr->p, nf * 4, hipMemcpyDeviceToDevice));
            fr.swap(a);
            fr2.swap(b);
            cap = nc;
        }
        hipLaunchKernelGGL(k_tbfs_emit, dim3((unsigned)nb), dim3(256), 0, st, slot, ns, bcount, cand, ccp, cdist, cbdy, nt,
                           vis, I.t_set.as<int32_t>(), I.t_cpc.as<double>(), I.t_dist.as<double>(),
                           I.t_bdy.as<int32_t>(), fr2->as<int32_t>());
        fr.swap(fr2);
        nt += cnt;
        nf = cnt;
    }
    HIPCHK(hipStreamSynchronize(st));
    I.nt = nt;
    // subset index of every tube node: vis holds it for members, sentinels elsewhere
    realloc<int32_t>(I.t_sub, N);
    {
        std::vector<int32_t> h(N);
        HIPCHK(hipMemcpy(h.data(), vis, N * 4, hipMemcpyDeviceToHost));
        for (auto& v : h) if (v >= kCand || v < 0) v = -1;
        HIPCHK(hipMemcpy(I.t_sub.p, h.data(), N * 4, hipMemcpyHostToDevice));
    }
    if (nt == 0) throw GeomError("IBC tagging subset is empty");
    // cp_diff and its on-curve fix-up
    realloc<double>(I.t_cpdiff, nt * 3);
    uint8_t* onibc = reinterpret_cast<uint8_t*>(ctx->workspace("ibc_onibc", nt));
    hipLaunchKernelGGL(k_cpdiff, dim3(nblk(nt)), dim3(256), 0, st, nt, I.t_set.as<int32_t>(), t->cpx.as<double>(),
                       I.t_cpc.as<double>(), I.spec.tol, I.t_cpdiff.as<double>(), onibc);
    int32_t* list = reinterpret_cast<int32_t*>(ctx->workspace("ibc_list", nt * 4));
    const int64_t nl = select_flags(ctx, onibc, nt, list);
    if (nl > 0)
        hipLaunchKernelGGL(k_onibc_fix, dim3(1), dim3(64), 0, st, list, nl, I.t_set.as<int32_t>(), I.t_sub.as<int32_t>(),
                           t->nbr.as<int32_t>(), I.spec.tol, I.t_cpdiff.as<double>());
    // frame-based directions
    const int S = t->p + 1, W = S * S * S;
    int* ec = reinterpret_cast<int*>(ctx->workspace("ibc_ec", (size_t)W * nt * 4));
    double* ev = ctx->workspace("ibc_ev", (size_t)W * nt * 8);
    build_E_ell(*t, I.t_cpc.as<double>(), nt, ec, ev, nullptr, nt, st, ctx);
    double* gsub = ctx->workspace("ibc_gsub", (size_t)9 * nt * 8);
    realloc<double>(I.t_dir, nt * 3);
    realloc<double>(I.t_bn, nt * 3);
    hipLaunchKernelGGL(k_frame, dim3(nblk(nt)), dim3(256), 0, st, nt, ec, ev, W, nt, tube_normals, I.t_set.as<int32_t>(),
                       I.t_sub.as<int32_t>(), t->nbr.as<int32_t>(), I.t_cpc.as<double>(), I.t_cpdiff.as<double>(), t->dx,
                       gsub, 0, I.t_dir.as<double>(), I.t_bn.as<double>());
    hipLaunchKernelGGL(k_frame, dim3(nblk(nt)), dim3(256), 0, st, nt, ec, ev, W, nt, tube_normals, I.t_set.as<int32_t>(),
                       I.t_sub.as<int32_t>(), t->nbr.as<int32_t>(), I.t_cpc.as<double>(), I.t_cpdiff.as<double>(), t->dx,
                       gsub, 1, I.t_dir.as<double>(), I.t_bn.as<double>());
    HIPCHK(hipStreamSynchronize(st));
}

// DOF subset (IBCDOFSubset::ComputeDOFSubset, IBCDOFSubset.cpp:10-93).
static void dof_subset(Tube* t, Ctx* ctx, IbcData& I) {
    const hipStream_t st = ctx->stream;
    uint8_t* f = reinterpret_cast<uint8_t*>(ctx->workspace("ibc_dflag", I.nt));
    hipLaunchKernelGGL(k_dof_flag, dim3(nblk(I.nt)), dim3(256), 0, st, I.nt, I.t_dist.as<double>(),
                       I.spec.dof_scale * t->r_t, f);
    int32_t* sel = reinterpret_cast<int32_t*>(ctx->workspace("ibc_dsel", I.nt * 4));
    I.nd = select_flags(ctx, f, I.nt, sel);
    if (I.nd == 0) throw GeomError("IBC DOF subset is empty");
    realloc<int32_t>(I.d_set, I.nd);
    realloc<double>(I.d_cpc, I.nd * 3);
    realloc<double>(I.d_dir, I.nd * 3);
    realloc<int32_t>(I.d_sub, t->N);
    hipLaunchKernelGGL(k_fill_i32, dim3(std::min<unsigned>(nblk(t->N), 4096)), dim3(256), 0, st, I.d_sub.as<int32_t>(),
                       t->N, -1);
    hipLaunchKernelGGL(k_dof_fill, dim3(nblk(I.nd)), dim3(256), 0, st, I.nd, sel, I.t_set.as<int32_t>(),
                       I.t_cpc.as<double>(), I.t_dir.as<double>(), I.d_set.as<int32_t>(), I.d_cpc.as<double>(),
                       I.d_dir.as<double>(), I.d_sub.as<int32_t>());
    HIPCHK(hipStreamSynchronize(st));
}

Problem* problem_build(Tube* t, const std::vector<IbcSpec>& specs, Ctx* bctx) {
    auto t0 = std::chrono::steady_clock::now();
    if ((int)specs.size() > kMaxIbc) throw ArgError("too many IBC curves");
    Ctx* ctx = bctx ? bctx : t->ctx;
    if (ctx->device != t->ctx->device) throw ArgError("the problem's context is on another device than the tube");
    const hipStream_t st = ctx->stream;
    std::unique_ptr<Problem> P(new Problem());
    P->ctx = ctx;
    P->tube = t;
    const int64_t N = t->N;
    P->N = N;
    const int S = t->p + 1, W = S * S * S;
    // E of the tube (tube-index columns)
    int* ect = reinterpret_cast<int*>(ctx->workspace("prob_ect", (size_t)W * N * 4));
    double* evt = ctx->workspace("prob_evt", (size_t)W * N * 8);
    build_E_ell(*t, t->cpx.as<double>(), N, ect, evt, nullptr, N, st, ctx);
    // whole-tube normals from the Jacobian of cp
    realloc<double>(P->normals, N * 3);
    double* G = ctx->workspace("prob_grad", (size_t)9 * N * 8);
    hipLaunchKernelGGL(k_grad_tube, dim3(nblk(N)), dim3(256), 0, st, N, t->nbr.as<int32_t>(), t->cpx.as<double>(), t->dx, G);
    hipLaunchKernelGGL(k_tube_normals, dim3(nblk(N)), dim3(256), 0, st, N, ect, evt, W, N, G, P->normals.as<double>());
    // subsets per curve
    P->ibc.resize(specs.size());
    for (size_t c = 0; c < specs.size(); ++c) {
        IbcData& I = P->ibc[c];
        I.spec = specs[c];
        if (I.spec.order != 1 && I.spec.order != 2) throw ArgError("boundary_order must be 1 or 2");
        if (I.spec.surf.type == "Polyline") {   // segment hierarchy (its vertex copy serves the parameter)
            segment_build(I.bvh, I.spec.surf.verts);
            const MeshDev md = I.bvh.dev();
            I.sd = make_surfdev(I.spec.surf, I.bvh.v.as<double>(), &md);
        } else {
            if (!I.spec.surf.verts.empty()) {
                realloc<double>(I.verts, I.spec.surf.verts.size());
                HIPCHK(hipMemcpy(I.verts.p, I.spec.surf.verts.data(), I.spec.surf.verts.size() * 8, hipMemcpyHostToDevice));
            }
            I.sd = make_surfdev(I.spec.surf, I.verts.as<double>());
        }
        tagging_subset(t, ctx, I, P->normals.as<double>());
    }
    for (auto& I : P->ibc) dof_subset(t, ctx, I);
    // IBCSubsets::IBCStartingColumns / TotalDOFs (IBCSubsets.cpp:74-116)
    int64_t G_ = 0, K_ = 0, ident_start = 0;
    bool any2 = false;
    for (size_t c = 0; c < P->ibc.size(); ++c) {
        IbcData& I = P->ibc[c];
        I.start = N + G_;
        G_ += I.nd;
        I.ndir = I.second_dirichlet() ? I.nd : 0;   // not oriented: every DOF takes boundary_type[0]
        if (I.second_dirichlet() && I.spec.btype[0] != 0) throw ArgError("oriented IBCs are not supported");
        any2 |= I.second_dirichlet();
    }
    for (auto& I : P->ibc) {
        I.dir_start = K_;
        K_ += I.ndir;
    }
    ident_start = any2 ? N + G_ : 0;
    P->G = G_;
    P->K = K_;
    P->D = N + G_ + K_;
    P->identity_rows_start = ident_start;
    const int64_t D = P->D;
    P->ew = W + (any2 ? 1 : 0);
    realloc<int>(P->ecol, (size_t)P->ew * D);
    realloc<double>(P->eval64, (size_t)P->ew * D);
    realloc<float>(P->eval32, (size_t)P->ew * D);
    realloc<int>(P->lcol, (size_t)6 * D);
    realloc<double>(P->lval64, (size_t)6 * D);
    realloc<float>(P->lval32, (size_t)6 * D);
    realloc<double>(P->ldiag64, D);
    realloc<float>(P->ldiag32, D);
    realloc<uint8_t>(P->flag, D);
    AsmArgs a{};
    a.nibc = (int)P->ibc.size();
    for (int c = 0; c < a.nibc; ++c) {
        IbcData& I = P->ibc[c];
        a.ib[c] = IbcDev{I.t_sub.as<int32_t>(), I.t_bdy.as<int32_t>(), I.t_dir.as<double>(), I.d_sub.as<int32_t>(),
                         I.d_dir.as<double>(), I.start};
    }
    a.N = N;
    a.D = D;
    a.ld = D;
    a.ew = P->ew;
    a.h = t->dx;
    int* err = reinterpret_cast<int*>(ctx->workspace("prob_err", 64));
    HIPCHK(hipMemsetAsync(err, 0, 4, st));
    hipLaunchKernelGGL(k_asm_tube, dim3(nblk(N)), dim3(256), 0, st, a, ect, evt, W, N, t->nbr.as<int32_t>(),
                       P->ecol.as<int>(), P->eval64.as<double>(), P->eval32.as<float>(), P->lcol.as<int>(),
                       P->lval64.as<double>(), P->lval32.as<float>(), P->ldiag64.as<double>(), P->ldiag32.as<float>(),
                       P->flag.as<uint8_t>(), err);
    realloc<double>(P->rhs_pts, D * 3);
    realloc<int32_t>(P->rhs_kind, D);
    realloc<double>(P->qpts, D * 3);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * D)), dim3(256), 0, st, P->qpts.as<double>(), 3 * D, (double)NAN);
    HIPCHK(hipMemcpyAsync(P->qpts.p, t->cpx.p, (size_t)N * 24, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_rhs_tube, dim3(nblk(N)), dim3(256), 0, st, N, t->cpx.as<double>(), P->rhs_pts.as<double>(),
                       P->rhs_kind.as<int32_t>());
    for (int c = 0; c < a.nibc; ++c) {
        IbcData& I = P->ibc[c];
        double* q = ctx->workspace("prob_q", I.nd * 24);
        hipLaunchKernelGGL(k_ghost_query, dim3(nblk(I.nd)), dim3(256), 0, st, I.nd, I.spec.order, I.d_set.as<int32_t>(),
                           I.d_cpc.as<double>(), t->rec.as<int64_t>(), t->N1d, t->xstart, t->dx, t->sd, q);
        int* ecg = reinterpret_cast<int*>(ctx->workspace("prob_ecg", (size_t)W * I.nd * 4));
        double* evg = ctx->workspace("prob_evg", (size_t)W * I.nd * 8);
        build_E_ell(*t, q, I.nd, ecg, evg, nullptr, I.nd, st, ctx);
        HIPCHK(hipMemcpyAsync(P->qpts.as<double>() + 3 * I.start, q, (size_t)I.nd * 24, hipMemcpyDeviceToDevice, st));
        const int64_t ident0 = ident_start + I.dir_start;
        hipLaunchKernelGGL(k_asm_ghost, dim3(nblk(I.nd)), dim3(256), 0, st, a, c, I.nd, I.spec.order, I.spec.btype[0],
                           ident0, ecg, evg, W, I.d_set.as<int32_t>(), t->nbr.as<int32_t>(), P->ecol.as<int>(),
                           P->eval64.as<double>(), P->eval32.as<float>(), P->lcol.as<int>(), P->lval64.as<double>(),
                           P->lval32.as<float>(), P->ldiag64.as<double>(), P->ldiag32.as<float>(), P->flag.as<uint8_t>(),
                           err);
        hipLaunchKernelGGL(k_rhs_points, dim3(nblk(I.nd)), dim3(256), 0, st, I.nd, I.start, ident0, I.spec.order,
                           I.spec.btype[0], I.d_cpc.as<double>(), P->rhs_pts.as<double>(), P->rhs_kind.as<int32_t>());
        HIPCHK(hipStreamSynchronize(st));
    }
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h & 1) throw GeomError("a stencil entry crosses more than one IBC (curves too close together)");
    P->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return P.release();
}

// Lfull rows N..D-1 (ghost and identity rows, whose internal index equals the
// reference one) for explicit stepping: slot 0 the diagonal, then the six
// off-diagonal slots; columns of tube nodes renumbered by rlex when the
// operator runs in the lexicographic order.
__global__ void k_held_rows(int64_t N, int64_t D, const int* lcol, const double* lval, const double* ldiag,
                            const int32_t* rlex, int* hrow, int* hcol, double* hval) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nh = D - N;
    if (k >= nh) return;
    const int64_t r = N + k;
    hrow[k] = (int)r;
    hcol[k] = (int)r;
    hval[k] = ldiag[r];
    for (int s = 0; s < 6; ++s) {
        const int c = lcol[(int64_t)s * D + r];
        hcol[(s + 1) * nh + k] = (rlex && c < N) ? rlex[c] : c;
        hval[(s + 1) * nh + k] = lval[(int64_t)s * D + r];
    }
}

static void attach_held_rows(Op& op, Problem* P, const int32_t* rlex) {
    const int64_t nh = P->D - P->N;
    if (nh <= 0) return;
    const hipStream_t st = P->ctx->stream;
    op.hrow = Buffer((size_t)nh * 4);
    op.hcol = Buffer((size_t)nh * 7 * 4);
    op.hval = Buffer((size_t)nh * 7 * 8);
    hipLaunchKernelGGL(k_held_rows, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, P->N, P->D,
                       P->lcol.as<int>(), P->lval64.as<double>(), P->ldiag64.as<double>(), rlex,
                       op.hrow.as<int>(), op.hcol.as<int>(), op.hval.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    op.nh = nh;
    op.hw = 7;
}

Op* op_from_problem(Problem* P, double m, double n, int precond, int precond_iters) {
    const hipStream_t st = P->ctx->stream;
    const int64_t D = P->D;
    std::unique_ptr<Op> op(new Op(P->ctx, D, 0));
    op->m = m;
    op->n = n;
    op->precond = precond;
    op->precond_iters = precond_iters;
    TpSource src;
    src.D = D;
    src.N = P->N;
    src.ecol = P->ecol.as<int>();
    src.eval = P->eval64.as<double>();
    src.ew = P->ew;
    src.lcol = P->lcol.as<int>();
    src.lval = P->lval64.as<double>();
    src.lw = 6;
    src.ldiag = P->ldiag64.as<double>();
    src.flag = P->flag.as<uint8_t>();
    src.qpts = P->qpts.as<double>();
    if (!tp_disabled() && op_build_tp(*op, *P->tube, src)) {
        attach_held_rows(*op, P, P->tube->rlex.as<int32_t>());
        return op.release();
    }
    op->ew = P->ew;
    op->lw = 6;
    auto cp = [&](Buffer& dst, const Buffer& src) {
        dst.~Buffer();
        new (&dst) Buffer(src.bytes);
        HIPCHK(hipMemcpyAsync(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice, st));
    };
    cp(op->ecol, P->ecol);
    cp(op->eval64, P->eval64);
    cp(op->eval32, P->eval32);
    cp(op->lcol, P->lcol);
    cp(op->lval64, P->lval64);
    cp(op->lval32, P->lval32);
    cp(op->ldiag64, P->ldiag64);
    cp(op->ldiag32, P->ldiag32);
    cp(op->flag, P->flag);
    HIPCHK(hipStreamSynchronize(st));
    attach_held_rows(*op, P, nullptr);
    op->m = m;
    op->n = n;
    op->precond = precond;
    op->precond_iters = precond_iters;
    return op.release();
}

// ---------------------------------------------------------------- plotting rows
// Per-curve tables of the plotting interpolation (device array).
struct PlotIbc {
    SurfDev sd;
    const int32_t* t_sub;
    const double* t_cpdiff;
    const double* t_bn;
    double tol;
};

// AddIBCPlotInterpolation, not oriented (IBCMatrixManipulations.cpp:74-106
// with the plot points' directions of PlotPointsCPDiff :254-416): for every
// curve whose distance from the plot point is below 2 r_t, cp_diff = x - cp_c
// (the nearest tube node's tagging cp_diff when x lies on the curve), the
// binormal interpolated without orientation from the tagging subset at cp_c
// (sign-aligned to the entry of smallest subset index, normalised), and the
// direction (b . cp_diff) b; stencil entries then move to ghost columns as
// for the tube rows.  Columns ascending in the output row.
__global__ void k_plot_rows(const PlotIbc* pc, const IbcDev* ib, int nibc, GridDev g, int p, const int32_t* gidx,
                            int64_t nq, const double* xq, const int* ecol, const double* eval, int W, int* col_out,
                            double* val_out, int* err) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double* x = xq + 3 * q;
    double dirs[kMaxIbc][3];
    const double* rowdir[kMaxIbc];
    const int S = p + 1;
    for (int c = 0; c < nibc; ++c) {
        rowdir[c] = nullptr;
        double cc[3], dist;
        int bd;
        cp_surface(pc[c].sd, x, cc, dist, bd);
        double diff[3] = {x[0] - cc[0], x[1] - cc[1], x[2] - cc[2]};
        const double nd = sqrt(dot3(diff, diff));
        if (!(nd < 2.0 * g.r_t)) continue;
        if (nd < pc[c].tol) {   // Tube.h:105-110 nearest node
            int64_t t[3];
            bool in = true;
            for (int d = 0; d < 3; ++d) {
                t[d] = (int64_t)rint((x[d] - g.xstart) / g.dx);
                in = in && t[d] >= 0 && t[d] < g.N1d;
            }
            if (in) {
                const int32_t ni = gidx[(t[0] * g.N1d + t[1]) * g.N1d + t[2]];
                const int32_t ts = (ni >= 0 && ni < kCand) ? pc[c].t_sub[ni] : -1;
                if (ts >= 0)
                    for (int k = 0; k < 3; ++k) diff[k] = pc[c].t_cpdiff[3 * (int64_t)ts + k];
            }
        }
        double bn[3] = {0.0, 0.0, 0.0};
        int64_t b[3];
        double w[3][5];
        if (stencil_weights(g, p, cc, b, w)) {
            int32_t first = 0x7fffffff;
            for (int i = 0; i < S; ++i)
                for (int j = 0; j < S; ++j)
                    for (int k = 0; k < S; ++k) {
                        const int32_t gv = gidx[((b[0] + i) * g.N1d + (b[1] + j)) * g.N1d + (b[2] + k)];
                        if (gv < 0 || gv >= kCand) continue;
                        const int32_t ts = pc[c].t_sub[gv];
                        if (ts >= 0 && ts < first) first = ts;
                    }
            if (first != 0x7fffffff) {
                const double* f0 = pc[c].t_bn + 3 * (int64_t)first;
                for (int i = 0; i < S; ++i)
                    for (int j = 0; j < S; ++j)
                        for (int k = 0; k < S; ++k) {
                            const int32_t gv = gidx[((b[0] + i) * g.N1d + (b[1] + j)) * g.N1d + (b[2] + k)];
                            if (gv < 0 || gv >= kCand) continue;
                            const int32_t ts = pc[c].t_sub[gv];
                            if (ts < 0) continue;
                            const double* v = pc[c].t_bn + 3 * (int64_t)ts;
                            const double wv = w[2][k] * w[1][j] * w[0][i];
                            const double sg = dot3(f0, v) > 0.0 ? 1.0 : -1.0;
                            for (int d = 0; d < 3; ++d) bn[d] += sg * (wv * v[d]);
                        }
                const double nb = sqrt(dot3(bn, bn));
                if (nb > 0.0)
                    for (int d = 0; d < 3; ++d) bn[d] /= nb;
            }
        }
        const double sb = dot3(bn, diff);
        for (int d = 0; d < 3; ++d) dirs[c][d] = sb * bn[d];
        rowdir[c] = dirs[c];
    }
    int cols[64];
    double vals[64];
    for (int s = 0; s < W; ++s) {
        const int64_t col = ecol[(int64_t)s * nq + q];
        int nm;
        const int64_t nc = moved_col(ib, nibc, rowdir, col, 1.0, &nm);
        if (nm > 1) atomicOr(err, 1);
        const double v = eval[(int64_t)s * nq + q];
        int k = s;   // insertion by column
        while (k > 0 && cols[k - 1] > (int)nc) {
            cols[k] = cols[k - 1];
            vals[k] = vals[k - 1];
            --k;
        }
        cols[k] = (int)nc;
        vals[k] = v;
    }
    for (int s = 0; s < W; ++s) {
        col_out[q * W + s] = cols[s];
        val_out[q * W + s] = vals[s];
    }
}

void problem_plot_matrix(Problem* P, const double* xq, int64_t nq, int32_t* rowptr, int32_t* col, double* val) {
    Tube* t = P->tube;
    Ctx* ctx = P->ctx;
    const hipStream_t st = ctx->stream;
    const int S = t->p + 1, W = S * S * S;
    if (W > 64) throw ArgError("plotting rows support interpolation degree <= 3");
    const int nibc = (int)P->ibc.size();
    std::vector<PlotIbc> hp((size_t)nibc);
    std::vector<IbcDev> hi((size_t)nibc);
    for (int c = 0; c < nibc; ++c) {
        IbcData& I = P->ibc[(size_t)c];
        hp[(size_t)c] = PlotIbc{I.sd, I.t_sub.as<int32_t>(), I.t_cpdiff.as<double>(), I.t_bn.as<double>(), I.spec.tol};
        hi[(size_t)c] = IbcDev{I.t_sub.as<int32_t>(), I.t_bdy.as<int32_t>(), I.t_dir.as<double>(),
                               I.d_sub.as<int32_t>(), I.d_dir.as<double>(), I.start};
    }
    const size_t tb = (sizeof(PlotIbc) * nibc + 255) / 256 * 256, ib_bytes = (sizeof(IbcDev) * nibc + 255) / 256 * 256;
    char* w = reinterpret_cast<char*>(ctx->workspace("plot_tables", tb + ib_bytes + 256));
    auto* dpc = reinterpret_cast<PlotIbc*>(w);
    auto* dib = reinterpret_cast<IbcDev*>(w + tb);
    int* err = reinterpret_cast<int*>(w + tb + ib_bytes);
    HIPCHK(hipMemcpyAsync(dpc, hp.data(), sizeof(PlotIbc) * nibc, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dib, hi.data(), sizeof(IbcDev) * nibc, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(err, 0, 4, st));
    double* dxq = ctx->workspace("plot_xq", (size_t)nq * 24);
    HIPCHK(hipMemcpyAsync(dxq, xq, (size_t)nq * 24, hipMemcpyHostToDevice, st));
    int* ec = reinterpret_cast<int*>(ctx->workspace("plot_ec", (size_t)W * nq * 4));
    double* ev = ctx->workspace("plot_ev", (size_t)W * nq * 8);
    build_E_ell(*t, dxq, nq, ec, ev, nullptr, nq, st, ctx);
    int* oc = reinterpret_cast<int*>(ctx->workspace("plot_oc", (size_t)W * nq * 4));
    double* ov = ctx->workspace("plot_ov", (size_t)W * nq * 8);
    GridDev g{t->N1d, t->xstart, t->dx, t->r_t};
    hipLaunchKernelGGL(k_plot_rows, dim3(nblk(nq)), dim3(256), 0, st, dpc, dib, nibc, g, t->p, t->gidx.as<int32_t>(),
                       nq, dxq, ec, ev, W, oc, ov, err);
    HIPCHK(hipGetLastError());
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(col, oc, (size_t)W * nq * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(val, ov, (size_t)W * nq * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h & 1) throw GeomError("a plotting stencil entry crosses more than one IBC (curves too close together)");
    for (int64_t r = 0; r <= nq; ++r) rowptr[r] = (int32_t)(r * W);
}

// ---------------------------------------------------------------- curve parameter
// Surface::IBCParameter (src/Surface.cpp:2731-2788) of points on a curve.
// Parameterisation points of the curve (Surface::GetSurfaceParameterization,
// :193-366): a polyline's vertices, 100 points of a Circle3D (paramCircle3D,
// :1883-1904) or of an Arc (paramArc, :1907-1924), computed on the host in the
// reference's order.
static std::vector<double> curve_param_points(const SurfaceHost& s) {
    if (s.type == "Polyline") {
        // a closed polyline's vertex list holds each vertex once in the reference (its last edge
        // returns to vertex 0, SurfaceSpecifier.cpp:334-341); the chain here repeats vertex 0
        std::vector<double> v = s.verts;
        if (!s.is_open && v.size() >= 6) v.resize(v.size() - 3);
        return v;
    }
    const auto& P = s.params;
    const int Np = 100;
    std::vector<double> xp(3 * Np, 0.0);
    if (s.type == "Circle3D") {
        double nr[3] = {P[4], P[5], P[6]}, t1[3], t2[3];
        if (std::fabs(nr[0]) > std::fabs(nr[1])) { t1[0] = -nr[2]; t1[1] = 0; t1[2] = nr[0]; }
        else { t1[0] = 0; t1[1] = nr[2]; t1[2] = -nr[1]; }
        t2[0] = nr[1] * t1[2] - nr[2] * t1[1];
        t2[1] = nr[2] * t1[0] - nr[0] * t1[2];
        t2[2] = nr[0] * t1[1] - nr[1] * t1[0];
        const double n1 = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
        const double n2 = std::sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
        for (int k = 0; k < 3; ++k) { t1[k] /= n1; t2[k] /= n2; }
        const double dth = 2 * M_PI / (Np - 1);
        for (int i = 0; i < Np; ++i) {
            const double th = i * dth, a = std::cos(th), b = std::sin(th);
            for (int k = 0; k < 3; ++k) xp[3 * i + k] = P[1 + k] + P[0] * (a * t1[k] + b * t2[k]);
        }
    } else if (s.type == "Arc") {
        const double dth = (P[2] - P[1]) / (Np - 1);
        for (int i = 0; i < Np; ++i) {
            const double th = P[1] + i * dth;
            xp[3 * i] = P[0] * std::cos(th);
            xp[3 * i + 1] = P[0] * std::sin(th);
        }
    } else {
        throw ArgError("no parameterisation points for IBC curve type '" + s.type + "'");
    }
    return xp;
}

__device__ __forceinline__ double dist3(const double* a, const double* b) {
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

// Normalised arc length of the point's closest point on the curve along the
// parameterisation points (the reference's rule for polylines and curves
// with parameterisation points): IBCNearestTwoXpIndices (:2661-2728) finds
// the nearest point xp[mi] to the closest point and the neighbour xp[nn]
// bounding the edge it lies on; the parameter is the length up to the edge
// plus the distance along it, over the total length (prefix[k]: length of the
// first k edges, summed in order).  err: no edge found (the reference throws).
__global__ void k_ibc_param_arclength(SurfDev sd, const double* xp, int64_t nxp, const double* prefix,
                                      const double* pts, int64_t n, double* out, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    double c[3], dist;
    int b;
    cp_surface(sd, x, c, dist, b);
    int64_t mi = 0;
    double md = DBL_MAX;
    for (int64_t k = 0; k < nxp; ++k) {
        const double d = dist3(xp + 3 * k, c);
        if (d < md) { md = d; mi = k; }
    }
    int64_t nn;
    if (mi == 0) {
        nn = 1;
    } else if (mi == nxp - 1) {
        nn = mi - 1;
    } else if (dist3(c, xp + 3 * mi) < 1e-6) {   // on the point: the shorter edge
        nn = dist3(xp + 3 * (mi + 1), c) > dist3(xp + 3 * (mi - 1), c) ? mi - 1 : mi + 1;
    } else {
        const double* q = xp + 3 * mi;
        const double a[3] = {c[0] - q[0], c[1] - q[1], c[2] - q[2]};
        double e[3], t[3];
        for (int k = 0; k < 3; ++k) e[k] = xp[3 * (mi - 1) + k] - q[k];
        t[0] = a[1] * e[2] - a[2] * e[1];
        t[1] = a[2] * e[0] - a[0] * e[2];
        t[2] = a[0] * e[1] - a[1] * e[0];
        if (sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) < 1e-4) {
            nn = mi - 1;
        } else {
            for (int k = 0; k < 3; ++k) e[k] = xp[3 * (mi + 1) + k] - q[k];
            t[0] = a[1] * e[2] - a[2] * e[1];
            t[1] = a[2] * e[0] - a[0] * e[2];
            t[2] = a[0] * e[1] - a[1] * e[0];
            if (sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) < 1e-4) {
                nn = mi + 1;
            } else {
                atomicOr(err, 1);
                out[i] = NAN;
                return;
            }
        }
    }
    const double to_seg = prefix[mi < nn ? mi : nn];
    const double par = mi < nn ? to_seg + md : to_seg + dist3(xp + 3 * nn, xp + 3 * mi) - md;
    out[i] = par / prefix[nxp - 1];
}

// The nearest parameterisation point's parameter: the reference's rule for
// curves given by parameterised samples (Dumbbell, Planar Curve, Torus Knot,
// :2737-2751): the first point at the smallest distance from x itself.
__global__ void k_ibc_param_vertex(const double* xp, int64_t nxp, const double* vparam, const double* pts, int64_t n,
                                   double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    int64_t mi = 0;
    double md = DBL_MAX;
    for (int64_t k = 0; k < nxp; ++k) {
        const double d = dist3(x, xp + 3 * k);
        if (d < md) { md = d; mi = k; }
    }
    out[i] = vparam[mi];
}

void problem_ibc_parameter(Problem* P, int c, const double* pts, int64_t n, const double* vparam, double* out) {
    if (c < 0 || c >= (int)P->ibc.size()) throw ArgError("no such IBC");
    if (n < 0 || (n > 0 && (!pts || !out))) throw ArgError("bad points");
    if (n == 0) return;
    const IbcData& I = P->ibc[(size_t)c];
    if (I.spec.surf.type == "Point") {   // :2733-2736
        std::fill(out, out + n, 0.0);
        return;
    }
    const std::vector<double> xp = curve_param_points(I.spec.surf);
    const int64_t nxp = (int64_t)xp.size() / 3;
    if (nxp < 2) throw ArgError("a curve parameter needs two parameterisation points");
    std::vector<double> prefix((size_t)nxp, 0.0);
    for (int64_t k = 0; k + 1 < nxp; ++k) {   // total length summed edge by edge (:2763-2771)
        double d2 = 0.0;
        for (int q = 0; q < 3; ++q) d2 += (xp[3 * k + q] - xp[3 * (k + 1) + q]) * (xp[3 * k + q] - xp[3 * (k + 1) + q]);
        prefix[(size_t)k + 1] = prefix[(size_t)k] + std::sqrt(d2);
    }
    Ctx* ctx = P->ctx;
    const hipStream_t st = ctx->stream;
    Buffer dxp(xp.size() * 8), dpre(prefix.size() * 8), dpts((size_t)n * 24), dout((size_t)n * 8), derr(64),
        dvp(vparam ? (size_t)nxp * 8 : 0);
    HIPCHK(hipMemcpyAsync(dxp.p, xp.data(), xp.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dpre.p, prefix.data(), prefix.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dpts.p, pts, (size_t)n * 24, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(derr.p, 0, 4, st));
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (vparam) {
        HIPCHK(hipMemcpyAsync(dvp.p, vparam, (size_t)nxp * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_ibc_param_vertex, dim3(nb), dim3(256), 0, st, dxp.as<double>(), nxp, dvp.as<double>(),
                           dpts.as<double>(), n, dout.as<double>());
    } else {
        hipLaunchKernelGGL(k_ibc_param_arclength, dim3(nb), dim3(256), 0, st, I.sd, dxp.as<double>(), nxp,
                           dpre.as<double>(), dpts.as<double>(), n, dout.as<double>(), derr.as<int>());
    }
    HIPCHK(hipGetLastError());
    int e = 0;
    HIPCHK(hipMemcpyAsync(out, dout.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&e, derr.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (e) throw GeomError("IBC parameter: a point's closest point lies on no edge of the curve "
                           "(Surface::IBCNearestTwoXpIndices)");
}

}  // namespace cpmgpu
