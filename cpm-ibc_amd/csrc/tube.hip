// GPU tube construction (reference: src/Tube.cpp, src/Grid.cpp,
// src/Interpolation.cpp, src/FDMatrices.cpp, src/Surface.cpp).
//
// The reference grows the tube by a FIFO breadth-first traversal of the
// implicit (2^n+1)^3 grid (Tube.cpp:429-454), so a node's index is its
// discovery order.  Here the traversal is level-synchronous: every node of
// level l+1 is discovered by a level-l parent, and the FIFO gives it the
// position of its first discoverer (parent rank, then direction +x,-x,+y,-y,
// +z,-z).  Each level therefore
//   1. expands the frontier: each (parent, direction) claims its neighbour
//      with atomicMin(key = rank*6 + dir) on a dense int32 grid,
//   2. evaluates the closest point of each newly claimed node and keeps it if
//      |dist| <= r_t, writing its slot key,
//   3. compacts the slot keys in order (block counts, scan, emit), which
//      assigns tube indices identical to the reference's.
// Closest points and weights are computed with FP contraction off so that
// FP64 values follow the reference's rounding sequence.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "tube.h"

#pragma clang fp contract(off)

namespace cpmgpu {


// ---------------------------------------------------------------- closest points
__device__ __forceinline__ double dsq(double a) { return a * a; }

// Ellipsoid closest point: Eberly, "Distance from a Point to an Ellipse, an
// Ellipsoid, or a Hyperellipsoid" (Geometric Tools, 2013).  In the first
// octant with semi-axes e0 >= e1 >= e2 the closest point is
// x_i = r_i y_i / (s + r_i), r_i = (e_i / e_last)^2, where s is the root of
// sum_i (r_i z_i / (s + r_i))^2 = 1 (z_i = y_i / e_i), found by bisection until
// the bracket stops shrinking.  Only + - * / and sqrt: the host oracle gets
// the same bits.  (The reference runs LBFGS from the nearest of 10000
// parameter points, Surface.cpp:1363-1400,1459-1530, to tolerance 1e-6.)
__device__ double ell_robust_length(const double* v, int n) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = fmax(m, fabs(v[i]));
    if (m == 0.0) return 0.0;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (v[i] / m) * (v[i] / m);
    return m * sqrt(s);
}

__device__ double ell_root(int n, const double* r, const double* z, double g) {
    double num[3];
    for (int i = 0; i < n - 1; ++i) num[i] = r[i] * z[i];
    num[n - 1] = z[n - 1];
    double s0 = z[n - 1] - 1.0, s1 = g < 0.0 ? 0.0 : ell_robust_length(num, n) - 1.0, s = 0.0;
    for (int it = 0; it < 1100; ++it) {
        s = 0.5 * (s0 + s1);
        if (s == s0 || s == s1) break;
        double gs = -1.0;
        for (int i = 0; i < n; ++i) {
            const double q = num[i] / (s + r[i]);
            gs += q * q;
        }
        if (gs > 0.0) s0 = s;
        else if (gs < 0.0) s1 = s;
        else break;
    }
    return s;
}

__device__ void ell_point2(double e0, double e1, double y0, double y1, double& x0, double& x1) {
    if (y1 > 0.0) {
        if (y0 > 0.0) {
            const double z[2] = {y0 / e0, y1 / e1};
            const double g = z[0] * z[0] + z[1] * z[1] - 1.0;
            if (g != 0.0) {
                const double r[2] = {(e0 / e1) * (e0 / e1), 1.0};
                const double s = ell_root(2, r, z, g);
                x0 = r[0] * y0 / (s + r[0]);
                x1 = y1 / (s + 1.0);
            } else {
                x0 = y0;
                x1 = y1;
            }
        } else {
            x0 = 0.0;
            x1 = e1;
        }
    } else {
        const double num0 = e0 * y0, den0 = e0 * e0 - e1 * e1;
        if (num0 < den0) {
            const double xd0 = num0 / den0;
            x0 = e0 * xd0;
            x1 = e1 * sqrt(1.0 - xd0 * xd0);
        } else {
            x0 = e0;
            x1 = 0.0;
        }
    }
}

__device__ void ell_point3(const double* e, const double* y, double* x) {
    if (y[2] > 0.0) {
        if (y[1] > 0.0) {
            if (y[0] > 0.0) {
                const double z[3] = {y[0] / e[0], y[1] / e[1], y[2] / e[2]};
                const double g = z[0] * z[0] + z[1] * z[1] + z[2] * z[2] - 1.0;
                if (g != 0.0) {
                    const double r[3] = {(e[0] / e[2]) * (e[0] / e[2]), (e[1] / e[2]) * (e[1] / e[2]), 1.0};
                    const double s = ell_root(3, r, z, g);
                    x[0] = r[0] * y[0] / (s + r[0]);
                    x[1] = r[1] * y[1] / (s + r[1]);
                    x[2] = y[2] / (s + 1.0);
                } else {
                    x[0] = y[0];
                    x[1] = y[1];
                    x[2] = y[2];
                }
            } else {
                x[0] = 0.0;
                ell_point2(e[1], e[2], y[1], y[2], x[1], x[2]);
            }
        } else if (y[0] > 0.0) {
            x[1] = 0.0;
            ell_point2(e[0], e[2], y[0], y[2], x[0], x[2]);
        } else {
            x[0] = 0.0;
            x[1] = 0.0;
            x[2] = e[2];
        }
    } else {
        const double den0 = e[0] * e[0] - e[2] * e[2], den1 = e[1] * e[1] - e[2] * e[2];
        const double num0 = e[0] * y[0], num1 = e[1] * y[1];
        bool done = false;
        if (num0 < den0 && num1 < den1) {
            const double xd0 = num0 / den0, xd1 = num1 / den1;
            const double disc = 1.0 - xd0 * xd0 - xd1 * xd1;
            if (disc > 0.0) {
                x[0] = e[0] * xd0;
                x[1] = e[1] * xd1;
                x[2] = e[2] * sqrt(disc);
                done = true;
            }
        }
        if (!done) {
            x[2] = 0.0;
            ell_point2(e[0], e[1], y[0], y[1], x[0], x[1]);
        }
    }
}

__device__ void cp_ellipsoid(const double* abc, const double* x, double* cp) {
    int perm[3] = {0, 1, 2};   // axes by decreasing semi-axis
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (abc[perm[j]] > abc[perm[i]]) {
                const int t = perm[i];
                perm[i] = perm[j];
                perm[j] = t;
            }
    double e[3], y[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e[k] = abc[perm[k]];
        y[k] = fabs(x[perm[k]]);
    }
    ell_point3(e, y, q);
    for (int k = 0; k < 3; ++k) cp[perm[k]] = copysign(q[k], x[perm[k]]);
}

__device__ void cp_surface(const SurfDev& s, const double* x, double* c, double& dist, int& bdy) {
    const double* P = s.P;
    bdy = 0;
    switch (s.type) {
    case SURF_SPHERE: {   // Surface.cpp:710-740
        const double R = P[0];
        const double cen[3] = {P[1], P[2], P[3]};
        const double xs[3] = {x[0] - cen[0], x[1] - cen[1], x[2] - cen[2]};
        const double theta = atan2(xs[1], xs[0]);
        dist = sqrt(dsq(xs[0]) + dsq(xs[1]) + dsq(xs[2]));
        const double phi = dist != 0 ? acos(xs[2] / dist) : 0.0;
        c[0] = R * sin(phi) * cos(theta) + cen[0];
        c[1] = R * sin(phi) * sin(theta) + cen[1];
        c[2] = R * cos(phi) + cen[2];
        dist -= R;
        break;
    }
    case SURF_HEMISPHERE: {   // :744-781
        const double R = P[0];
        const double theta = atan2(x[1], x[0]);
        if (x[2] > 0) {
            dist = sqrt(dsq(x[0]) + dsq(x[1]) + dsq(x[2]));
            const double phi = dist != 0 ? acos(x[2] / dist) : 0.0;
            c[0] = R * sin(phi) * cos(theta);
            c[1] = R * sin(phi) * sin(theta);
            c[2] = R * cos(phi);
            dist -= R;
        } else {
            c[0] = R * cos(theta);
            c[1] = R * sin(theta);
            c[2] = 0;
            dist = sqrt(dsq(x[0] - c[0]) + dsq(x[1] - c[1]) + dsq(x[2] - c[2]));
            bdy = 1;
        }
        break;
    }
    case SURF_TORUS: {   // :800-837 (meridian circle of radius P[1] centred at (P[0], 0))
        const double theta = atan2(x[1], x[0]);
        const double r = sqrt(x[0] * x[0] + x[1] * x[1]);
        const double cyl[2] = {r, x[2]};
        const double sh[2] = {cyl[0] - P[0], cyl[1] - 0.0};
        const double th2 = atan2(sh[1], sh[0]);
        const double cc[2] = {P[1] * cos(th2) + P[0], P[1] * sin(th2) + 0.0};
        dist = sqrt(dsq(cyl[0] - cc[0]) + dsq(cyl[1] - cc[1]));
        c[0] = cc[0] * cos(theta);
        c[1] = cc[0] * sin(theta);
        c[2] = cc[1];
        break;
    }
    case SURF_POINT: {   // :840-848
        c[0] = P[0]; c[1] = P[1]; c[2] = P[2];
        dist = sqrt(dsq(x[0] - c[0]) + dsq(x[1] - c[1]) + dsq(x[2] - c[2]));
        break;
    }
    case SURF_CIRCLE3D: {   // :464-503
        const double R = P[0];
        const double cen[3] = {P[1], P[2], P[3]}, nr[3] = {P[4], P[5], P[6]};
        double t1[3], t2[3];
        if (fabs(nr[0]) > fabs(nr[1])) { t1[0] = -nr[2]; t1[1] = 0; t1[2] = nr[0]; }
        else { t1[0] = 0; t1[1] = nr[2]; t1[2] = -nr[1]; }
        t2[0] = nr[1] * t1[2] - nr[2] * t1[1];
        t2[1] = nr[2] * t1[0] - nr[0] * t1[2];
        t2[2] = nr[0] * t1[1] - nr[1] * t1[0];
        const double cx[3] = {x[0] - cen[0], x[1] - cen[1], x[2] - cen[2]};
        const double a = (nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]) * (t1[0] * cx[0] + t1[1] * cx[1] + t1[2] * cx[2]);
        const double b = t2[0] * cx[0] + t2[1] * cx[1] + t2[2] * cx[2];
        const double q[3] = {a * t1[0] + b * t2[0], a * t1[1] + b * t2[1], a * t1[2] + b * t2[2]};
        const double nq = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (nq > 0) {
            for (int d = 0; d < 3; ++d) c[d] = cen[d] + R * q[d] / nq;
            const double h = nr[0] * cx[0] + nr[1] * cx[1] + nr[2] * cx[2];
            const double ncx[3] = {nr[1] * cx[2] - nr[2] * cx[1], nr[2] * cx[0] - nr[0] * cx[2], nr[0] * cx[1] - nr[1] * cx[0]};
            const double radial = sqrt(ncx[0] * ncx[0] + ncx[1] * ncx[1] + ncx[2] * ncx[2]) - R;
            dist = sqrt(h * h + radial * radial);
        } else {
            const double n1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
            const double n2 = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
            const double* t = n1 > n2 ? t1 : t2;
            const double nt = n1 > n2 ? n1 : n2;
            for (int d = 0; d < 3; ++d) c[d] = cen[d] + R * t[d] / nt;
            dist = R;
        }
        break;
    }
    case SURF_ARC: {   // Surface.cpp:507-579: the arc of radius P[0] about the origin in z = 0
        // between the angles P[1] < P[2]; points outside the arc's angles go to the nearer end
        // along the circle (split opposite the arc's midpoint), tagged 1 (first end) or 2
        const double R = P[0], a1 = P[1], a2 = P[2];
        const double half = 0.5 * (2 * M_PI - (a2 - a1));
        const double th = atan2(x[1], x[0]);
        double ang;
        if (th > a1 && th < a2) {
            ang = th;
        } else if (th <= a1) {
            bdy = th >= a1 - half ? 1 : 2;
            ang = bdy == 1 ? a1 : a2;
        } else {
            bdy = th < a2 + half ? 2 : 1;
            ang = bdy == 1 ? a1 : a2;
        }
        c[0] = R * cos(ang);
        c[1] = R * sin(ang);
        c[2] = 0.0;
        dist = sqrt(dsq(x[0] - c[0]) + dsq(x[1] - c[1]) + dsq(x[2] - c[2]));
        break;
    }
    case SURF_ELLIPSOID: {
        cp_ellipsoid(s.P, x, c);
        dist = sqrt((x[0] - c[0]) * (x[0] - c[0]) + (x[1] - c[1]) * (x[1] - c[1]) + (x[2] - c[2]) * (x[2] - c[2]));
        break;
    }
    case SURF_TRIANGULATION:   // Surface.cpp:1751-1779 (|signed distance|, boundary tag 0)
        cp_mesh(s.mesh, x, c, dist);
        break;
    case SURF_POLYLINE: {   // segment projections through the hierarchy: the exact minimiser, lowest index
        cp_polyline(s.mesh, x, c, dist);
        if (s.is_open) {   // Surface.cpp:2614-2636
            const double e0 = sqrt(dsq(s.ep0[0] - c[0]) + dsq(s.ep0[1] - c[1]) + dsq(s.ep0[2] - c[2]));
            const double e1 = sqrt(dsq(s.ep1[0] - c[0]) + dsq(s.ep1[1] - c[1]) + dsq(s.ep1[2] - c[2]));
            bdy = e0 < 1e-7 ? 1 : (e1 < 1e-7 ? 2 : 0);
        }
        break;
    }
    default:
        c[0] = c[1] = c[2] = NAN;
        dist = NAN;
    }
}

// ---------------------------------------------------------------- grid helpers
__device__ __forceinline__ void trip_of(int64_t l, int64_t N1d, int64_t* t) {
    t[2] = l % N1d;
    t[1] = (l / N1d) % N1d;
    t[0] = l / (N1d * N1d);
}

// ---------------------------------------------------------------- BFS kernels
__global__ void k_fill_i32(int32_t* a, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += stride) a[k] = v;
}

// FindStartingPoint (Tube.cpp:540-551) + first node (:328-331).
__global__ void k_seed(SurfDev s, GridDev g, int32_t* gidx, int64_t* rec, double* cpx, double* dist, int32_t* bdy,
                       int32_t* frontier, int* err) {
    double xs[3] = {g.xstart, g.xstart, g.xstart}, c[3], d;
    int b;
    cp_surface(s, xs, c, d, b);
    int64_t t[3];
    for (int k = 0; k < 3; ++k) t[k] = (int64_t)rint((c[k] - g.xstart) / g.dx);
    for (int k = 0; k < 3; ++k)
        if (t[k] < 0 || t[k] >= g.N1d) { *err = 1; return; }
    const int64_t l = (t[0] * g.N1d + t[1]) * g.N1d + t[2];
    double x[3];
    for (int k = 0; k < 3; ++k) x[k] = g.xstart + t[k] * g.dx;
    cp_surface(s, x, c, d, b);
    gidx[l] = 0;
    rec[0] = l;
    for (int k = 0; k < 3; ++k) cpx[k] = c[k];
    dist[0] = d;
    bdy[0] = b;
    frontier[0] = 0;
}

// 1. every (parent, direction) claims its neighbour; the first claimer records it.
__global__ void k_expand(const int32_t* frontier, int64_t nf, const int64_t* rec, GridDev g, int32_t* gidx,
                         int64_t* cand, int* ncand, int* err) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * 6) return;
    const int64_t fi = t / 6;
    const int dir = (int)(t - fi * 6);
    int64_t tr[3];
    trip_of(rec[frontier[fi]], g.N1d, tr);
    tr[dir >> 1] += (dir & 1) ? -1 : 1;
    if (tr[dir >> 1] < 0 || tr[dir >> 1] >= g.N1d) { atomicOr(err, 2); return; }   // grid padding too small
    const int64_t l = (tr[0] * g.N1d + tr[1]) * g.N1d + tr[2];
    const int32_t v = __hip_atomic_load(gidx + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v < kCand || v == kOutside) return;          // in the tube already, or visited outside
    const int32_t key = kCand + (int32_t)t;           // t = rank * 6 + dir
    const int32_t old = atomicMin(gidx + l, key);
    if (old == kUnvisited) {
        const int pos = atomicAdd(ncand, 1);
        cand[pos] = l;
    }
}

// 2. closest point of each candidate; inside ones publish their slot.
__global__ void k_evaluate(const int64_t* cand, const int* ncand, SurfDev s, GridDev g, int32_t* gidx,
                           int32_t* slot, double* ccp, double* cdist, int32_t* cbdy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *ncand) return;
    const int64_t l = cand[i];
    int64_t tr[3];
    trip_of(l, g.N1d, tr);
    double x[3], c[3], d;
    int b;
    for (int k = 0; k < 3; ++k) x[k] = g.xstart + tr[k] * g.dx;   // Grid.cpp:35
    cp_surface(s, x, c, d, b);
    if (fabs(d) <= g.r_t) {   // Tube.cpp:444
        const int32_t key = gidx[l] - kCand;
        slot[key] = (int32_t)i;
        for (int k = 0; k < 3; ++k) ccp[3 * i + k] = c[k];
        cdist[i] = d;
        cbdy[i] = b;
    } else {
        gidx[l] = kOutside;
    }
}

// 3a. occupied slots per block of 256
__global__ void k_count(const int32_t* slot, int64_t ns, int* bcount) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int occ = (i < ns && slot[i] >= 0) ? 1 : 0;
    const unsigned long long m = __ballot(occ);
    __shared__ int ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// 3b. exclusive scan of the block counts in one workgroup; total to *total.
__global__ void k_scan(int* bcount, int64_t nb, int* total) {
    __shared__ int part[256];
    const int64_t per = (nb + 255) / 256;
    const int64_t b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    int s = 0;
    for (int64_t b = b0; b < b1; ++b) s += bcount[b];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < 256; ++k) { int v = part[k]; part[k] = run; run += v; }
        *total = run;
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int64_t b = b0; b < b1; ++b) { int v = bcount[b]; bcount[b] = run; run += v; }
}

// 3c. assign tube indices in slot order, append node data and next frontier.
__global__ void k_emit(const int32_t* slot, int64_t ns, const int* boff, const int64_t* cand, const double* ccp,
                       const double* cdist, const int32_t* cbdy, int64_t N0, int32_t* gidx, int64_t* rec,
                       double* cpx, double* dist, int32_t* bdy, int32_t* next) {
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
    const int64_t t = N0 + before;
    const int64_t l = cand[ci];
    gidx[l] = (int32_t)t;
    rec[t] = l;
    for (int k = 0; k < 3; ++k) cpx[3 * t + k] = ccp[3 * ci + k];
    dist[t] = cdist[ci];
    bdy[t] = cbdy[ci];
    next[before] = (int32_t)t;
}

// SetNeighbours (Tube.cpp:676-694) and ComputeBasePoints (:464-483).
__global__ void k_nbr_bpt(int64_t N, const int64_t* rec, const double* cpx, const int32_t* gidx, GridDev g, int p,
                          int32_t* nbr, int32_t* bpt, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int64_t tr[3];
    trip_of(rec[i], g.N1d, tr);
    for (int dir = 0; dir < 6; ++dir) {
        int64_t u[3] = {tr[0], tr[1], tr[2]};
        u[dir >> 1] += (dir & 1) ? -1 : 1;
        int32_t v = -1;
        if (u[dir >> 1] >= 0 && u[dir >> 1] < g.N1d) {
            const int32_t gv = gidx[(u[0] * g.N1d + u[1]) * g.N1d + u[2]];
            if (gv >= 0 && gv < kCand) v = gv;
        }
        nbr[6 * i + dir] = v;
    }
    const int64_t shift = (p % 2 == 0) ? p / 2 : (p - 1) / 2;
    int64_t b[3];
    for (int k = 0; k < 3; ++k) {
        const double s = (cpx[3 * i + k] - g.xstart) / g.dx;
        b[k] = (int64_t)((p % 2 == 0) ? rint(s) : floor(s)) - shift;
        if (b[k] < 0 || b[k] >= g.N1d) { atomicOr(err, 4); bpt[i] = -1; return; }
    }
    const int32_t gv = gidx[(b[0] * g.N1d + b[1]) * g.N1d + b[2]];
    if (gv < 0 || gv >= kCand) { atomicOr(err, 4); bpt[i] = -1; return; }
    bpt[i] = gv;
}

// Barycentric Lagrange weights, Interpolation.cpp:128-200.
__device__ void weights_1d(const double* xs, double xq, int S, double* w) {
    double cw[5];
    switch (S) {
    case 1: cw[0] = 1; break;
    case 2: cw[0] = 1; cw[1] = -1; break;
    case 3: cw[0] = 1; cw[1] = -2; cw[2] = 1; break;
    case 4: cw[0] = 1; cw[1] = -3; cw[2] = 3; cw[3] = -1; break;
    default: cw[0] = 1; cw[1] = -4; cw[2] = 6; cw[3] = -4; cw[4] = 1; break;
    }
    int hit = -1;
    for (int j = 0; j < S; ++j)
        if (fabs(xq - xs[j]) < 1e-17) hit = j;
    if (hit >= 0) {
        for (int j = 0; j < S; ++j) w[j] = 0.0;
        w[hit] = 1.0;
        return;
    }
    double sum = 0.0;
    for (int j = 0; j < S; ++j) { w[j] = cw[j] / (xq - xs[j]); sum += w[j]; }
    for (int j = 0; j < S; ++j) w[j] /= sum;
}

// E rows at query points (Interpolation.cpp:52-121, 204-225): base point,
// 1-D weights per axis, value (w_z w_y) w_x, column = tube index of
// (b + (i, j, k)); x outer, z inner.
__device__ bool stencil_weights(const GridDev& g, int p, const double* xq, int64_t* b, double (*w)[5]) {
    const int S = p + 1;
    const int64_t shift = (p % 2 == 0) ? p / 2 : (p - 1) / 2;
    for (int d = 0; d < 3; ++d) {
        const double s = (xq[d] - g.xstart) / g.dx;
        b[d] = (int64_t)((p % 2 == 0) ? rint(s) : floor(s)) - shift;
        if (!(b[d] >= 0 && b[d] + S <= g.N1d)) return false;
        double xs[5];
        for (int i = 0; i < S; ++i) xs[i] = g.xstart + (b[d] + i) * g.dx;
        weights_1d(xs, xq[d], S, w[d]);
    }
    return true;
}

__global__ void k_erows(int64_t nq, const double* xq, const int32_t* gidx, GridDev g, int p, int* ecol, double* eval,
                        float* eval32, int64_t ld, int* err) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int S = p + 1;
    int64_t b[3];
    double w[3][5];
    if (!stencil_weights(g, p, xq + 3 * q, b, w)) { atomicOr(err, 8); return; }
    int s = 0;
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            for (int k = 0; k < S; ++k, ++s) {
                const int32_t gv = gidx[((b[0] + i) * g.N1d + (b[1] + j)) * g.N1d + (b[2] + k)];
                int col = 0;
                if (gv >= 0 && gv < kCand) col = gv;
                else atomicOr(err, 8);   // stencil leaves the tube
                const double v = w[2][k] * w[1][j] * w[0][i];
                ecol[s * ld + q] = col;
                eval[s * ld + q] = v;
                if (eval32) eval32[s * ld + q] = (float)v;
            }
}

// L rows (FDMatrices.cpp:25-49): diagonal -2 dim / h^2, neighbours 1 / h^2.
__global__ void k_lrows(int64_t N, const int32_t* nbr, double h, int* lcol, double* lval, float* lval32,
                        double* ldiag, float* ldiag32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double hh = h * h;
    const double dg = -6.0 / hh, off = 1.0 / hh;
    ldiag[i] = dg;
    ldiag32[i] = (float)dg;
    for (int s = 0; s < 6; ++s) {
        const int32_t v = nbr[6 * i + s];
        lcol[s * N + i] = v >= 0 ? v : (int)i;
        lval[s * N + i] = v >= 0 ? off : 0.0;
        lval32[s * N + i] = v >= 0 ? (float)off : 0.0f;
    }
}

// ---------------------------------------------------------------- host side
SurfaceHost surface_from_desc(const char* type, const double* params, int nparams, double bb0, double bb1, int dim,
                              const double* verts, int64_t nverts) {
    SurfaceHost s;
    s.type = type ? type : "";
    s.dim = dim;
    if (dim != 3) throw ArgError("the GPU tube supports 3-D embeddings");
    if (s.type == "Polyline") {
        if (!verts || nverts < 2) throw ArgError("polyline needs >= 2 vertices");
        s.verts.assign(verts, verts + 3 * nverts);
        const bool closed = std::memcmp(verts, verts + 3 * (nverts - 1), 3 * sizeof(double)) == 0;
        s.is_open = !closed;   // SurfaceSpecifier.cpp:334-341 (faces are consecutive vertex pairs)
        s.bbox[0] = bb0;
        s.bbox[1] = bb1;
        return s;
    }
    // defaults of SurfaceSpecifier::SetSurface(string) (SurfaceSpecifier.cpp:65-242)
    if (nparams <= 0 || !params) {
        if (s.type == "Sphere" || s.type == "Hemisphere") { s.params = {1.0}; bb0 = -1; bb1 = 1; }
        else if (s.type == "Torus") { s.params = {2.0, 1.0}; bb0 = -3; bb1 = 3; }
        else if (s.type == "Circle3D") { s.params = {1, 0, 0, 0, 0, 0, 1}; bb0 = -1; bb1 = 1; }
        else if (s.type == "Ellipsoid") { s.params = {1.5, 1.0, 0.75}; bb0 = -1.5; bb1 = 1.5; }
        else if (s.type == "Arc") { s.params = {1.0, -3 * M_PI_4, M_PI_4}; bb0 = -1; bb1 = 1; }
        else throw ArgError("no default parameters for surface '" + s.type + "'");
    } else {
        s.params.assign(params, params + nparams);
    }
    s.is_open = s.type == "Hemisphere" || s.type == "Arc";   // SetIsOpen (SurfaceSpecifier.cpp:370-381)
    s.bbox[0] = bb0;
    s.bbox[1] = bb1;
    return s;
}

SurfaceHost surface_mesh(const double* verts, int64_t nverts, const int32_t* faces, int64_t nfaces, double bb0,
                         double bb1) {
    if (!verts || !faces || nverts < 3 || nfaces < 1) throw ArgError("mesh needs vertices and triangles");
    SurfaceHost s;
    s.type = "Triangulation";
    s.dim = 3;
    s.verts.assign(verts, verts + 3 * nverts);
    s.faces.assign(faces, faces + 3 * nfaces);
    s.is_open = false;   // SurfaceSpecifier.cpp:319
    if (!(bb0 < bb1)) {  // SurfaceSpecifier::SetSurface() (:346-363)
        bb0 = 0.0;
        bb1 = 0.0;
        for (double x : s.verts) {
            bb0 = std::min(bb0, x);
            bb1 = std::max(bb1, x);
        }
    }
    s.bbox[0] = bb0;
    s.bbox[1] = bb1;
    return s;
}

SurfDev make_surfdev(const SurfaceHost& s, const double* dverts, const MeshDev* mesh) {
    SurfDev d{};
    d.dim = 3;
    for (size_t i = 0; i < s.params.size() && i < 16; ++i) d.P[i] = s.params[i];
    if (s.type == "Sphere") {
        d.type = SURF_SPHERE;
        if (s.params.size() <= 1) d.P[1] = d.P[2] = d.P[3] = 0.0;
    } else if (s.type == "Hemisphere") d.type = SURF_HEMISPHERE;
    else if (s.type == "Torus") d.type = SURF_TORUS;
    else if (s.type == "Circle3D") d.type = SURF_CIRCLE3D;
    else if (s.type == "Point") d.type = SURF_POINT;
    else if (s.type == "Arc") {   // Surface::cpArc's argument checks (Surface.cpp:513-527)
        if (s.params.size() < 3) throw ArgError("Arc needs radius, angle1, angle2");
        if (s.params[1] < -M_PI || s.params[2] > M_PI) throw ArgError("Arc angles must lie in (-pi, pi]");
        if (!(s.params[1] < s.params[2])) throw ArgError("Arc: angle1 must be less than angle2");
        d.type = SURF_ARC;
        d.is_open = 1;
    }
    else if (s.type == "Ellipsoid") {
        if (s.params.size() < 3 || !(s.params[0] > 0 && s.params[1] > 0 && s.params[2] > 0))
            throw ArgError("Ellipsoid needs three positive semi-axes");
        d.type = SURF_ELLIPSOID;
    }
    else if (s.type == "Triangulation") {
        if (!mesh) throw ArgError("triangulation without its hierarchy");
        d.type = SURF_TRIANGULATION;
        d.mesh = *mesh;
    } else if (s.type == "Polyline") {
        if (!mesh) throw ArgError("polyline without its hierarchy");
        d.type = SURF_POLYLINE;
        d.mesh = *mesh;
        d.verts = dverts;
        d.nseg = (int)(s.verts.size() / 3) - 1;
        d.is_open = s.is_open;
        for (int k = 0; k < 3; ++k) {
            d.ep0[k] = s.verts[k];
            d.ep1[k] = s.verts[s.verts.size() - 3 + k];
        }
    } else throw ArgError("surface '" + s.type + "' is not supported by the GPU closest-point kernels");
    return d;
}

Tube* tube_build(Ctx* ctx, const SurfaceHost& s, int p, double dx) {
    if (p < 1 || p > 4) throw ArgError("interpolation degree must be 1..4");
    if (!(dx > 0)) throw ArgError("dx must be positive");
    auto t0 = std::chrono::steady_clock::now();
    const hipStream_t st = ctx->stream;
    std::unique_ptr<Tube> T(new Tube());
    T->ctx = ctx;
    T->surf = s;
    T->dim = 3;
    T->p = p;
    T->dx = dx;
    if (s.type == "Triangulation" || s.type == "Polyline") {
        if (s.type == "Triangulation") mesh_build(T->mesh, s.verts, s.faces, ctx->stream);
        else segment_build(T->mesh, s.verts);
        const MeshDev md = T->mesh.dev();
        T->sd = make_surfdev(s, T->mesh.v.as<double>(), &md);
    } else {
        if (!s.verts.empty()) {
            T->surf_verts.~Buffer();
            new (&T->surf_verts) Buffer(s.verts.size() * sizeof(double));
            HIPCHK(hipMemcpy(T->surf_verts.p, s.verts.data(), s.verts.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        T->sd = make_surfdev(s, T->surf_verts.as<double>());
    }
    // Tube.cpp:131-157 (m_order = 2)
    const int order = 2;
    T->rfac = 1.0001 * std::sqrt((T->dim - 1) * std::pow(0.5 * (p + 1), 2) + std::pow(0.5 * (order + p + 1), 2));
    T->r_t = T->rfac * dx;
    const size_t pad = 5;
    const size_t spacings = (size_t)std::ceil((s.bbox[1] - s.bbox[0]) / dx);
    const size_t n = (size_t)std::ceil(std::log2((double)(spacings + 2 * pad)));
    const size_t sdiff = (size_t)std::ceil(0.5 * (std::pow(2, (double)n) - spacings - 2 * pad));
    T->xstart = s.bbox[0] - (pad + sdiff) * dx;
    T->n = (int64_t)n;
    T->N1d = (int64_t)std::pow(2.0, (double)n) + 1;
    if (T->n > 10) throw ArgError("grid 2^n+1 with n > 10 exceeds the dense index grid");
    const int64_t G = T->N1d * T->N1d * T->N1d;
    T->gidx.~Buffer();
    new (&T->gidx) Buffer((size_t)G * sizeof(int32_t));
    hipLaunchKernelGGL(k_fill_i32, dim3(4096), dim3(256), 0, st, T->gidx.as<int32_t>(), G, kUnvisited);

    GridDev g{T->N1d, T->xstart, dx, T->r_t};
    // growable node arrays
    int64_t cap = 1 << 16;
    std::unique_ptr<Buffer> rec(new Buffer(cap * 8)), cpx(new Buffer(cap * 24)), dist(new Buffer(cap * 8)),
        bdy(new Buffer(cap * 4));
    auto grow = [&](int64_t need, int64_t have) {
        if (need <= cap) return;
        int64_t nc = cap;
        while (nc < need) nc *= 2;
        auto mv = [&](std::unique_ptr<Buffer>& b, size_t esz) {
            std::unique_ptr<Buffer> nb(new Buffer(nc * esz));
            if (have) HIPCHK(hipMemcpyAsync(nb->p, b->p, have * esz, hipMemcpyDeviceToDevice, st));
            b.swap(nb);
            HIPCHK(hipStreamSynchronize(st));
        };
        mv(rec, 8);
        mv(cpx, 24);
        mv(dist, 8);
        mv(bdy, 4);
        cap = nc;
    };
    int* dint = reinterpret_cast<int*>(ctx->workspace("tube_int", 64));
    int* err = dint;          // [0]
    int* ncand = dint + 1;    // [1]
    int* total = dint + 2;    // [2]
    HIPCHK(hipMemsetAsync(dint, 0, 64, st));
    int64_t fcap = 1 << 14;
    std::unique_ptr<Buffer> fr(new Buffer(fcap * 4)), fr2(new Buffer(fcap * 4));
    hipLaunchKernelGGL(k_seed, dim3(1), dim3(1), 0, st, T->sd, g, T->gidx.as<int32_t>(), rec->as<int64_t>(),
                       cpx->as<double>(), dist->as<double>(), bdy->as<int32_t>(), fr->as<int32_t>(), err);
    int64_t N = 1, nf = 1, levels = 0;
    int* hbuf = reinterpret_cast<int*>(ctx->host_states);   // pinned scratch
    while (nf > 0) {
        HIPCHK(hipStreamSynchronize(st));   // the previous level's emit still reads the level workspaces
        const int64_t ns = nf * 6;
        auto* cand = reinterpret_cast<int64_t*>(ctx->workspace("tube_cand", ns * 8));
        auto* slot = reinterpret_cast<int32_t*>(ctx->workspace("tube_slot", ns * 4));
        auto* ccp = ctx->workspace("tube_ccp", ns * 24);
        auto* cdist = ctx->workspace("tube_cdist", ns * 8);
        auto* cbdy = reinterpret_cast<int32_t*>(ctx->workspace("tube_cbdy", ns * 4));
        const int64_t nb = (ns + 255) / 256;
        int* bcount = reinterpret_cast<int*>(ctx->workspace("tube_bcount", nb * 4));
        HIPCHK(hipMemsetAsync(ncand, 0, 4, st));
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<int64_t>(nb, 4096)), dim3(256), 0, st, slot, ns, -1);
        hipLaunchKernelGGL(k_expand, dim3((unsigned)nb), dim3(256), 0, st, fr->as<int32_t>(), nf, rec->as<int64_t>(), g,
                           T->gidx.as<int32_t>(), cand, ncand, err);
        hipLaunchKernelGGL(k_evaluate, dim3((unsigned)nb), dim3(256), 0, st, cand, ncand, T->sd, g, T->gidx.as<int32_t>(),
                           slot, ccp, cdist, cbdy);
        hipLaunchKernelGGL(k_count, dim3((unsigned)nb), dim3(256), 0, st, slot, ns, bcount);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, bcount, nb, total);
        HIPCHK(hipMemcpyAsync(hbuf, dint, 16, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hbuf[0]) break;
        const int64_t cnt = hbuf[2];
        if (N + cnt >= kCand) throw ArgError("tube exceeds 2^30 nodes");
        grow(N + cnt, N);
        if (cnt > fcap) {
            while (fcap < cnt) fcap *= 2;
            fr2.reset(new Buffer(fcap * 4));
            std::unique_ptr<Buffer> keep(new Buffer(fcap * 4));
            HIPCHK(hipMemcpyAsync(keep->p, fr->p, nf * 4, hipMemcpyDeviceToDevice, st));
            fr.swap(keep);
            HIPCHK(hipStreamSynchronize(st));
        }
        hipLaunchKernelGGL(k_emit, dim3((unsigned)nb), dim3(256), 0, st, slot, ns, bcount, cand, ccp, cdist, cbdy, N,
                           T->gidx.as<int32_t>(), rec->as<int64_t>(), cpx->as<double>(), dist->as<double>(),
                           bdy->as<int32_t>(), fr2->as<int32_t>());
        fr.swap(fr2);
        N += cnt;
        nf = cnt;
        ++levels;
    }
    HIPCHK(hipMemcpyAsync(hbuf, dint, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hbuf[0] & 1) throw GeomError("starting point outside the grid");
    if (hbuf[0] & 2) throw GeomError("tube reaches the edge of the 2^n+1 grid (padding too small)");
    T->N = N;
    T->levels = levels;
    // final node arrays sized N
    auto fit = [&](Buffer& dst, const std::unique_ptr<Buffer>& src, size_t esz) {
        dst.~Buffer();
        new (&dst) Buffer(N * esz);
        HIPCHK(hipMemcpyAsync(dst.p, src->p, N * esz, hipMemcpyDeviceToDevice, st));
    };
    fit(T->rec, rec, 8);
    fit(T->cpx, cpx, 24);
    fit(T->dist, dist, 8);
    fit(T->bdy, bdy, 4);
    T->nbr.~Buffer();
    new (&T->nbr) Buffer(N * 6 * 4);
    T->bpt.~Buffer();
    new (&T->bpt) Buffer(N * 4);
    hipLaunchKernelGGL(k_nbr_bpt, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, T->rec.as<int64_t>(),
                       T->cpx.as<double>(), T->gidx.as<int32_t>(), g, p, T->nbr.as<int32_t>(), T->bpt.as<int32_t>(), err);
    HIPCHK(hipMemcpyAsync(hbuf, dint, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hbuf[0] & 4)
        throw GeomError("Tube::ComputeBasePoints - Interpolation base point is not within the tube");
    T->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return T.release();
}

void build_E_ell(const Tube& t, const double* d_xq, int64_t nq, int* ecol, double* eval, float* eval32, int64_t ld,
                 hipStream_t st, Ctx* wctx) {
    int* err = reinterpret_cast<int*>((wctx ? wctx : t.ctx)->workspace("tube_int", 64));
    HIPCHK(hipMemsetAsync(err, 0, 4, st));
    GridDev g{t.N1d, t.xstart, t.dx, t.r_t};
    hipLaunchKernelGGL(k_erows, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, nq, d_xq, t.gidx.as<int32_t>(),
                       g, t.p, ecol, eval, eval32, ld, err);
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h) throw GeomError("interpolation stencil leaves the tube (base point not within the tube)");
}

bool tp_disabled() {
    const char* e = std::getenv("CPMGPU_OP_FORMAT");
    return e && std::string(e) == "ell";
}

Op* op_from_tube(Tube* t, double m, double n, int precond, int precond_iters) {
    const hipStream_t st = t->ctx->stream;
    const int64_t N = t->N;
    const int S = t->p + 1;
    std::unique_ptr<Op> op(new Op(t->ctx, N, 0));
    op->ew = S * S * S;
    op->lw = 6;
    auto alloc = [](Buffer& b, size_t bytes) { b.~Buffer(); new (&b) Buffer(bytes); };
    alloc(op->ecol, (size_t)op->ew * N * 4);
    alloc(op->eval64, (size_t)op->ew * N * 8);
    alloc(op->eval32, (size_t)op->ew * N * 4);
    alloc(op->lcol, (size_t)6 * N * 4);
    alloc(op->lval64, (size_t)6 * N * 8);
    alloc(op->lval32, (size_t)6 * N * 4);
    alloc(op->ldiag64, (size_t)N * 8);
    alloc(op->ldiag32, (size_t)N * 4);
    alloc(op->flag, (size_t)N);
    HIPCHK(hipMemsetAsync(op->flag.p, 0, N, st));
    build_E_ell(*t, t->cpx.as<double>(), N, op->ecol.as<int>(), op->eval64.as<double>(), op->eval32.as<float>(), N, st);
    hipLaunchKernelGGL(k_lrows, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, t->nbr.as<int32_t>(), t->dx,
                       op->lcol.as<int>(), op->lval64.as<double>(), op->lval32.as<float>(), op->ldiag64.as<double>(),
                       op->ldiag32.as<float>());
    HIPCHK(hipStreamSynchronize(st));
    op->m = m;
    op->n = n;
    op->precond = precond;
    op->precond_iters = precond_iters;
    TpSource src;
    src.D = src.N = N;
    src.ecol = op->ecol.as<int>();
    src.eval = op->eval64.as<double>();
    src.ew = op->ew;
    src.lcol = op->lcol.as<int>();
    src.lval = op->lval64.as<double>();
    src.lw = op->lw;
    src.ldiag = op->ldiag64.as<double>();
    src.flag = op->flag.as<uint8_t>();
    src.qpts = t->cpx.as<double>();
    if (!tp_disabled()) op_build_tp(*op, *t, src);   // ELL stays when a row does not fit
    return op.release();
}

}  // namespace cpmgpu
