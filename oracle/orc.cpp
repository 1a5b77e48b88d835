// CPU oracle: surfaces, tube, E, L, operator, BiCGStab, Poisson problem.
// TEST INFRASTRUCTURE ONLY (see orc.h).
#include "orc.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <list>
#include <numeric>

namespace orc {

// ---------------------------------------------------------------- CSR
CSR CSR::from_triplets(i64 rows, i64 cols, std::vector<Trip> t) {
    // Eigen setFromTriplets: duplicates summed in insertion order, columns
    // ascending in a RowMajor matrix.
    std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    CSR A;
    A.rows = rows;
    A.cols = cols;
    A.ptr.assign((size_t)rows + 1, 0);
    for (size_t k = 0; k < t.size();) {
        size_t e = k;
        double v = t[k].v;
        while (e + 1 < t.size() && t[e + 1].r == t[k].r && t[e + 1].c == t[k].c) {
            ++e;
            v += t[e].v;
        }
        if (t[k].r < 0 || t[k].r >= rows || t[k].c < 0 || t[k].c >= cols)
            throw std::runtime_error("CSR::from_triplets: index out of range");
        A.col.push_back(t[k].c);
        A.val.push_back(v);
        A.ptr[(size_t)t[k].r + 1]++;
        k = e + 1;
    }
    for (i64 r = 0; r < rows; ++r) A.ptr[(size_t)r + 1] += A.ptr[(size_t)r];
    return A;
}

double CSR::coeff(i64 r, i64 c) const {
    auto b = col.begin() + ptr[(size_t)r], e = col.begin() + ptr[(size_t)r + 1];
    auto it = std::lower_bound(b, e, c);
    if (it != e && *it == c) return val[(size_t)(it - col.begin())];
    return 0.0;
}

void CSR::matvec(const double* x, double* y) const {
    for (i64 r = 0; r < rows; ++r) {
        double s = 0.0;
        for (i64 k = ptr[(size_t)r]; k < ptr[(size_t)r + 1]; ++k) s += val[(size_t)k] * x[col[(size_t)k]];
        y[r] = s;
    }
}

void CSR::prune_zeros() {
    std::vector<i64> np{0};
    std::vector<i64> nc;
    Vec nv;
    for (i64 r = 0; r < rows; ++r) {
        for (i64 k = ptr[(size_t)r]; k < ptr[(size_t)r + 1]; ++k)
            if (val[(size_t)k] != 0.0) {
                nc.push_back(col[(size_t)k]);
                nv.push_back(val[(size_t)k]);
            }
        np.push_back((i64)nc.size());
    }
    ptr.swap(np);
    col.swap(nc);
    val.swap(nv);
}

// ---------------------------------------------------------------- helpers
static inline double sq(double a) { return a * a; }   // pow(a,2) folds to a*a in the reference build

static double distance(const double* a, const double* b, int d) {   // VectorMath.cpp:77-89
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += sq(a[i] - b[i]);
    return std::sqrt(s);
}
static double dot(const double* a, const double* b, int d) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += a[i] * b[i];
    return s;
}
static double norm(const double* a, int d) { return std::sqrt(dot(a, a, d)); }
static void cross(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---------------------------------------------------------------- SurfaceSpec
SurfaceSpec SurfaceSpec::defaults(const std::string& type) {
    SurfaceSpec s;
    s.type = type;
    if (type == "Circle") {
        s.params = {1.0}; s.bbox = {-1.0, 1.0}; s.dim = 2;
    } else if (type == "Circle3D") {
        s.params = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0}; s.bbox = {-1.0, 1.0}; s.dim = 3;
    } else if (type == "Sphere" || type == "Hemisphere") {
        s.params = {1.0}; s.bbox = {-1.0, 1.0}; s.dim = 3;
    } else if (type == "Torus") {
        s.params = {2.0, 1.0}; s.bbox = {-3.0, 3.0}; s.dim = 3;
    } else if (type == "Ellipsoid") {
        s.params = {1.5, 1.0, 0.75}; s.bbox = {-1.5, 1.5}; s.dim = 3;
    } else if (type == "Arc") {                    // :104-111: radius, angle1, angle2
        s.params = {1.0, -3 * M_PI_4, M_PI_4}; s.bbox = {-1.0, 1.0}; s.dim = 2;
    } else {
        throw std::runtime_error("No default parameters for surface '" + type + "'");
    }
    // SetIsOpen (:370-381): of these, Hemisphere and Arc are open.
    s.is_open = (type == "Hemisphere" || type == "Arc");
    return s;
}

SurfaceSpec SurfaceSpec::polyline(const std::vector<Vec>& v, const Vec& bbox) {
    SurfaceSpec s;
    s.type = "Polyline";
    s.dim = (int)v[0].size();
    s.verts = v;
    s.bbox = bbox;
    for (size_t i = 0; i + 1 < v.size(); ++i) s.faces.push_back({(i64)i, (i64)i + 1});
    // SurfaceSpecifier.cpp:334-341: closed when first == last vertex or the
    // last face touches vertex 0.
    const auto& lf = s.faces.back();
    s.is_open = !(v.front() == v.back() || lf[0] == 0 || lf[1] == 0);
    return s;
}

// ---------------------------------------------------------------- Surface
Surface::Surface(const SurfaceSpec& s) : s_(s) { parameterise(); }

void Surface::parameterise() {
    const auto& P = s_.params;
    const std::string& ty = s_.type;
    if (ty == "Polyline" || ty == "Triangulation") {
        xp = s_.verts;
        return;
    }
    if (ty == "Circle") {                         // Surface.cpp:1868-1880, Np = 100
        size_t Np = 100;
        double R = P[0], dth = 2 * M_PI / (Np - 1);
        xp.assign(Np, Vec(2));
        thetap.assign(Np, 0);
        for (size_t i = 0; i < Np; ++i) {
            thetap[i] = i * dth;
            xp[i][0] = R * std::cos(thetap[i]);
            xp[i][1] = R * std::sin(thetap[i]);
        }
    } else if (ty == "Sphere" || ty == "Hemisphere") {   // :1973-2029
        size_t Np = ty == "Sphere" ? 40000 : 10000;
        size_t sq_n = (size_t)std::sqrt((double)Np);
        double R = P[0];
        double dth = 2 * M_PI / (sq_n - 1);
        double dph = (ty == "Sphere" ? M_PI : M_PI_2) / (sq_n - 1);
        xp.assign(Np, Vec(3));
        thetap.assign(Np, 0);
        phip.assign(Np, 0);
        for (size_t i = 0; i < sq_n; ++i)
            for (size_t j = 0; j < sq_n; ++j) {
                thetap[i * sq_n + j] = i * dth;
                phip[i * sq_n + j] = j * dph;
            }
        for (size_t i = 0; i < Np; ++i) {
            xp[i][0] = R * std::sin(phip[i]) * std::cos(thetap[i]);
            xp[i][1] = R * std::sin(phip[i]) * std::sin(thetap[i]);
            xp[i][2] = R * std::cos(phip[i]);
        }
    } else if (ty == "Torus") {                   // :2063-2093, Np = 90000
        size_t Np = 90000, sq_n = 300;
        double R = P[0], r = P[1], du = 2 * M_PI / (sq_n - 1);
        xp.assign(Np, Vec(3));
        thetap.assign(Np, 0);
        phip.assign(Np, 0);
        for (size_t i = 0; i < sq_n; ++i)
            for (size_t j = 0; j < sq_n; ++j) {
                phip[i * sq_n + j] = i * du;
                thetap[i * sq_n + j] = j * du;
            }
        for (size_t i = 0; i < Np; ++i) {
            xp[i][0] = (R + r * std::cos(thetap[i])) * std::cos(phip[i]);
            xp[i][1] = (R + r * std::cos(thetap[i])) * std::sin(phip[i]);
            xp[i][2] = r * std::sin(thetap[i]);
        }
    } else if (ty == "Ellipsoid") {               // :341-349, 2480-2511, Np = 10000
        size_t Np = 10000, sq_n = 100;
        double a = P[0], b = P[1], c = P[2], du = 2 * M_PI / (sq_n - 1), dv = M_PI / (sq_n - 1);
        xp.assign(Np, Vec(3));
        thetap.assign(Np, 0);
        phip.assign(Np, 0);
        for (size_t i = 0; i < sq_n; ++i)
            for (size_t j = 0; j < sq_n; ++j) {
                phip[i * sq_n + j] = i * du;
                thetap[i * sq_n + j] = j * dv;
            }
        for (size_t i = 0; i < Np; ++i) {
            xp[i][0] = a * std::sin(thetap[i]) * std::cos(phip[i]);
            xp[i][1] = b * std::sin(thetap[i]) * std::sin(phip[i]);
            xp[i][2] = c * std::cos(thetap[i]);
        }
    } else if (ty == "Arc") {                     // paramArc :1907-1924, Np = 100 (z = 0 in 3-D)
        size_t Np = 100;
        double R = P[0], a1 = P[1], a2 = P[2], dth = (a2 - a1) / (Np - 1);
        xp.assign(Np, Vec((size_t)s_.dim, 0.0));
        thetap.assign(Np, 0);
        for (size_t i = 0; i < Np; ++i) {
            thetap[i] = a1 + i * dth;
            xp[i][0] = R * std::cos(thetap[i]);
            xp[i][1] = R * std::sin(thetap[i]);
        }
    } else if (ty == "Point") {
        xp.assign(1, Vec(P.begin(), P.begin() + s_.dim));
    } else if (ty == "Circle3D") {                // :1883-1904, Np = 100
        size_t Np = 100;
        double R = P[0];
        double c[3] = {P[1], P[2], P[3]}, nrm[3] = {P[4], P[5], P[6]}, t1[3], t2[3];
        if (std::fabs(nrm[0]) > std::fabs(nrm[1])) { t1[0] = -nrm[2]; t1[1] = 0; t1[2] = nrm[0]; }
        else { t1[0] = 0; t1[1] = nrm[2]; t1[2] = -nrm[1]; }
        cross(nrm, t1, t2);
        double n1 = norm(t1, 3), n2 = norm(t2, 3);
        for (int d = 0; d < 3; ++d) { t1[d] /= n1; t2[d] /= n2; }
        double dth = 2 * M_PI / (Np - 1);
        xp.assign(Np, Vec(3));
        thetap.assign(Np, 0);
        for (size_t i = 0; i < Np; ++i) {
            thetap[i] = i * dth;
            double a = std::cos(thetap[i]), b = std::sin(thetap[i]);
            for (int d = 0; d < 3; ++d) xp[i][d] = c[d] + R * (a * t1[d] + b * t2[d]);
        }
    }
    // other surfaces: no parameterisation needed by the oracle
}

// ---------------------------------------------------------------- ellipsoid closest point
// Eberly, "Distance from a Point to an Ellipse, an Ellipsoid, or a
// Hyperellipsoid" (Geometric Tools, 2013): reduce to the first octant with
// semi-axes e0 >= e1 >= e2, then find the root s of
//   sum_i (r_i z_i / (s + r_i))^2 = 1,  r_i = (e_i / e_last)^2, z_i = y_i / e_i,
// by bisection until the interval stops shrinking; x_i = r_i y_i / (s + r_i).
static double robust_length(const double* v, int n) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = std::max(m, std::fabs(v[i]));
    if (m == 0.0) return 0.0;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (v[i] / m) * (v[i] / m);
    return m * std::sqrt(s);
}

static double ellipsoid_root(int n, const double* r, const double* z, double g) {
    // r[n-1] == 1
    double num[3];
    for (int i = 0; i < n - 1; ++i) num[i] = r[i] * z[i];
    num[n - 1] = z[n - 1];
    double s0 = z[n - 1] - 1.0, s1 = g < 0.0 ? 0.0 : robust_length(num, n) - 1.0, s = 0.0;
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

static void ellipse_point(double e0, double e1, double y0, double y1, double& x0, double& x1) {
    if (y1 > 0.0) {
        if (y0 > 0.0) {
            const double z[2] = {y0 / e0, y1 / e1};
            const double g = z[0] * z[0] + z[1] * z[1] - 1.0;
            if (g != 0.0) {
                const double r[2] = {(e0 / e1) * (e0 / e1), 1.0};
                const double s = ellipsoid_root(2, r, z, g);
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
            x1 = e1 * std::sqrt(1.0 - xd0 * xd0);
        } else {
            x0 = e0;
            x1 = 0.0;
        }
    }
}

static void ellipsoid_point(const double* e, const double* y, double* x) {
    if (y[2] > 0.0) {
        if (y[1] > 0.0) {
            if (y[0] > 0.0) {
                const double z[3] = {y[0] / e[0], y[1] / e[1], y[2] / e[2]};
                const double g = z[0] * z[0] + z[1] * z[1] + z[2] * z[2] - 1.0;
                if (g != 0.0) {
                    const double r[3] = {(e[0] / e[2]) * (e[0] / e[2]), (e[1] / e[2]) * (e[1] / e[2]), 1.0};
                    const double s = ellipsoid_root(3, r, z, g);
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
                ellipse_point(e[1], e[2], y[1], y[2], x[1], x[2]);
            }
        } else if (y[0] > 0.0) {
            x[1] = 0.0;
            ellipse_point(e[0], e[2], y[0], y[2], x[0], x[2]);
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
                x[2] = e[2] * std::sqrt(disc);
                done = true;
            }
        }
        if (!done) {
            x[2] = 0.0;
            ellipse_point(e[0], e[1], y[0], y[1], x[0], x[1]);
        }
    }
}

// Ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1 centred at the origin.
void cp_ellipsoid(const double* abc, const double* x, double* cp) {
    int perm[3] = {0, 1, 2};   // axes by decreasing semi-axis (stable)
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (abc[perm[j]] > abc[perm[i]]) std::swap(perm[i], perm[j]);
    double e[3], y[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e[k] = abc[perm[k]];
        y[k] = std::fabs(x[perm[k]]);
    }
    ellipsoid_point(e, y, q);
    for (int k = 0; k < 3; ++k) cp[perm[k]] = std::copysign(q[k], x[perm[k]]);
}

void Surface::cp(const double* x, double* cpx, double& dist, i64& bdy) const {
    const auto& P = s_.params;
    const std::string& ty = s_.type;
    bdy = 0;
    if (ty == "Sphere") {                          // Surface.cpp:710-740
        double R = P[0];
        double cen[3] = {0, 0, 0};
        if (P.size() > 1) { cen[0] = P[1]; cen[1] = P[2]; cen[2] = P[3]; }
        double xs[3] = {x[0] - cen[0], x[1] - cen[1], x[2] - cen[2]};
        double theta = std::atan2(xs[1], xs[0]);
        dist = std::sqrt(sq(xs[0]) + sq(xs[1]) + sq(xs[2]));
        double phi = dist != 0 ? std::acos(xs[2] / dist) : 0.0;
        cpx[0] = R * std::sin(phi) * std::cos(theta);
        cpx[1] = R * std::sin(phi) * std::sin(theta);
        cpx[2] = R * std::cos(phi);
        for (int d = 0; d < 3; ++d) cpx[d] += cen[d];
        dist -= R;
    } else if (ty == "Hemisphere") {               // :744-781
        double R = P[0];
        double theta = std::atan2(x[1], x[0]);
        if (x[2] > 0) {
            dist = std::sqrt(sq(x[0]) + sq(x[1]) + sq(x[2]));
            double phi = dist != 0 ? std::acos(x[2] / dist) : 0.0;
            cpx[0] = R * std::sin(phi) * std::cos(theta);
            cpx[1] = R * std::sin(phi) * std::sin(theta);
            cpx[2] = R * std::cos(phi);
            dist -= R;
            bdy = 0;
        } else {
            cpx[0] = R * std::cos(theta);
            cpx[1] = R * std::sin(theta);
            cpx[2] = 0;
            dist = distance(x, cpx, 3);
            bdy = 1;
        }
    } else if (ty == "Torus") {                    // :800-837
        double theta = std::atan2(x[1], x[0]);
        double r = std::sqrt(x[0] * x[0] + x[1] * x[1]);
        double cyl[2] = {r, x[2]};
        double cen[2] = {P[0], 0.0};
        double sh[2] = {cyl[0] - cen[0], cyl[1] - cen[1]};
        double th2 = std::atan2(sh[1], sh[0]);
        double cpc[2] = {P[1] * std::cos(th2) + cen[0], P[1] * std::sin(th2) + cen[1]};
        dist = distance(cyl, cpc, 2);
        cpx[0] = cpc[0] * std::cos(theta);
        cpx[1] = cpc[0] * std::sin(theta);
        cpx[2] = cpc[1];
    } else if (ty == "Circle") {                   // :421-440
        double R = P[0];
        double theta = std::atan2(x[1], x[0]);
        cpx[0] = R * std::cos(theta);
        cpx[1] = R * std::sin(theta);
        dist = 0.0;
        for (int d = 0; d < 2; ++d) dist += sq(x[d] - cpx[d]);
        dist = std::sqrt(dist);
    } else if (ty == "Arc") {                      // cpArc :507-579
        cp_arc(x, cpx, dist, bdy);
    } else if (ty == "Point") {                    // :840-848
        for (int d = 0; d < s_.dim; ++d) cpx[d] = P[d];
        dist = distance(x, cpx, s_.dim);
    } else if (ty == "Circle3D") {                 // :464-503
        double R = P[0];
        double c[3] = {P[1], P[2], P[3]}, nrm[3] = {P[4], P[5], P[6]}, t1[3], t2[3];
        if (std::fabs(nrm[0]) > std::fabs(nrm[1])) { t1[0] = -nrm[2]; t1[1] = 0; t1[2] = nrm[0]; }
        else { t1[0] = 0; t1[1] = nrm[2]; t1[2] = -nrm[1]; }
        cross(nrm, t1, t2);
        double cx[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
        double a = dot(nrm, nrm, 3) * dot(t1, cx, 3), b = dot(t2, cx, 3);
        double q[3];
        for (int d = 0; d < 3; ++d) q[d] = a * t1[d] + b * t2[d];
        double nq = norm(q, 3);
        if (nq > 0) {
            for (int d = 0; d < 3; ++d) cpx[d] = c[d] + R * q[d] / nq;
            double height = dot(nrm, cx, 3);
            double nc[3];
            cross(nrm, cx, nc);
            double radial = norm(nc, 3) - R;
            dist = std::sqrt(height * height + radial * radial);
        } else {
            const double* t = norm(t1, 3) > norm(t2, 3) ? t1 : t2;
            double nt = norm(t, 3);
            for (int d = 0; d < 3; ++d) cpx[d] = c[d] + R * t[d] / nt;
            dist = R;
        }
    } else if (ty == "Ellipsoid") {
        // The reference minimises |r(u,v) - x|^2 with LBFGS from the nearest of
        // 10000 parameter points (Surface.cpp:1363-1400,1459-1530; LBFGSpp,
        // epsilon 1e-6).  Here: the exact closest point by Eberly's robust
        // bisection ("Distance from a Point to an Ellipse, an Ellipsoid, or a
        // Hyperellipsoid"), which the LBFGS iteration approximates.
        cp_ellipsoid(P.data(), x, cpx);
        dist = distance(x, cpx, 3);
    } else if (ty == "Polyline") {
        cp_polyline(x, cpx, dist, bdy);
    } else if (ty == "Triangulation") {
        cp_triangles(x, cpx, dist);
    } else {
        throw std::runtime_error("Invalid surface '" + ty + "' in Surface::cp");
    }
}

// Arc of the circle of radius R about the origin in the xy-plane between the
// angles a1 < a2 (both in (-pi, pi]; z = 0 in 3-D), Surface::cpArc
// (Surface.cpp:507-579): points whose polar angle lies strictly inside the
// arc project radially (bdy 0); the others go to the nearer end along the
// circle, split at the angle opposite the arc's midpoint: end 1 at a1
// (bdy 1), end 2 at a2 (bdy 2).
void Surface::cp_arc(const double* x, double* cpx, double& dist, i64& bdy) const {
    const auto& P = s_.params;
    const double R = P[0], a1 = P[1], a2 = P[2];
    if (a1 < -M_PI) throw std::runtime_error("Arc: angle1 must lie in (-pi, pi]");
    if (a2 > M_PI) throw std::runtime_error("Arc: angle2 must lie in (-pi, pi]");
    if (a1 >= a2) throw std::runtime_error("Arc: angle1 must be less than angle2");
    const double half = 0.5 * (2 * M_PI - (a2 - a1));
    const double th = std::atan2(x[1], x[0]);
    double ang;
    if (th > a1 && th < a2) {
        ang = th;
        bdy = 0;
    } else if (th <= a1) {
        const bool first = th >= a1 - half;
        bdy = first ? 1 : 2;
        ang = first ? a1 : a2;
    } else {   // th >= a2
        const bool second = th < a2 + half;
        bdy = second ? 2 : 1;
        ang = second ? a2 : a1;
    }
    cpx[0] = R * std::cos(ang);
    cpx[1] = R * std::sin(ang);
    if (s_.dim == 3) cpx[2] = 0.0;
    double d2 = 0.0;
    for (int d = 0; d < s_.dim; ++d) d2 += sq(x[d] - cpx[d]);
    dist = std::sqrt(d2);
}

// Polyline closest point.  The reference delegates to fcpw (FP32 SAH BVH,
// Surface.cpp:1751-1779); here: exact FP64 brute force over segments, first
// minimum wins.  Parity with fcpw is unpinned (fcpw is not vendored).
void Surface::cp_polyline(const double* x, double* cpx, double& dist, i64& bdy) const {
    const int D = s_.dim;
    double best = std::numeric_limits<double>::infinity();
    for (const auto& f : s_.faces) {
        const double* a = s_.verts[(size_t)f[0]].data();
        const double* b = s_.verts[(size_t)f[1]].data();
        double ab[3] = {0, 0, 0}, ax[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) { ab[d] = b[d] - a[d]; ax[d] = x[d] - a[d]; }
        double L2 = dot(ab, ab, D);
        double t = L2 > 0 ? dot(ax, ab, D) / L2 : 0.0;
        t = std::min(1.0, std::max(0.0, t));
        double p[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) p[d] = a[d] + t * ab[d];
        double dd = distance(x, p, D);
        if (dd < best) {
            best = dd;
            for (int d = 0; d < D; ++d) cpx[d] = p[d];
        }
    }
    dist = best;
    bdy = 0;
    if (s_.is_open) {                              // Surface.cpp:2614-2636
        const double tol = 1e-7;
        if (distance(xp.front().data(), cpx, D) < tol) bdy = 1;
        else if (distance(xp.back().data(), cpx, D) < tol) bdy = 2;
    }
}

// Closest point on a triangle, Ericson "Real-Time Collision Detection" 5.1.5.
static void cp_tri(const double* p, const double* a, const double* b, const double* c, double* out) {
    double ab[3], ac[3], ap[3];
    for (int d = 0; d < 3; ++d) { ab[d] = b[d] - a[d]; ac[d] = c[d] - a[d]; ap[d] = p[d] - a[d]; }
    double d1 = dot(ab, ap, 3), d2 = dot(ac, ap, 3);
    if (d1 <= 0 && d2 <= 0) { for (int d = 0; d < 3; ++d) out[d] = a[d]; return; }
    double bp[3];
    for (int d = 0; d < 3; ++d) bp[d] = p[d] - b[d];
    double d3 = dot(ab, bp, 3), d4 = dot(ac, bp, 3);
    if (d3 >= 0 && d4 <= d3) { for (int d = 0; d < 3; ++d) out[d] = b[d]; return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        for (int d = 0; d < 3; ++d) out[d] = a[d] + v * ab[d];
        return;
    }
    double cp_[3];
    for (int d = 0; d < 3; ++d) cp_[d] = p[d] - c[d];
    double d5 = dot(ab, cp_, 3), d6 = dot(ac, cp_, 3);
    if (d6 >= 0 && d5 <= d6) { for (int d = 0; d < 3; ++d) out[d] = c[d]; return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        for (int d = 0; d < 3; ++d) out[d] = a[d] + w * ac[d];
        return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int d = 0; d < 3; ++d) out[d] = b[d] + w * (c[d] - b[d]);
        return;
    }
    double den = 1.0 / (va + vb + vc);
    double v = vb * den, w = vc * den;
    for (int d = 0; d < 3; ++d) out[d] = a[d] + ab[d] * v + ac[d] * w;
}

// Triangle-mesh closest point; fcpw in the reference (parity unpinned).
void Surface::cp_triangles(const double* x, double* cpx, double& dist) const {
    double best = std::numeric_limits<double>::infinity();
    for (const auto& f : s_.faces) {
        double q[3];
        cp_tri(x, s_.verts[(size_t)f[0]].data(), s_.verts[(size_t)f[1]].data(), s_.verts[(size_t)f[2]].data(), q);
        double dd = distance(x, q, 3);
        if (dd < best) { best = dd; cpx[0] = q[0]; cpx[1] = q[1]; cpx[2] = q[2]; }
    }
    dist = best;
}

void Surface::cpbar(const double* x, double* cpx, double& dist) const {   // :399-409
    double proj[3];
    for (int d = 0; d < s_.dim; ++d) proj[d] = 2.0 * cpx[d] - x[d];
    i64 bdy;
    cp(proj, cpx, dist, bdy);
}

std::vector<i64> Surface::nearest_two_xp(const double* x, double& min_dist, size_t np) const {   // :2661-2728
    const int D = s_.dim;
    double cpx[3], dist;
    i64 bdy;
    cp(x, cpx, dist, bdy);
    size_t mi = 0;
    min_dist = std::numeric_limits<double>::max();
    for (size_t i = 0; i < np; ++i) {
        double cd = distance(xp[i].data(), cpx, D);
        if (cd < min_dist) { min_dist = cd; mi = i; }
    }
    size_t nn;
    if (mi == 0) nn = 1;
    else if (mi == np - 1) nn = mi - 1;
    else if (distance(cpx, xp[mi].data(), D) < 1e-6) {
        nn = distance(xp[mi + 1].data(), cpx, D) > distance(xp[mi - 1].data(), cpx, D) ? mi - 1 : mi + 1;
    } else {
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, c[3] = {0, 0, 0}, t[3];
        for (int d = 0; d < D; ++d) {
            a[d] = cpx[d] - xp[mi][d];
            b[d] = xp[mi - 1][d] - xp[mi][d];
            c[d] = xp[mi + 1][d] - xp[mi][d];
        }
        cross(a, b, t);
        if (norm(t, 3) < 1e-4) nn = mi - 1;
        else {
            cross(a, c, t);
            if (norm(t, 3) < 1e-4) nn = mi + 1;
            else throw std::runtime_error("IBCNearestTwoXpIndices: problem finding the edge");
        }
    }
    return {(i64)mi, (i64)nn};
}

double Surface::ibc_parameter(const double* x, const Vec* vparam) const {   // :2731-2788
    const std::string& ty = s_.type;
    if (ty == "Point") return 0.0;
    // a closed polyline's vertex list holds each vertex once in the reference (its last edge
    // returns to vertex 0); the vertex chain here repeats vertex 0 at the end
    const size_t np = (ty == "Polyline" && !s_.is_open && xp.size() > 2) ? xp.size() - 1 : xp.size();
    if (vparam) {                                  // :2737-2751: nearest parameterisation point to x
        if (vparam->size() != np) throw std::runtime_error("one parameter per parameterisation point");
        double md = std::numeric_limits<double>::max();
        size_t mi = 0;
        for (size_t i = 0; i < np; ++i) {
            const double d = distance(x, xp[i].data(), s_.dim);
            if (d < md) { md = d; mi = i; }
        }
        return (*vparam)[mi];
    }
    if (ty == "Circle") return std::atan2(x[1], x[0]);
    const int D = s_.dim;
    double min_dist;
    auto ni = nearest_two_xp(x, min_dist, np);
    double total = 0.0, to_seg = 0.0;
    i64 lo = std::min(ni[0], ni[1]);
    for (size_t i = 0; i + 1 < np; ++i) {
        if ((i64)i == lo) to_seg = total;
        total += distance(xp[i].data(), xp[i + 1].data(), D);
    }
    double par;
    if (ni[0] < ni[1]) par = to_seg + min_dist;
    else par = to_seg + distance(xp[(size_t)ni[1]].data(), xp[(size_t)ni[0]].data(), D) - min_dist;
    return par / total;
}

// ---------------------------------------------------------------- Tube
i64 Tube::lin(const i64* t) const {
    for (int d = 0; d < dim; ++d)
        if (t[d] < 0 || t[d] >= N1d) throw std::runtime_error("grid index outside the virtual grid");
    return dim == 2 ? t[0] * N1d + t[1] : (t[0] * N1d + t[1]) * N1d + t[2];
}

void Tube::trip(i64 l, i64* t) const {
    if (dim == 2) { t[1] = l % N1d; t[0] = l / N1d; }
    else { t[2] = l % N1d; t[1] = (l / N1d) % N1d; t[0] = l / (N1d * N1d); }
}

void Tube::base_triplet(const double* xq, i64* t) const {   // Tube.cpp:502-531
    i64 shifts = (p % 2 == 0) ? p / 2 : (p - 1) / 2;
    for (int d = 0; d < dim; ++d) {
        double s = (xq[d] - xstart) / dx;
        t[d] = (i64)((p % 2 == 0) ? std::round(s) : std::floor(s)) - shifts;
    }
}

std::vector<i64> Tube::base_points(const std::vector<Vec>& xq) const {   // Tube.cpp:464-483
    std::vector<i64> b(xq.size());
    for (size_t i = 0; i < xq.size(); ++i) {
        i64 t[3];
        base_triplet(xq[i].data(), t);
        i64 ti = tube_index(lin(t));
        if (ti < 0 || ti >= N())
            throw std::runtime_error("Tube::ComputeBasePoints - Interpolation base point is not within the tube");
        b[i] = ti;
    }
    return b;
}

i64 Tube::nearest_tube_index(const double* xq) const {   // Tube.h:105-110, Grid.cpp:147-156
    i64 t[3];
    for (int d = 0; d < dim; ++d) t[d] = (i64)std::round((xq[d] - xstart) / dx);
    return tube_index(lin(t));
}

Tube build_tube(const SurfaceSpec& s, int p, double dx) {
    Tube T;
    T.dim = s.dim;
    T.p = p;
    T.dx = dx;
    T.surf = Surface(s);
    // ComputeTubeRadius, Tube.cpp:131-138 (m_order = 2)
    const int order = 2;
    T.rfac = 1.0001 * std::sqrt((T.dim - 1) * sq(0.5 * (p + 1)) + sq(0.5 * (order + p + 1)));
    T.r_t = T.rfac * dx;
    // Tube.cpp:148-157
    const size_t pad = 5;
    size_t spacings = (size_t)std::ceil((s.bbox[1] - s.bbox[0]) / dx);
    size_t n = (size_t)std::ceil(std::log2((double)(spacings + 2 * pad)));
    size_t sdiff = (size_t)std::ceil(0.5 * (std::pow(2, (double)n) - spacings - 2 * pad));
    T.xstart = s.bbox[0] - (pad + sdiff) * dx;
    T.n = (i64)n;
    T.N1d = (i64)std::pow(2.0, (double)n) + 1;
    i64 total = T.dim == 2 ? T.N1d * T.N1d : T.N1d * T.N1d * T.N1d;
    T.grid.assign((size_t)total, -1);

    const int D = T.dim;
    auto add_node = [&](i64 l, const double* x, const double* c, double dist, i64 bdy) {   // :554-605
        T.grid[(size_t)l] = (int32_t)T.rec.size();
        T.rec.push_back(l);
        for (int d = 0; d < D; ++d) { T.x.push_back(x[d]); T.cpx.push_back(c[d]); }
        T.dist.push_back(dist);
        T.bdy.push_back(bdy);
    };
    auto coord = [&](const i64* t, double* x) {   // Grid.cpp:35
        for (int d = 0; d < D; ++d) x[d] = T.xstart + t[d] * dx;
    };

    // FindStartingPoint, Tube.cpp:540-551
    double xs[3] = {T.xstart, T.xstart, T.xstart}, c[3], dist;
    i64 bdy;
    T.surf.cp(xs, c, dist, bdy);
    i64 idx[3];
    for (int d = 0; d < D; ++d) idx[d] = (i64)std::round((c[d] - T.xstart) / dx);
    i64 l0 = T.lin(idx);
    double x[3];
    coord(idx, x);
    T.surf.cp(x, c, dist, bdy);
    add_node(l0, x, c, dist, bdy);

    // FIFO breadth-first traversal, Tube.cpp:429-454
    std::list<i64> queue{l0};
    while (!queue.empty()) {
        i64 cur[3];
        T.trip(queue.front(), cur);
        for (int sd = 0; sd < D; ++sd)
            for (int s2 = 0; s2 < 2; ++s2) {
                i64 nb[3] = {cur[0], cur[1], cur[2]};
                nb[sd] += s2 == 0 ? 1 : -1;
                i64 l = T.lin(nb);
                if (T.grid[(size_t)l] == -1) {
                    T.grid[(size_t)l] = -2;
                    coord(nb, x);
                    T.surf.cp(x, c, dist, bdy);
                    if (std::fabs(dist) <= T.r_t) {
                        queue.push_back(l);
                        add_node(l, x, c, dist, bdy);
                    }
                }
            }
        queue.pop_front();
    }

    // SetNeighbours, Tube.cpp:676-694
    const i64 N = T.N();
    T.nbr.assign((size_t)N * 6, -1);
    for (i64 i = 0; i < N; ++i) {
        i64 cur[3];
        T.trip(T.rec[(size_t)i], cur);
        for (int d = 0; d < D; ++d)
            for (int s2 = 0; s2 < 2; ++s2) {
                i64 nb[3] = {cur[0], cur[1], cur[2]};
                nb[d] += s2 == 0 ? 1 : -1;
                i64 ti = T.grid[(size_t)T.lin(nb)];
                if (ti >= 0 && ti < N) T.nbr[(size_t)i * 6 + 2 * d + s2] = ti;
            }
    }
    // ComputeBasePoints, Tube.cpp:459-460
    std::vector<Vec> cps((size_t)N, Vec(D));
    for (i64 i = 0; i < N; ++i)
        for (int d = 0; d < D; ++d) cps[(size_t)i][d] = T.cpx[(size_t)(i * D + d)];
    T.bpt = T.base_points(cps);
    return T;
}

// ---------------------------------------------------------------- E and L
void weights_1d(const double* xs, double xq, int S, double* w) {   // Interpolation.cpp:128-200
    static const double c[6][5] = {{1}, {1, -1}, {1, -2, 1}, {1, -3, 3, -1}, {1, -4, 6, -4, 1}};
    bool hit = false;
    int hp = 0;
    for (int j = 0; j < S; ++j)
        if (std::fabs(xq - xs[j]) < 1e-17) { hit = true; hp = j; }
    if (hit) {
        for (int j = 0; j < S; ++j) w[j] = 0.0;
        w[hp] = 1.0;
        return;
    }
    double sum = 0.0;
    for (int j = 0; j < S; ++j) {
        w[j] = c[S - 1][j] / (xq - xs[j]);
        sum += w[j];
    }
    for (int j = 0; j < S; ++j) w[j] /= sum;
}

CSR build_E(const Tube& t, const std::vector<Vec>& xq, const std::vector<i64>* subset_of, i64 ncols) {
    // Interpolation ctor (:8-35) + BuildInterpolationWeights (:204-225) +
    // BuildInterpolationMatrix (:52-121).  The stencil walk by neighbour
    // pointers visits grid node (b + offset); a missing node would be a null
    // dereference in the reference, here it throws.
    const int D = t.dim, S = t.p + 1;
    std::vector<i64> b = t.base_points(xq);
    std::vector<Trip> trip;
    trip.reserve(xq.size() * (size_t)std::pow(S, D));
    std::vector<double> w((size_t)D * S);
    for (size_t q = 0; q < xq.size(); ++q) {
        i64 bt[3];
        t.trip(t.rec[(size_t)b[q]], bt);
        for (int d = 0; d < D; ++d) {
            double xs[8];
            for (int i = 0; i < S; ++i) xs[i] = t.xstart + (bt[d] + i) * t.dx;   // t.x() of the walked node
            weights_1d(xs, xq[q][d], S, &w[(size_t)d * S]);
        }
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j)
                for (int k = 0; k < (D == 3 ? S : 1); ++k) {
                    i64 g[3] = {bt[0] + i, bt[1] + j, D == 3 ? bt[2] + k : 0};
                    i64 ti = t.tube_index(t.lin(g));
                    if (ti < 0) throw std::runtime_error("interpolation stencil leaves the tube");
                    i64 colx = subset_of ? (*subset_of)[(size_t)ti] : ti;
                    if (colx < 0) continue;
                    double v = D == 3 ? w[2 * S + k] * w[S + j] * w[i] : w[S + j] * w[i];
                    trip.push_back({(i64)q, colx, v});
                }
    }
    return CSR::from_triplets((i64)xq.size(), ncols >= 0 ? ncols : t.N(), std::move(trip));
}

CSR build_E_tube(const Tube& t) {
    std::vector<Vec> cps((size_t)t.N(), Vec(t.dim));
    for (i64 i = 0; i < t.N(); ++i)
        for (int d = 0; d < t.dim; ++d) cps[(size_t)i][d] = t.cpx[(size_t)(i * t.dim + d)];
    return build_E(t, cps);
}

CSR build_L_subset(const Tube& t, const std::vector<i64>& set_of_sub, const std::vector<i64>& sub_of_set) {
    // FDMatrices.cpp:25-49
    std::vector<Trip> tr;
    const double h2 = t.dx * t.dx;
    for (size_t row = 0; row < set_of_sub.size(); ++row) {
        tr.push_back({(i64)row, (i64)row, -2.0 * t.dim});
        i64 si = set_of_sub[row];
        for (int nb = 0; nb < 2 * t.dim; ++nb) {
            i64 ni = t.nbr[(size_t)si * 6 + nb];
            if (ni >= 0 && sub_of_set[(size_t)ni] >= 0) tr.push_back({(i64)row, sub_of_set[(size_t)ni], 1.0});
        }
    }
    CSR L = CSR::from_triplets((i64)set_of_sub.size(), (i64)set_of_sub.size(), std::move(tr));
    for (auto& v : L.val) v /= h2;
    return L;
}

CSR build_L(const Tube& t) {
    std::vector<i64> id((size_t)t.N());
    std::iota(id.begin(), id.end(), 0);
    return build_L_subset(t, id, id);
}

std::vector<CSR> build_grad_subset(const Tube& t, const std::vector<i64>& set_of_sub,
                                   const std::vector<i64>& sub_of_set) {   // FDMatrices.cpp:66-92
    std::vector<std::vector<Trip>> tr((size_t)t.dim);
    for (size_t row = 0; row < set_of_sub.size(); ++row) {
        i64 si = set_of_sub[row];
        for (int nb = 0; nb < 2 * t.dim; ++nb) {
            i64 ni = t.nbr[(size_t)si * 6 + nb];
            if (ni >= 0 && sub_of_set[(size_t)ni] >= 0)
                tr[(size_t)nb / 2].push_back({(i64)row, sub_of_set[(size_t)ni], std::pow(-1.0, nb) * 0.5});
        }
    }
    std::vector<CSR> Dc;
    for (int d = 0; d < t.dim; ++d) {
        CSR A = CSR::from_triplets((i64)set_of_sub.size(), (i64)set_of_sub.size(), std::move(tr[(size_t)d]));
        for (auto& v : A.val) v /= t.dx;
        Dc.push_back(std::move(A));
    }
    return Dc;
}

CSR plot_E(const Tube& t) { return build_E(t, t.surf.xp); }

// ---------------------------------------------------------------- operator
Parallel g_parallel;

void System::multiply(const double* x, double* y, double mm, double nn) const {
    // GeneralizedBiCGStabSystem.h:213-233 with Interpolation :235-261,
    // LnodiagTimesX :93-115, LdiagTimesX :65-91, FinalizeMultiplication :166-189
    const i64 D = L->rows, dof = dofs();
    const bool par = g_parallel.apply;
    Vec tmp((size_t)D, 0.0);
#pragma omp parallel for schedule(static) if (par)
    for (i64 k = 0; k < E->rows; ++k) {
        if (ident[(size_t)k]) { tmp[(size_t)k] = x[k]; continue; }
        double s = 0.0;
        for (i64 e = E->ptr[(size_t)k]; e < E->ptr[(size_t)k + 1]; ++e) s += E->val[(size_t)e] * x[E->col[(size_t)e]];
        tmp[(size_t)k] = s;
    }
#pragma omp parallel for schedule(static) if (par)
    for (i64 k = 0; k < D; ++k) {      // result = Loff * tmp
        double s = 0.0;
        for (i64 e = L->ptr[(size_t)k]; e < L->ptr[(size_t)k + 1]; ++e)
            if (L->col[(size_t)e] != k) s += L->val[(size_t)e] * tmp[(size_t)L->col[(size_t)e]];
        y[k] = s;
    }
    Vec ld((size_t)D, 0.0);              // tmp = Ldiag * x
#pragma omp parallel for schedule(static) if (par)
    for (i64 k = 0; k < dof; ++k)
        for (i64 e = L->ptr[(size_t)k]; e < L->ptr[(size_t)k + 1]; ++e)
            if (L->col[(size_t)e] == k) ld[(size_t)k] += L->val[(size_t)e] * x[k];
    for (int i = 0; i < nD; ++i) ld[(size_t)(dof + i)] = (mm - nn) * x[dof + i];
#pragma omp parallel for schedule(static) if (par)
    for (i64 q = 0; q < dof; ++q) {
        if (ident[(size_t)q]) y[q] = x[q];
        else y[q] = mm * x[q] + nn * (ld[(size_t)q] + y[q]);
    }
    for (int i = 0; i < nD; ++i) y[dof + i] = (mm - nn) * x[dof + i];
}

void System::diag_inverse(const double* in, double* out) const {   // :281-340 (computed == false)
    const i64 D = L->rows, dof = dofs();
#pragma omp parallel for schedule(static) if (g_parallel.apply)
    for (i64 k = 0; k < D; ++k) {
        if (ident[(size_t)k]) { out[k] = in[k]; continue; }
        bool found = false;
        for (i64 e = L->ptr[(size_t)k]; e < L->ptr[(size_t)k + 1]; ++e)
            if (L->col[(size_t)e] == k) {
                double v = m + L->val[(size_t)e] * n;
                out[k] = in[k] / v;
                found = true;
                break;
            }
        if (!found) out[k] = in[k];
    }
    for (int i = 0; i < nD; ++i) out[dof + i] = in[dof + i] / (m - n);
}

void System::precondition(const double* in, double* out) const {   // :350-385
    const i64 D = L->rows;
    if (precond == PRECOND_DIAGONAL) {
        diag_inverse(in, out);
    } else if (precond == PRECOND_JACOBI) {
        // damped Jacobi, omega = 2/3, precond_iters sweeps of a full apply
        const double omega = 2.0 / 3.0;
        Vec tmp((size_t)D), res((size_t)D);
        for (i64 q = 0; q < D; ++q) out[q] = 0.0;
        for (int it = 0; it < precond_iters; ++it) {
            multiply(out, tmp.data());
            for (i64 q = 0; q < D; ++q) res[(size_t)q] = in[q] - tmp[(size_t)q];
            diag_inverse(res.data(), res.data());
            for (i64 q = 0; q < dofs(); ++q) out[q] += omega * res[(size_t)q];
        }
    } else {
        for (i64 q = 0; q < D; ++q) out[q] = in[q];
    }
}

static double vdot(const Vec& a, const Vec& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

SolveInfo bicgstab(const System& sys, const double* bp, double* xp, double tol, int max_it) {
    // BiCGStab.h:30-108.  x starts at zero (Solver.cpp:71).
    const size_t D = (size_t)sys.L->rows;
    Vec x(D, 0.0), b(bp, bp + D), r(D), r0(D, 0.0), v(D, 0.0), p(D, 0.0), s(D), t(D), y(D), z(D);
    SolveInfo info;
    sys.multiply(x.data(), r0.data());
    for (size_t i = 0; i < D; ++i) r[i] = b[i] - r0[i];
    r0 = r;
    double r0_sq = vdot(r0, r0);
    double rhs_sq = vdot(b, b);
    if (rhs_sq == 0.0) {
        std::fill(xp, xp + D, 0.0);
        return info;
    }
    double rho = 1, alpha = 1, w = 1;
    const double tol2 = tol * tol * rhs_sq;
    const double eps = std::numeric_limits<double>::epsilon();
    const double eps2 = eps * eps;
    int iters = 0, restarts = 0;
    while (vdot(r, r) > tol2 && iters < max_it) {
        double rho_old = rho;
        rho = vdot(r0, r);
        if (std::fabs(rho) < eps2 * r0_sq) {
            sys.multiply(x.data(), r0.data());
            for (size_t i = 0; i < D; ++i) r[i] = b[i] - r0[i];
            r0 = r;
            rho = r0_sq = vdot(r, r);
            if (restarts++ == 0) iters = 0;
        }
        double beta = (rho / rho_old) * (alpha / w);
        const bool pv = g_parallel.vec;
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) p[i] = -w * v[i] + p[i];
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) p[i] *= beta;
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) p[i] += r[i];
        sys.precondition(p.data(), y.data());
        sys.multiply(y.data(), v.data());
        alpha = rho / vdot(r0, v);
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) s[i] = -alpha * v[i] + r[i];
        sys.precondition(s.data(), z.data());
        sys.multiply(z.data(), t.data());
        double val = vdot(t, t);
        w = val > 0.0 ? vdot(t, s) / val : 0.0;
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) x[i] = alpha * y[i] + x[i];
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) x[i] = w * z[i] + x[i];
#pragma omp parallel for schedule(static) if (pv)
        for (size_t i = 0; i < D; ++i) r[i] = -w * t[i] + s[i];
        ++iters;
    }
    info.iters = iters;
    info.restarts = restarts;
    info.rel_res = std::sqrt(vdot(r, r) / rhs_sq);
    std::copy(x.begin(), x.end(), xp);
    return info;
}

void set_identity_rows(CSR& A, const std::vector<uint8_t>& ident) {   // Helpers.cpp:403-442
    std::vector<Trip> tr;
    for (i64 r = 0; r < A.rows; ++r) {
        bool has_diag = false;
        for (i64 e = A.ptr[(size_t)r]; e < A.ptr[(size_t)r + 1]; ++e) {
            double v = A.val[(size_t)e];
            if (ident[(size_t)r]) {
                if (A.col[(size_t)e] == r) { has_diag = true; v = 1.0; }
                else v = 0.0;
            }
            tr.push_back({r, A.col[(size_t)e], v});
        }
        if (!has_diag && ident[(size_t)r]) tr.push_back({r, r, 1.0});
    }
    A = CSR::from_triplets(A.rows, A.cols, std::move(tr));
    A.prune_zeros();
}

CSR explicit_operator(const CSR& L, const CSR& E, double m, double n,
                      const std::vector<uint8_t>& ident, int nD) {
    // m I + n (diag(L) + (L - diag L) E), identity rows set, trailing nD rows
    // (m - n) on the diagonal (GeneralizedBiCGStabSystem.h:185-188).
    const i64 D = L.rows, dof = D - nD;
    std::vector<Trip> tr;
    for (i64 r = 0; r < D; ++r) {
        if (r >= dof) { tr.push_back({r, r, m - n}); continue; }
        if (ident[(size_t)r]) { tr.push_back({r, r, 1.0}); continue; }
        tr.push_back({r, r, m});
        for (i64 e = L.ptr[(size_t)r]; e < L.ptr[(size_t)r + 1]; ++e) {
            i64 c = L.col[(size_t)e];
            double lv = L.val[(size_t)e];
            if (c == r) { tr.push_back({r, r, n * lv}); continue; }
            // row c of the interpolation (identity rows of E pass x through)
            if (ident[(size_t)c]) { tr.push_back({r, c, n * lv}); continue; }
            for (i64 f = E.ptr[(size_t)c]; f < E.ptr[(size_t)c + 1]; ++f)
                tr.push_back({r, E.col[(size_t)f], n * lv * E.val[(size_t)f]});
        }
    }
    return CSR::from_triplets(D, D, std::move(tr));
}

// ---------------------------------------------------------------- Poisson
PoissonProblem make_poisson(const SurfaceSpec& s, double dx, int bdy_order, std::vector<i64> which_bdy, int p) {
    // problems/Poisson.cpp:5-18, 164-238
    PoissonProblem P;
    P.tube = build_tube(s, p, dx);
    P.bdy_order = bdy_order;
    P.which_bdy = which_bdy;
    Tube& t = P.tube;
    const i64 N = t.N();
    if (bdy_order == 2) {   // Helpers.cpp:267-288, ReplaceClosestPointWithcpBar
        for (i64 i = 0; i < N; ++i)
            if (t.bdy[(size_t)i] > 0) {
                double c[3], dist;
                for (int d = 0; d < t.dim; ++d) c[d] = t.cpx[(size_t)(i * t.dim + d)];
                t.surf.cpbar(&t.x[(size_t)(i * t.dim)], c, dist);
                for (int d = 0; d < t.dim; ++d) t.cpx[(size_t)(i * t.dim + d)] = c[d];
            }
    }
    P.L = build_L(t);
    P.E = build_E_tube(t);
    auto is_bdy = [&](i64 i) {
        for (i64 b : which_bdy) if (t.bdy[(size_t)i] == b) return true;
        return false;
    };
    if (bdy_order == 2 && !which_bdy.empty())
        for (i64 r = 0; r < P.E.rows; ++r)
            if (is_bdy(r))
                for (i64 e = P.E.ptr[(size_t)r]; e < P.E.ptr[(size_t)r + 1]; ++e) P.E.val[(size_t)e] *= -1.0;
    P.ident.assign((size_t)N, 0);
    if (!which_bdy.empty()) {
        if (bdy_order == 1) {
            for (i64 i = 0; i < N; ++i) if (is_bdy(i)) P.ident[(size_t)i] = 1;
            set_identity_rows(P.E, P.ident);
            P.Efull = P.E;
            P.Lfull = P.L;
        } else {
            throw std::runtime_error("make_poisson: second-order boundary rows are 2-D only in the reference");
        }
    } else {
        P.Lfull = P.L;
        P.Efull = P.E;
    }
    return P;
}

double average_order(const Vec& err, const Vec& dx) {   // Helpers.cpp:446-463
    double s = 0.0;
    for (size_t i = 1; i < err.size(); ++i) s += std::log(err[i - 1] / err[i]) / std::log(dx[i - 1] / dx[i]);
    return s / (double)(err.size() - 1);
}

}  // namespace orc
