// Interior boundary conditions: restatement of src/IBCSubset.cpp,
// src/IBCDOFSubset.cpp, src/IBCSubsets.cpp, src/IBCMatrixManipulations.cpp,
// src/Geometry.cpp and problems/PoissonWithIBC.cpp (3-D embeddings).
// TEST INFRASTRUCTURE ONLY (see orc.h).
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <list>
#include <numeric>

#include "orc_ibc.h"

namespace orc {

using V3 = std::array<double, 3>;

static V3 sub3(const double* a, const double* b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double norm3(const V3& a) { return std::sqrt(dot3(a, a)); }
static V3 cross3(const V3& u, const V3& v) {
    return {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
}
static void normalize3(V3& u, double tol) {   // VectorMath.cpp:40-58
    double n = norm3(u);
    if (n > tol) for (auto& c : u) c /= n;
    else u = {0, 0, 0};
}

// ---------------------------------------------------------------- 3x3 symmetric eigen
// Cyclic Jacobi rotations; eigenvalues ascending, eigenvectors as columns
// (Eigen's SelfAdjointEigenSolver convention: columns ordered like the
// values; signs are arbitrary and every use below is sign-invariant).
void sym_eig3(const double A_in[3][3], double w[3], double V[3][3]) {
    double A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { A[i][j] = A_in[i][j]; V[i][j] = i == j; }
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double scale = 0;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) scale += A[i][j] * A[i][j];
        if (off <= 1e-40 * scale || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {   // A = J^T A J
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int a, int b) { return A[a][a] < A[b][b]; });
    double W[3][3];
    for (int j = 0; j < 3; ++j) {
        w[j] = A[idx[j]][idx[j]];
        for (int i = 0; i < 3; ++i) W[i][j] = V[i][idx[j]];
    }
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) V[i][j] = W[i][j];
}

// ---------------------------------------------------------------- Geometry
// EigenDecompositionOfJcp, Geometry.cpp:27-69 + EigenDecomposition :72-98.
static void eig_jcp(const Tube& t, const std::vector<Vec>& subset_cp, const std::vector<i64>& set_of_sub,
                    const std::vector<i64>& sub_of_set, std::vector<std::array<double, 9>>& vecs,
                    std::vector<V3>& vals) {
    const size_t n = set_of_sub.size();
    std::vector<CSR> Dc = build_grad_subset(t, set_of_sub, sub_of_set);
    CSR E = build_E(t, subset_cp, &sub_of_set, (i64)n);
    std::vector<Vec> cp(3, Vec(n));
    for (size_t i = 0; i < n; ++i) for (int d = 0; d < 3; ++d) cp[d][i] = subset_cp[i][d];
    Vec tmp(n);
    std::vector<Vec> J(9, Vec(n));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            Dc[c].matvec(cp[r].data(), tmp.data());
            E.matvec(tmp.data(), J[r * 3 + c].data());
        }
    vecs.resize(n);
    vals.resize(n);
    for (size_t i = 0; i < n; ++i) {
        double A[3][3], V[3][3], w[3];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A[r][c] = J[r * 3 + c][i];
        double S[3][3];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) S[r][c] = 0.5 * (A[r][c] + A[c][r]);
        sym_eig3(S, w, V);
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) vecs[i][r * 3 + c] = V[r][c];
        vals[i] = {w[0], w[1], w[2]};
    }
}

// ApplyInterpolationToUnorientedVectors, Geometry.cpp:239-266.
static std::vector<V3> interp_unoriented(const Tube& t, const std::vector<i64>& sub_of_set, i64 nsub,
                                         const std::vector<V3>& v, const std::vector<Vec>& xq, bool normalize) {
    CSR E = build_E(t, xq, &sub_of_set, nsub);
    std::vector<V3> out(xq.size(), V3{0, 0, 0});
    for (i64 r = 0; r < E.rows; ++r) {
        V3 first{0, 0, 0};
        bool have = false;
        for (i64 e = E.ptr[(size_t)r]; e < E.ptr[(size_t)r + 1]; ++e) {
            const V3& vc = v[(size_t)E.col[(size_t)e]];
            if (!have) { first = vc; have = true; }
            double s = dot3(first, vc) > 0.0 ? 1.0 : -1.0;
            for (int d = 0; d < 3; ++d) out[(size_t)r][d] += s * (E.val[(size_t)e] * vc[d]);
        }
    }
    if (normalize) for (auto& o : out) normalize3(o, 0.0);
    else {
        double ns = std::pow(t.p + 1, 3);
        for (auto& o : out) for (auto& c : o) c /= ns;
    }
    return out;
}

// InterpolateUnorientedVectorFromNeighbours, Geometry.cpp:144-188.
static void interp_from_neighbours(const Tube& t, std::vector<V3>& sv, const std::vector<i64>& set_of_sub,
                                   const std::vector<i64>& sub_of_set, size_t si, double tol) {
    V3 acc{0, 0, 0}, first{0, 0, 0};
    size_t num = 0;
    i64 idx = set_of_sub[si];
    for (int nb = 0; nb < 6; ++nb) {
        i64 ni = t.nbr[(size_t)idx * 6 + nb];
        if (ni < 0) continue;
        i64 sn = sub_of_set[(size_t)ni];
        if (sn < 0) continue;
        const V3& w = sv[(size_t)sn];
        if (norm3(w) >= tol) {
            if (num == 0) first = w;
            if (dot3(first, w) > 0) for (int d = 0; d < 3; ++d) acc[d] += w[d];
            else for (int d = 0; d < 3; ++d) acc[d] -= w[d];
            ++num;
        }
    }
    if (num > 0) for (int d = 0; d < 3; ++d) sv[si][d] = acc[d] / (double)num;
}

// ---------------------------------------------------------------- subsets

// Whole-tube unoriented surface normals from Jcp (IBCSubset.cpp:192-226).
std::vector<V3> tube_jcp_normals(const Tube& t) {
    std::vector<i64> id((size_t)t.N());
    std::iota(id.begin(), id.end(), 0);
    std::vector<Vec> cps((size_t)t.N(), Vec(3));
    for (i64 i = 0; i < t.N(); ++i) for (int d = 0; d < 3; ++d) cps[(size_t)i][d] = t.cpx[(size_t)(3 * i + d)];
    std::vector<std::array<double, 9>> vecs;
    std::vector<V3> vals;
    eig_jcp(t, cps, id, id, vecs, vals);
    std::vector<V3> n((size_t)t.N());
    for (size_t i = 0; i < n.size(); ++i) n[i] = {vecs[i][0], vecs[i][3], vecs[i][6]};
    return n;
}

void compute_tagging_subset(const Tube& t, const Surface& ibc, const IBCMeta& meta, const std::vector<V3>* tube_normals,
                            TagSubset& S) {
    // ComputeSubsetNearIBC, IBCSubset.cpp:39-109
    const double radius = 2.0 * t.r_t;
    i64 start = t.nearest_tube_index(ibc.xp[0].data());   // FindStartingPoint :228-232
    if (start < 0) throw std::runtime_error("IBC starting point is not in the tube");
    std::vector<uint8_t> visited((size_t)t.N(), 0);
    std::list<i64> queue{start};
    while (!queue.empty()) {
        i64 f = queue.front();
        for (int nb = 0; nb < 6; ++nb) {
            i64 ni = t.nbr[(size_t)f * 6 + nb];
            if (ni < 0 || visited[(size_t)ni]) continue;
            visited[(size_t)ni] = 1;
            double c[3], dist;
            i64 bdy;
            ibc.cp(&t.x[(size_t)(3 * ni)], c, dist, bdy);
            if (dist < radius) {
                S.cpc.push_back({c[0], c[1], c[2]});
                S.dist.push_back(dist);
                S.bdy.push_back(bdy);
                S.set_of_sub.push_back(ni);
                queue.push_back(ni);
            }
        }
        queue.pop_front();
    }
    const size_t n = S.set_of_sub.size();
    S.sub_of_set.assign((size_t)t.N(), -1);
    for (size_t i = 0; i < n; ++i) S.sub_of_set[(size_t)S.set_of_sub[i]] = (i64)i;
    // ComputeCPdiff, IBCSubset.cpp:156-187
    S.cp_diff.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const i64 si = S.set_of_sub[i];
        if (S.bdy[i] > 0 && meta.use_cp_diff_directions) {
            double proj[3], a[3], b[3], d;
            i64 bb;
            for (int k = 0; k < 3; ++k) proj[k] = 2.0 * S.cpc[i][k] - t.x[(size_t)(3 * si + k)];
            t.surf.cp(proj, a, d, bb);
            ibc.cp(proj, b, d, bb);
            S.cp_diff[i] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        } else {
            S.cp_diff[i] = sub3(&t.cpx[(size_t)(3 * si)], S.cpc[i].data());
        }
    }
    for (size_t i = 0; i < n; ++i)
        if (norm3(S.cp_diff[i]) < meta.on_ibc_tolerance)
            interp_from_neighbours(t, S.cp_diff, S.set_of_sub, S.sub_of_set, i, meta.on_ibc_tolerance);
    if (meta.use_cp_diff_directions) {
        S.directions = S.cp_diff;
        return;
    }
    // ComputeFrameBasedDirections, IBCSubset.cpp:190-268 (3-D)
    if (!meta.use_Jcp_surface_normals) throw std::runtime_error("x - cp surface normals are not restated");
    std::vector<i64> id((size_t)t.N());
    std::iota(id.begin(), id.end(), 0);
    std::vector<Vec> cpc_v(n, Vec(3));
    for (size_t i = 0; i < n; ++i) for (int d = 0; d < 3; ++d) cpc_v[i][d] = S.cpc[i][d];
    S.normals = interp_unoriented(t, id, t.N(), *tube_normals, cpc_v, true);
    std::vector<std::array<double, 9>> vecs;
    std::vector<V3> vals;
    eig_jcp(t, cpc_v, S.set_of_sub, S.sub_of_set, vecs, vals);
    S.tangents.assign(n, V3{0, 0, 0});
    S.binormals.resize(n);
    S.directions.resize(n);
    for (size_t i = 0; i < n; ++i) {
        if (std::fabs(vals[i][2]) < 1e-8) {   // point IBC
            S.binormals[i] = {0, 0, 0};
            const double s = dot3(S.normals[i], S.cp_diff[i]);
            for (int d = 0; d < 3; ++d) S.directions[i][d] = S.cp_diff[i][d] - s * S.normals[i][d];
        } else {
            S.tangents[i] = {vecs[i][2], vecs[i][5], vecs[i][8]};
            S.binormals[i] = cross3(S.tangents[i], S.normals[i]);
            const double s = dot3(S.binormals[i], S.cp_diff[i]);
            for (int d = 0; d < 3; ++d) S.directions[i][d] = s * S.binormals[i][d];
        }
    }
}

void compute_dof_subset(const Tube& t, const Surface& ibc, const IBCMeta& meta, const TagSubset& T, DofSubset& S) {
    // IBCDOFSubset::ComputeDOFSubset, IBCDOFSubset.cpp:10-93
    for (size_t i = 0; i < T.set_of_sub.size(); ++i)
        if (T.dist[i] < meta.ibc_DOFSubset_radius_scale * t.r_t) {
            S.set_of_sub.push_back(T.set_of_sub[i]);
            S.cpc.push_back(T.cpc[i]);
            S.cp_diff.push_back(T.cp_diff[i]);
            S.dist.push_back(T.dist[i]);
            S.bdy.push_back(T.bdy[i]);
            S.directions.push_back(T.directions[i]);
        }
    const size_t n = S.set_of_sub.size();
    S.sub_of_set.assign((size_t)t.N(), -1);
    for (size_t i = 0; i < n; ++i) S.sub_of_set[(size_t)S.set_of_sub[i]] = (i64)i;
    S.which_side.assign(n, 0);
    if (meta.is_oriented) throw std::runtime_error("oriented IBCs are not restated in 3-D");
    if ((meta.boundary_type[0] == 0 || meta.boundary_type[1] == 0) && meta.boundary_order == 2) {
        for (size_t i = 0; i < n; ++i)
            if (meta.boundary_type[0] == 0) S.dir_set_of_sub.push_back((i64)i);
        S.dir_sub_of_set.assign(n, -1);
        for (size_t i = 0; i < S.dir_set_of_sub.size(); ++i) S.dir_sub_of_set[(size_t)S.dir_set_of_sub[i]] = (i64)i;
    }
    std::vector<Vec> xq(n, Vec(3));
    if (meta.boundary_order == 2) {
        S.cpbar.resize(n);
        for (size_t i = 0; i < n; ++i) {
            double proj[3], c[3], d;
            i64 b;
            for (int k = 0; k < 3; ++k) proj[k] = 2.0 * S.cpc[i][k] - t.x[(size_t)(3 * S.set_of_sub[i] + k)];
            t.surf.cp(proj, c, d, b);
            S.cpbar[i] = {c[0], c[1], c[2]};
            for (int k = 0; k < 3; ++k) xq[i][k] = c[k];
        }
    } else {
        for (size_t i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) xq[i][k] = S.cpc[i][k];
    }
    S.E = build_E(t, xq);   // ComputeExtensionMatrix :110-125
}

// ---------------------------------------------------------------- matrix manipulations
// TagStencils, IBCMatrixManipulations.cpp:423-463: returns the tag (1 or 2) of every entry of A (CSR order).
static std::vector<uint8_t> tag_stencils(const CSR& A, const IBCData& I, const std::vector<V3>& director_dirs,
                                         const std::vector<i64>& director_sub_of_row, size_t c, double flip) {
    const TagSubset& T = I.tag[c];
    std::vector<uint8_t> tag((size_t)A.nnz(), 1);
    for (i64 r = 0; r < A.rows; ++r) {
        const i64 di = r < (i64)director_sub_of_row.size() ? director_sub_of_row[(size_t)r] : -1;
        if (di < 0) continue;
        for (i64 e = A.ptr[(size_t)r]; e < A.ptr[(size_t)r + 1]; ++e) {
            const i64 col = A.col[(size_t)e];
            const i64 sp = col < (i64)T.sub_of_set.size() ? T.sub_of_set[(size_t)col] : -1;
            if (sp < 0 || T.bdy[(size_t)sp] != 0) continue;
            const double dp = flip * dot3(director_dirs[(size_t)di], T.directions[(size_t)sp]);
            if (dp < 0.0) tag[(size_t)e] = 2;
        }
    }
    return tag;
}

// MoveToIBCColumns, IBCMatrixManipulations.cpp:189-216.
static void move_to_ibc_columns(const CSR& A, const std::vector<std::vector<uint8_t>>& tags, const IBCData& I,
                                std::vector<Trip>& out) {
    for (i64 r = 0; r < A.rows; ++r)
        for (i64 e = A.ptr[(size_t)r]; e < A.ptr[(size_t)r + 1]; ++e) {
            const i64 col = A.col[(size_t)e];
            size_t moved = 0;
            for (size_t c = 0; c < I.dof.size(); ++c) {
                const i64 ds = I.dof[c].sub_of_set[(size_t)col];
                if (tags[c][(size_t)e] == 2 && ds >= 0) {
                    ++moved;
                    out.push_back({r, ds + I.start[c], A.val[(size_t)e]});
                }
            }
            if (moved == 0) out.push_back({r, col, A.val[(size_t)e]});
        }
}

IBCData make_ibc(const Tube& t, const std::vector<SurfaceSpec>& specs) {
    IBCData I;
    std::vector<V3> normals;
    bool need_normals = false;
    for (const auto& s : specs) need_normals |= !s.meta.use_cp_diff_directions;
    if (need_normals) normals = tube_jcp_normals(t);
    for (const auto& s : specs) {
        I.surf.emplace_back(s);
        I.meta.push_back(s.meta);
    }
    // IBCSubsets::ComputeSubsets, IBCSubsets.cpp:54-71
    I.tag.resize(specs.size());
    I.dof.resize(specs.size());
    for (size_t c = 0; c < specs.size(); ++c) compute_tagging_subset(t, I.surf[c], I.meta[c], &normals, I.tag[c]);
    for (size_t c = 0; c < specs.size(); ++c) compute_dof_subset(t, I.surf[c], I.meta[c], I.tag[c], I.dof[c]);
    // IBCStartingColumns :74-95 and TotalDOFs :98-116
    const size_t C = specs.size();
    I.start.assign(C, 0);
    I.dir_start.assign(C, 0);
    if (C > 0) { I.start[0] = t.N(); I.dir_start[0] = t.N(); }
    auto second_dir = [&](size_t c) {
        return (I.meta[c].boundary_type[0] == 0 || I.meta[c].boundary_type[1] == 0) && I.meta[c].boundary_order == 2;
    };
    for (size_t c = 1; c < C; ++c) {
        I.start[c] = I.start[c - 1] + (i64)I.dof[c - 1].set_of_sub.size();
        if (second_dir(c)) I.dir_start[c] = I.dir_start[c - 1] + (i64)I.dof[c - 1].dir_set_of_sub.size();
    }
    I.identity_rows_start = 0;
    I.num2 = 0;
    I.total = 0;
    for (size_t c = 0; c < C; ++c) {
        if (second_dir(c)) {
            I.identity_rows_start += 1;
            I.num2 += (i64)I.dof[c].dir_set_of_sub.size();
        }
        I.total += (i64)I.dof[c].set_of_sub.size();
    }
    if (I.identity_rows_start != 0) I.identity_rows_start = I.total;
    return I;
}

// AddIBCExtension, IBCMatrixManipulations.cpp:12-35 (+ AddExtensionIBCRows :118-186).
CSR add_ibc_extension(const CSR& E, const Tube& t, const IBCData& I) {
    const size_t C = I.dof.size();
    std::vector<std::vector<uint8_t>> tags(C);
    for (size_t c = 0; c < C; ++c) tags[c] = tag_stencils(E, I, I.tag[c].directions, I.tag[c].sub_of_set, c, 1.0);
    std::vector<Trip> tr;
    move_to_ibc_columns(E, tags, I, tr);
    for (size_t c = 0; c < C; ++c) {
        const DofSubset& S = I.dof[c];
        const IBCMeta& m = I.meta[c];
        const size_t n = S.set_of_sub.size();
        std::vector<i64> ident_rows(n);
        std::iota(ident_rows.begin(), ident_rows.end(), 0);
        auto et = tag_stencils(S.E, I, S.directions, ident_rows, c, -1.0);
        for (i64 r = 0; r < S.E.rows; ++r) {
            const int type = m.boundary_type[0];   // not oriented
            const double scale = type == 0 ? -1.0 : 1.0;
            if (!(type == 1 || (type == 0 && m.boundary_order == 2))) continue;
            for (i64 e = S.E.ptr[(size_t)r]; e < S.E.ptr[(size_t)r + 1]; ++e) {
                const i64 col = S.E.col[(size_t)e];
                const i64 ds = S.sub_of_set[(size_t)col];
                if (et[(size_t)e] == 2 && ds >= 0) tr.push_back({r + I.start[c], ds + I.start[c], scale * S.E.val[(size_t)e]});
                else tr.push_back({r + I.start[c], col, scale * S.E.val[(size_t)e]});
            }
        }
        if (m.boundary_order == 2) {
            for (size_t i = 0; i < n; ++i)
                if (m.boundary_type[0] == 0)
                    tr.push_back({(i64)i + I.start[c], I.identity_rows_start + S.dir_sub_of_set[i] + I.dir_start[c], 2.0});
            for (size_t i = 0; i < n; ++i)
                if (m.boundary_type[0] == 0) {
                    const i64 rr = I.identity_rows_start + S.dir_sub_of_set[i] + I.dir_start[c];
                    tr.push_back({rr, rr, 1.0});
                }
        } else {
            for (size_t i = 0; i < n; ++i)
                if (m.boundary_type[0] == 0) {
                    const i64 rr = I.identity_rows_start + (i64)i + I.start[c];
                    tr.push_back({rr, rr, 1.0});
                }
        }
    }
    const i64 D = E.rows + I.total + I.num2;
    return CSR::from_triplets(D, D, std::move(tr));
}

// AddIBCFiniteDifference, IBCMatrixManipulations.cpp:38-71 (+ AddFiniteDifferenceIBCRows :219-251).
CSR add_ibc_finite_difference(const CSR& L, const Tube& t, const IBCData& I) {
    const size_t C = I.dof.size();
    std::vector<std::vector<uint8_t>> tags(C);
    for (size_t c = 0; c < C; ++c) tags[c] = tag_stencils(L, I, I.tag[c].directions, I.tag[c].sub_of_set, c, 1.0);
    std::vector<Trip> tr;
    move_to_ibc_columns(L, tags, I, tr);
    for (i64 r = 0; r < L.rows; ++r)
        for (i64 e = L.ptr[(size_t)r]; e < L.ptr[(size_t)r + 1]; ++e) {
            const i64 col = L.col[(size_t)e];
            for (size_t c = 0; c < C; ++c) {
                const i64 rs = I.dof[c].sub_of_set[(size_t)r];
                if (tags[c][(size_t)e] == 2 && rs >= 0) {
                    tr.push_back({rs + I.start[c], col, L.val[(size_t)e]});
                } else if (rs >= 0) {
                    const i64 cs = I.dof[c].sub_of_set[(size_t)col];
                    if (cs >= 0) tr.push_back({rs + I.start[c], cs + I.start[c], L.val[(size_t)e]});
                }
            }
        }
    for (size_t c = 0; c < C; ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i)
            if (I.meta[c].boundary_type[0] == 0 && I.meta[c].boundary_order == 2) {
                const i64 rr = I.identity_rows_start + I.dof[c].dir_sub_of_set[i] + I.dir_start[c];
                tr.push_back({rr, rr, 1.0});
            }
    const i64 D = L.rows + I.total + I.num2;
    return CSR::from_triplets(D, D, std::move(tr));
}

// PoissonWithIBC::Initialize, problems/PoissonWithIBC.cpp:162-192.
PoissonIBCProblem make_poisson_ibc(const SurfaceSpec& s, const std::vector<SurfaceSpec>& ibc_specs, double dx, int p) {
    PoissonIBCProblem P;
    P.tube = build_tube(s, p, dx);
    P.ibc = make_ibc(P.tube, ibc_specs);
    P.L = build_L(P.tube);
    P.E = build_E_tube(P.tube);
    P.Lfull = add_ibc_finite_difference(P.L, P.tube, P.ibc);
    P.Efull = add_ibc_extension(P.E, P.tube, P.ibc);
    const IBCData& I = P.ibc;
    P.ident.assign((size_t)P.Lfull.rows, 0);
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i)
            if (I.meta[c].boundary_type[0] == 0) {
                if (I.meta[c].boundary_order == 2)
                    P.ident[(size_t)(I.identity_rows_start + I.dof[c].dir_sub_of_set[i] + I.dir_start[c])] = 1;
                else
                    P.ident[(size_t)(I.identity_rows_start + (i64)i + I.start[c])] = 1;
            }
    return P;
}

// PoissonWithIBC::GetRHS(f, surface_exact), problems/PoissonWithIBC.cpp:50-90.
Vec poisson_ibc_rhs(const PoissonIBCProblem& P, const std::function<double(const double*)>& f,
                    const std::function<double(const double*)>& exact) {
    const IBCData& I = P.ibc;
    Vec rhs((size_t)P.Lfull.rows, 0.0);
    for (i64 i = 0; i < P.tube.N(); ++i) rhs[(size_t)i] = f(&P.tube.cpx[(size_t)(3 * i)]);
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i) {
            const int ty = I.meta[c].boundary_type[0];
            if (ty == 1 || (ty == 0 && I.meta[c].boundary_order == 2))
                rhs[(size_t)((i64)i + I.start[c])] = f(I.dof[c].cpc[i].data());
        }
    for (size_t c = 0; c < I.dof.size(); ++c)
        for (size_t i = 0; i < I.dof[c].set_of_sub.size(); ++i)
            if (I.meta[c].boundary_type[0] == 0) {
                const i64 r = I.meta[c].boundary_order == 2
                                  ? I.identity_rows_start + I.dof[c].dir_sub_of_set[i] + I.dir_start[c]
                                  : I.identity_rows_start + (i64)i + I.start[c];
                rhs[(size_t)r] = exact(I.dof[c].cpc[i].data());
            }
    return rhs;
}

// GetPlotInterpMatrix with AddIBCPlotInterpolation (not oriented),
// IBCMatrixManipulations.cpp:74-106 + PlotPointsCPDiff :254-416.
CSR ibc_plot_E(const PoissonIBCProblem& P) { return ibc_plot_E_at(P.tube, P.ibc, P.tube.surf.xp); }

// The same rows at arbitrary points xp (GetInterpMatrix(xp), problems/HeatWithIBC.cpp:117-128).
CSR ibc_plot_E_at(const Tube& t, const IBCData& I, const std::vector<Vec>& xp) {
    CSR A = build_E(t, xp);
    const size_t C = I.dof.size();
    std::vector<std::vector<uint8_t>> tags(C);
    for (size_t c = 0; c < C; ++c) {
        std::vector<V3> cpc, cpd;
        std::vector<i64> bdy, sub_of_set(xp.size(), -1), set_of_sub;
        for (size_t i = 0; i < xp.size(); ++i) {
            double cc[3], d;
            i64 b;
            I.surf[c].cp(xp[i].data(), cc, d, b);
            V3 diff = sub3(xp[i].data(), cc);
            if (norm3(diff) < 2.0 * t.r_t) {
                cpc.push_back({cc[0], cc[1], cc[2]});
                cpd.push_back(diff);
                bdy.push_back(b);
                sub_of_set[i] = (i64)set_of_sub.size();
                set_of_sub.push_back((i64)i);
            }
        }
        const size_t n = set_of_sub.size();
        for (size_t i = 0; i < n; ++i)
            if (norm3(cpd[i]) < I.meta[c].on_ibc_tolerance) {
                i64 ni = t.nearest_tube_index(xp[(size_t)set_of_sub[i]].data());
                i64 ts = I.tag[c].sub_of_set[(size_t)ni];
                if (ts >= 0) cpd[i] = I.tag[c].cp_diff[(size_t)ts];
            }
        std::vector<V3> dirs(n);
        if (I.meta[c].use_cp_diff_directions) {
            for (size_t i = 0; i < n; ++i) {
                if (bdy[i] > 0) {
                    double proj[3], a[3], b[3], d;
                    i64 bb;
                    for (int k = 0; k < 3; ++k) proj[k] = 2.0 * cpc[i][k] - xp[(size_t)set_of_sub[i]][k];
                    t.surf.cp(proj, a, d, bb);
                    I.surf[c].cp(proj, b, d, bb);
                    cpd[i] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                }
                dirs[i] = cpd[i];
            }
        } else {
            std::vector<Vec> q(n, Vec(3));
            for (size_t i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) q[i][k] = cpc[i][k];
            auto bn = interp_unoriented(t, I.tag[c].sub_of_set, (i64)I.tag[c].set_of_sub.size(), I.tag[c].binormals, q, true);
            for (size_t i = 0; i < n; ++i) {
                const double s = dot3(bn[i], cpd[i]);
                for (int k = 0; k < 3; ++k) dirs[i][k] = s * bn[i][k];
            }
        }
        tags[c] = tag_stencils(A, I, dirs, sub_of_set, c, 1.0);
    }
    std::vector<Trip> tr;
    move_to_ibc_columns(A, tags, I, tr);
    return CSR::from_triplets(A.rows, A.cols + I.total + I.num2, std::move(tr));
}

}  // namespace orc
