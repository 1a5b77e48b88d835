// Interior-boundary-condition part of the CPU oracle (3-D embeddings).
// TEST INFRASTRUCTURE ONLY (see orc.h).
#pragma once
#include <array>

#include "orc.h"

namespace orc {

// IBCSubset (src/IBCSubset.h:18-115): tube nodes within 2 r_t of the curve, in
// breadth-first discovery order.
struct TagSubset {
    std::vector<i64> set_of_sub, sub_of_set;
    std::vector<std::array<double, 3>> cpc, cp_diff, normals, tangents, binormals, directions;
    Vec dist;
    std::vector<i64> bdy;
};

// IBCDOFSubset (src/IBCDOFSubset.h:18-136): tagging nodes within r_t.
struct DofSubset {
    std::vector<i64> set_of_sub, sub_of_set, dir_set_of_sub, dir_sub_of_set, which_side;
    std::vector<std::array<double, 3>> cpc, cp_diff, cpbar, directions;
    Vec dist;
    std::vector<i64> bdy;
    CSR E;   // extension rows at cp_c (first order) or cp_c_bar (second order)
};

// IBCSubsets (src/IBCSubsets.h:18-106).
struct IBCData {
    std::vector<Surface> surf;
    std::vector<IBCMeta> meta;
    std::vector<TagSubset> tag;
    std::vector<DofSubset> dof;
    std::vector<i64> start, dir_start;
    i64 total = 0, identity_rows_start = 0, num2 = 0;
};

struct PoissonIBCProblem {
    Tube tube;
    IBCData ibc;
    CSR L, E, Lfull, Efull;
    std::vector<uint8_t> ident;
};

void sym_eig3(const double A[3][3], double w[3], double V[3][3]);
std::vector<std::array<double, 3>> tube_jcp_normals(const Tube& t);
IBCData make_ibc(const Tube& t, const std::vector<SurfaceSpec>& specs);
CSR add_ibc_extension(const CSR& E, const Tube& t, const IBCData& I);
CSR add_ibc_finite_difference(const CSR& L, const Tube& t, const IBCData& I);
PoissonIBCProblem make_poisson_ibc(const SurfaceSpec& s, const std::vector<SurfaceSpec>& ibc_specs, double dx, int p = 3);
Vec poisson_ibc_rhs(const PoissonIBCProblem& P, const std::function<double(const double*)>& f,
                    const std::function<double(const double*)>& exact);
CSR ibc_plot_E(const PoissonIBCProblem& P);
CSR ibc_plot_E_at(const Tube& t, const IBCData& I, const std::vector<Vec>& xp);

}  // namespace orc
