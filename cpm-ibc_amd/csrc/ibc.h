// Interior boundary conditions on the GPU: tagging / DOF subsets, stencil
// directions and the IBC-modified operator (PoissonWithIBC).
#pragma once
#include <vector>

#include "ops.h"
#include "tube.h"

namespace cpmgpu {

struct IbcSpec {
    SurfaceHost surf;
    int order = 1;              // IBCMetaData::boundary_order
    int btype[2] = {0, 0};      // IBCMetaData::boundary_type (0 Dirichlet, 1 Neumann)
    double tol = 0;             // IBCMetaData::on_ibc_tolerance
    double dof_scale = 1.0;     // IBCMetaData::ibc_DOFSubset_radius_scale
};

// Device tables of one IBC.
struct IbcData {
    IbcSpec spec;
    Buffer verts;
    MeshBvh bvh;                // polylines: hierarchy over the segments (and the vertices)
    SurfDev sd{};
    int64_t nt = 0, nd = 0;     // tagging / DOF subset sizes
    Buffer t_sub;               // int32 [N]: tagging index of a tube node or -1
    Buffer t_set, t_bdy;        // int32 [nt]
    Buffer t_cpc, t_cpdiff, t_dir, t_dist;   // double [nt*3] (dist [nt])
    Buffer t_bn;                // double [nt*3]: binormal t x n (0 for point curves)
    Buffer d_sub;               // int32 [N]: DOF index of a tube node or -1
    Buffer d_set;               // int32 [nd]
    Buffer d_cpc, d_dir;        // double [nd*3]
    int64_t start = 0, dir_start = 0, ndir = 0;
    bool second_dirichlet() const { return (spec.btype[0] == 0 || spec.btype[1] == 0) && spec.order == 2; }
};

// Device view used by the operator-assembly kernels.
struct IbcDev {
    const int32_t* t_sub;
    const int32_t* t_bdy;
    const double* t_dir;
    const int32_t* d_sub;
    const double* d_dir;
    int64_t start;
};

struct Problem {
    Ctx* ctx = nullptr;
    Tube* tube = nullptr;
    std::vector<IbcData> ibc;
    int64_t N = 0, G = 0, K = 0, D = 0, identity_rows_start = 0;
    int ew = 0;
    Buffer normals;             // whole-tube unoriented normals [N*3]
    Buffer ecol, eval64, eval32, lcol, lval64, lval32, ldiag64, ldiag32, flag;
    Buffer rhs_pts, rhs_kind;
    Buffer qpts;                // [3 * D] query point of each E row (NaN: identity rows)
    double build_ms = 0;
};

// bctx: the context whose stream and workspaces the build uses (the tube's when null; same device).
Problem* problem_build(Tube* tube, const std::vector<IbcSpec>& specs, Ctx* bctx = nullptr);
Op* op_from_problem(Problem* P, double m, double n, int precond, int precond_iters);
// Plotting rows at nq host points: tube interpolation modified by the curves
// (AddIBCPlotInterpolation), nq x D CSR with (p+1)^3 entries per row.
void problem_plot_matrix(Problem* P, const double* xq, int64_t nq, int32_t* rowptr, int32_t* col, double* val);
// Surface::IBCParameter of n host points (3 each) on curve c: normalised arc
// length along the curve (vparam null), or the parameter vparam[i] of the
// nearest of the curve's parameterisation points.
void problem_ibc_parameter(Problem* P, int c, const double* pts, int64_t n, const double* vparam, double* out);
// GeodesicDistance::ComputeDistance up to its plotting interpolation: phi on
// the D rows into d_phi; infos[0] the heat solve, infos[1..3] the Poisson solves.
void geodesic_distance(Problem* P, double dt, int sweeps, const cpm_solve_opts& o, double* d_phi,
                       cpm_solve_info* infos);

}  // namespace cpmgpu
