// Tube construction on the GPU: closest points, banded index set in the
// reference's breadth-first order, neighbours, base points, E and L rows.
#pragma once
#include <string>
#include <vector>

#include "engine.h"
#include "mesh.h"
#include "ops.h"

namespace cpmgpu {

enum SurfType : int {
    SURF_SPHERE = 1, SURF_HEMISPHERE = 2, SURF_TORUS = 3, SURF_CIRCLE3D = 4, SURF_POINT = 5,
    SURF_POLYLINE = 6, SURF_ELLIPSOID = 7, SURF_TRIANGULATION = 8, SURF_ARC = 9
};

// Device description of a surface (src/Surface.cpp:61-190 dispatch, resolved
// once on the host instead of per query by string compare).
struct SurfDev {
    int type;
    int dim;
    int is_open;
    int nseg;                 // polyline segments
    double P[16];             // parameters (SurfaceSpecifier order)
    const double* verts;      // polyline vertices, 3 per vertex
    double ep0[3], ep1[3];    // open-polyline end points (boundary tags)
    MeshDev mesh;             // triangulations
};

struct SurfaceHost {
    std::string type;
    std::vector<double> params;
    double bbox[2] = {0, 0};
    int dim = 3;
    std::vector<double> verts;   // polyline / triangulation vertices (x,y,z)...
    std::vector<int> faces;      // triangulation: 3 vertex indices per triangle
    bool is_open = false;
};

SurfaceHost surface_from_desc(const char* type, const double* params, int nparams, double bb0, double bb1, int dim,
                              const double* verts, int64_t nverts);
// "Triangulation" (SurfaceSpecifier.cpp:298-363): bbox_lo >= bbox_hi selects
// the reference's box [min(0, coordinates), max(0, coordinates)].
SurfaceHost surface_mesh(const double* verts, int64_t nverts, const int32_t* faces, int64_t nfaces, double bb0,
                         double bb1);

struct GridDev {
    int64_t N1d;
    double xstart, dx, r_t;
};

struct Tube {
    Ctx* ctx = nullptr;
    SurfaceHost surf;
    SurfDev sd{};
    Buffer surf_verts;
    MeshBvh mesh;       // triangulations: hierarchy over the triangles
    int dim = 3, p = 3;
    double dx = 0, xstart = 0, rfac = 0, r_t = 0;
    int64_t n = 0, N1d = 0, N = 0;
    Buffer gidx;        // int32 [N1d^3]: tube index of a grid node, or a sentinel
    Buffer rec;         // int64 [N]  linear grid index of node i
    Buffer cpx;         // double [N*3]
    Buffer dist;        // double [N]
    Buffer bdy;         // int32 [N]
    Buffer nbr;         // int32 [N*6], -1 = none
    Buffer bpt;         // int32 [N]
    int64_t levels = 0;
    double build_ms = 0;
    Buffer lex, rlex;   // int32 [N]: lexicographic position -> tube index and back (built on demand)
};

Tube* tube_build(Ctx* ctx, const SurfaceHost& s, int p, double dx);

// Shared kernels (defined in tube.hip).
constexpr int32_t kUnvisited = 0x7fffffff;
constexpr int32_t kOutside = 0x7ffffffe;
constexpr int32_t kCand = 0x40000000;   // candidate key base; indices are < kCand
__global__ void k_fill_i32(int32_t* a, int64_t n, int32_t v);
__global__ void k_count(const int32_t* slot, int64_t ns, int* bcount);
__global__ void k_scan(int* bcount, int64_t nb, int* total);
__device__ void cp_surface(const SurfDev& s, const double* x, double* c, double& dist, int& bdy);
__device__ void weights_1d(const double* xs, double xq, int S, double* w);
// Base point b (grid coordinates) and 1-D barycentric weights of the degree-p
// stencil at xq (Interpolation.cpp:52-121); false when it leaves the grid.
__device__ bool stencil_weights(const GridDev& g, int p, const double* xq, int64_t* b, double (*w)[5]);
SurfDev make_surfdev(const SurfaceHost& s, const double* dverts, const MeshDev* mesh = nullptr);

// E rows at query points xq (device, 3 doubles each) as slot-major ELL of width
// (p+1)^dim; columns are tube indices.  Throws GeomError when a stencil leaves
// the tube (the reference throws in Tube::ComputeBasePoints).
// wctx: the context whose workspace the launch uses (the tube's when null).
void build_E_ell(const Tube& t, const double* d_xq, int64_t nq, int* ecol, double* eval, float* eval32,
                 int64_t ld, hipStream_t st, Ctx* wctx = nullptr);

// Operator (m I + n (Ld + Loff E)) of the plain closest-point problem on a GPU tube.
Op* op_from_tube(Tube* t, double m, double n, int precond, int precond_iters);

// Reference-order operator rows handed to the tensor-product format builder.
struct TpSource {
    int64_t D = 0, N = 0;                     // rows; the first N are tube nodes
    const int* ecol = nullptr;                // E as slot-major ELL, ld = D, tube-order columns
    const double* eval = nullptr;
    int ew = 0;
    const int* lcol = nullptr;                // L off-diagonal ELL, ld = D
    const double* lval = nullptr;
    int lw = 0;
    const double* ldiag = nullptr;
    const uint8_t* flag = nullptr;
    const double* qpts = nullptr;             // [3 * D] query point of each non-identity E row
};

// Give op the tensor-product execution format (rows reordered lexicographically
// on the tube).  Returns false, leaving op untouched, when some row is not a
// tensor-product stencil with uniform L values (the caller keeps ELL).
bool op_build_tp(Op& op, Tube& t, const TpSource& src);
// CPMGPU_OP_FORMAT=ell keeps tube-built operators in the ELL format.
bool tp_disabled();

}  // namespace cpmgpu
