// Triangle-mesh and polyline closest points on the GPU: a bounding-volume
// hierarchy built on the host (binned surface-area heuristic) and a per-query
// FP64 traversal.
// The reference queries fcpw's SAH BVH (src/Surface.cpp:1751-1779,2537-2592,
// FP32 inside fcpw); here every distance is FP64 and the result is the exact
// minimiser of the computed distance, lowest triangle index on ties, i.e. the
// brute-force answer.
#pragma once
#include <vector>

#include "engine.h"

namespace cpmgpu {

// Binary BVH node, 64 bytes.  count > 0: leaf holding prim[first .. first+count);
// count == 0: internal node whose children are nodes first and second.
struct BvhNode {
    double lo[3], hi[3];
    int first;
    int count;
    int second;
    int pad;
};

struct MeshDev {
    const double* v = nullptr;      // vertices, 3 per vertex
    const int* f = nullptr;         // triangles, 3 vertex indices each
    const BvhNode* nodes = nullptr;
    const int* prim = nullptr;      // triangle indices in leaf order
    int ntri = 0;
    double slack = 0;               // distance slack of the box test (rounding of the triangle distances)
};

constexpr int kBvhStack = 64;       // traversal stack; the builder keeps the depth below it

struct MeshBvh {
    Buffer v, f, nodes, prim;
    int ntri = 0, nnodes = 0, depth = 0;
    double slack = 0;
    MeshDev dev() const;
};

// Validates the mesh, builds the hierarchy and uploads everything: on the GPU
// (linear BVH from Morton codes, mesh_build_gpu) for meshes of at least
// kGpuBvhTriangles triangles, on the host (binned SAH) for smaller ones or
// when the GPU tree is deeper than the traversal stack.  CPMGPU_BVH=host /
// gpu forces one builder.
void mesh_build(MeshBvh& m, const std::vector<double>& verts, const std::vector<int>& faces,
                hipStream_t stream = nullptr);
constexpr int kGpuBvhTriangles = 16384;
// The host builder alone.
void mesh_build_host(MeshBvh& m, const std::vector<double>& verts, const std::vector<int>& faces);
// The GPU builder alone; false (m untouched) if its tree is too deep.
bool mesh_build_gpu(MeshBvh& m, const std::vector<double>& verts, const std::vector<int>& faces, hipStream_t stream);
// The same over the segments of a polyline (vertices k, k + 1; f stays empty).
void segment_build(MeshBvh& m, const std::vector<double>& verts);

// Closest point of x on the mesh and its distance.
__device__ void cp_mesh(const MeshDev& m, const double* x, double* c, double& dist);
// Closest point of x on the polyline (segment projections, the oracle's
// formula; lowest segment index among equal distances) and its distance.
__device__ void cp_polyline(const MeshDev& m, const double* x, double* c, double& dist);

}  // namespace cpmgpu
