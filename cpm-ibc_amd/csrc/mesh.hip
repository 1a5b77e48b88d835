// Triangle-mesh closest points: host BVH builder and device traversal.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <string>

#include "mesh.h"
#include "ops.h"

namespace cpmgpu {

MeshDev MeshBvh::dev() const {
    MeshDev d;
    d.v = v.as<double>();
    d.f = f.as<int>();
    d.nodes = nodes.as<BvhNode>();
    d.prim = prim.as<int>();
    d.ntri = ntri;
    d.slack = slack;
    return d;
}

// ---------------------------------------------------------------- builder
namespace {

constexpr int kLeaf = 4;          // triangles per leaf
constexpr int kBins = 16;         // SAH bins per split
constexpr int kSahDepth = 32;     // deeper nodes split at the median, which bounds the depth

struct Box {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const double* a, const double* b) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], a[d]);
            hi[d] = std::max(hi[d], b[d]);
        }
    }
    void grow(const Box& o) { grow(o.lo, o.hi); }
    double area() const {
        const double e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
        return e0 < 0 ? 0.0 : 2.0 * (e0 * e1 + e1 * e2 + e2 * e0);
    }
};

// The hierarchy over primitives with boxes tb (triangles or segments);
// prim gets the primitive order of the leaves.  Returns the depth.
int build_nodes(const std::vector<Box>& tb, std::vector<int>& prim, std::vector<BvhNode>& nodes) {
    const int F = (int)tb.size();
    std::vector<double> cen((size_t)F * 3);
    for (int t = 0; t < F; ++t)
        for (int d = 0; d < 3; ++d) cen[(size_t)3 * t + d] = 0.5 * (tb[(size_t)t].lo[d] + tb[(size_t)t].hi[d]);
    prim.resize((size_t)F);
    std::iota(prim.begin(), prim.end(), 0);
    nodes.clear();
    nodes.reserve((size_t)2 * F);
    nodes.push_back(BvhNode{});
    struct Item { int node, start, end, depth; };
    std::vector<Item> work{{0, 0, F, 0}};
    int depth = 0;
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        depth = std::max(depth, it.depth);
        Box nb, cb;
        for (int k = it.start; k < it.end; ++k) {
            const int t = prim[(size_t)k];
            nb.grow(tb[(size_t)t]);
            cb.grow(&cen[(size_t)3 * t], &cen[(size_t)3 * t]);
        }
        BvhNode& nd = nodes[(size_t)it.node];
        for (int d = 0; d < 3; ++d) {
            nd.lo[d] = nb.lo[d];
            nd.hi[d] = nb.hi[d];
        }
        const int n = it.end - it.start;
        if (n <= kLeaf) {
            nd.first = it.start;
            nd.count = n;
            continue;
        }
        int axis = 0;
        for (int d = 1; d < 3; ++d)
            if (cb.hi[d] - cb.lo[d] > cb.hi[axis] - cb.lo[axis]) axis = d;
        const double lo = cb.lo[axis], ext = cb.hi[axis] - cb.lo[axis];
        int mid = -1;
        if (ext > 0 && it.depth < kSahDepth) {
            auto bin_of = [&](int t) {
                const int b = (int)(kBins * ((cen[(size_t)3 * t + axis] - lo) / ext));
                return std::min(kBins - 1, std::max(0, b));
            };
            int cnt[kBins] = {};
            Box bb[kBins];
            for (int k = it.start; k < it.end; ++k) {
                const int t = prim[(size_t)k];
                const int b = bin_of(t);
                ++cnt[b];
                bb[b].grow(tb[(size_t)t]);
            }
            double ra[kBins];
            int rn[kBins];
            Box acc;
            int an = 0;
            for (int b = kBins - 1; b > 0; --b) {   // ra[b], rn[b]: bins b..end
                acc.grow(bb[b]);
                an += cnt[b];
                ra[b] = acc.area();
                rn[b] = an;
            }
            Box lacc;
            int ln = 0, best_b = -1;
            double best_cost = INFINITY;
            for (int b = 0; b < kBins - 1; ++b) {   // split between bins b and b + 1
                lacc.grow(bb[b]);
                ln += cnt[b];
                if (ln == 0 || rn[b + 1] == 0) continue;
                const double cost = lacc.area() * ln + ra[b + 1] * rn[b + 1];
                if (cost < best_cost) { best_cost = cost; best_b = b; }
            }
            if (best_b >= 0) {
                auto pm = std::partition(prim.begin() + it.start, prim.begin() + it.end,
                                         [&](int t) { return bin_of(t) <= best_b; });
                mid = (int)(pm - prim.begin());
                if (mid == it.start || mid == it.end) mid = -1;
            }
        }
        if (mid < 0) {   // median split on the widest centroid axis
            mid = it.start + n / 2;
            std::nth_element(prim.begin() + it.start, prim.begin() + mid, prim.begin() + it.end, [&](int a, int b) {
                const double ca = cen[(size_t)3 * a + axis], cb2 = cen[(size_t)3 * b + axis];
                return ca < cb2 || (ca == cb2 && a < b);
            });
        }
        const int left = (int)nodes.size();
        nodes.push_back(BvhNode{});
        nodes.push_back(BvhNode{});
        nodes[(size_t)it.node].first = left;
        nodes[(size_t)it.node].second = left + 1;
        nodes[(size_t)it.node].count = 0;
        work.push_back({left + 1, mid, it.end, it.depth + 1});
        work.push_back({left, it.start, mid, it.depth + 1});
    }
    if (depth >= kBvhStack - 2) throw GeomError("closest-point hierarchy deeper than the traversal stack");
    return depth;
}

void upload(Buffer& b, const void* src, size_t bytes) {
    b = Buffer(bytes);
    HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
}

double coordinate_scale(const std::vector<double>& verts) {
    double scale = 0;
    for (double x : verts) {
        if (!std::isfinite(x)) throw ArgError("non-finite vertex coordinate");
        scale = std::max(scale, std::fabs(x));
    }
    return scale;
}

}  // namespace

static void check_mesh(const std::vector<double>& verts, const std::vector<int>& faces) {
    if (verts.size() % 3 || faces.size() % 3) throw ArgError("mesh arrays must hold 3 values per vertex / face");
    const int64_t nv = (int64_t)verts.size() / 3, nf = (int64_t)faces.size() / 3;
    if (nv < 3 || nf < 1) throw ArgError("mesh needs at least one triangle");
    if (nf > (1 << 28)) throw ArgError("too many triangles");
    for (int x : faces)
        if (x < 0 || x >= nv) throw ArgError("face index out of range");
}

void mesh_build(MeshBvh& m, const std::vector<double>& verts, const std::vector<int>& faces, hipStream_t stream) {
    check_mesh(verts, faces);
    const char* e = std::getenv("CPMGPU_BVH");
    const std::string mode = e ? e : "";
    const bool gpu = mode == "gpu" || (mode != "host" && (int64_t)faces.size() / 3 >= kGpuBvhTriangles);
    if (gpu && mesh_build_gpu(m, verts, faces, stream)) return;
    mesh_build_host(m, verts, faces);
}

void mesh_build_host(MeshBvh& m, const std::vector<double>& verts, const std::vector<int>& faces) {
    check_mesh(verts, faces);
    const int64_t nf = (int64_t)faces.size() / 3;
    const double scale = coordinate_scale(verts);
    const int F = (int)nf;
    std::vector<Box> tb((size_t)F);
    for (int t = 0; t < F; ++t)
        for (int k = 0; k < 3; ++k) {
            const double* p = &verts[(size_t)faces[(size_t)3 * t + k] * 3];
            tb[(size_t)t].grow(p, p);
        }
    std::vector<int> prim;
    std::vector<BvhNode> nodes;
    m.depth = build_nodes(tb, prim, nodes);
    upload(m.v, verts.data(), verts.size() * sizeof(double));
    upload(m.f, faces.data(), faces.size() * sizeof(int));
    upload(m.nodes, nodes.data(), nodes.size() * sizeof(BvhNode));
    upload(m.prim, prim.data(), prim.size() * sizeof(int));
    m.ntri = F;
    m.nnodes = (int)nodes.size();
    // A triangle's computed distance is within a few ulps of coordinate size
    // of the true one, which is at least the box distance.
    m.slack = 1e-12 * (1.0 + scale);
}

// ---------------------------------------------------------------- GPU builder
// A linear BVH (Karras, "Maximizing parallelism in the construction of BVHs,
// octrees, and k-d trees", HPG 2012): 30-bit Morton codes of the triangles'
// box centres in the vertices' bounding box, made unique by the triangle
// index in the low 32 bits and radix-sorted; internal node i (0 = root) covers
// the key range its common-prefix search finds and splits it where the prefix
// changes; leaves n - 1 + k hold one triangle each; boxes are filled bottom-up
// (the second child to arrive at a node forms its box).  Queries stay exact:
// they depend on the boxes being conservative, not on the tree's shape.
namespace {

__device__ __forceinline__ unsigned expand10(unsigned x) {   // bits 0..9 to every third bit
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ void tri_box(const double* v, const int* f, int t, double* lo, double* hi) {
    for (int d = 0; d < 3; ++d) {
        lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    for (int k = 0; k < 3; ++k) {
        const double* p = v + 3 * (int64_t)f[3 * (int64_t)t + k];
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
    }
}

__global__ void k_tri_keys(const double* v, const int* f, int F, double3 lo, double3 scale, unsigned long long* keys) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= F) return;
    double a[3], b[3];
    tri_box(v, f, t, a, b);
    const double c[3] = {0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]), 0.5 * (a[2] + b[2])};
    const double l[3] = {lo.x, lo.y, lo.z}, s[3] = {scale.x, scale.y, scale.z};
    unsigned q[3];
    for (int d = 0; d < 3; ++d) q[d] = (unsigned)fmin(1023.0, fmax(0.0, (c[d] - l[d]) * s[d]));
    const unsigned long long code = ((unsigned long long)expand10(q[0]) << 2) | ((unsigned long long)expand10(q[1]) << 1) |
                                    (unsigned long long)expand10(q[2]);
    keys[t] = (code << 32) | (unsigned)t;
}

// Length of the common prefix of keys i and j (-1 outside [0, n)); keys are unique.
__device__ __forceinline__ int prefix(const unsigned long long* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(k[i] ^ k[j]);
}

__global__ void k_lbvh_nodes(const unsigned long long* k, int n, BvhNode* nodes, int* parent, int* prim) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) prim[i] = (int)(k[i] & 0xffffffffull);
    if (i >= n - 1) return;
    const int d = prefix(k, n, i, i + 1) - prefix(k, n, i, i - 1) >= 0 ? 1 : -1;
    const int pmin = prefix(k, n, i, i - d);
    int lmax = 2;
    while (prefix(k, n, i, i + lmax * d) > pmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (prefix(k, n, i, i + (l + t) * d) > pmin) l += t;
    const int j = i + l * d;
    const int pnode = prefix(k, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (prefix(k, n, i, i + (s + t) * d) > pnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int left = min(i, j) == gamma ? n - 1 + gamma : gamma;
    const int right = max(i, j) == gamma + 1 ? n - 1 + gamma + 1 : gamma + 1;
    nodes[i].first = left;
    nodes[i].second = right;
    nodes[i].count = 0;
    parent[left] = i;
    parent[right] = i;
}

// Leaf boxes, then each node's box from its children's (the second child to arrive computes it).
__global__ void k_lbvh_boxes(const double* v, const int* f, const int* prim, int n, const int* parent, int* arrived,
                             BvhNode* nodes) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    int cur = n - 1 + k;
    BvhNode& leaf = nodes[cur];
    tri_box(v, f, prim[k], leaf.lo, leaf.hi);
    leaf.first = k;
    leaf.count = 1;
    leaf.second = -1;
    while (cur != 0) {
        const int p = parent[cur];
        if (p < 0) return;   // not reached: every node but the root has a parent
        __threadfence();   // this subtree's boxes before the arrival count (other XCDs read them)
        if (atomicAdd(&arrived[p], 1) == 0) return;
        __threadfence();   // the sibling's boxes after it
        BvhNode& np = nodes[p];
        const BvhNode& a = nodes[np.first];
        const BvhNode& b = nodes[np.second];
        for (int d = 0; d < 3; ++d) {
            np.lo[d] = fmin(a.lo[d], b.lo[d]);
            np.hi[d] = fmax(a.hi[d], b.hi[d]);
        }
        cur = p;
    }
}

__global__ void k_lbvh_depth(const int* parent, int n, int* depth) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    int cur = n - 1 + k, d = 0;
    while (cur > 0 && d <= kBvhStack) {
        cur = parent[cur];
        ++d;
    }
    if (cur < 0) d = kBvhStack + 1;   // not reached: a node without a parent rejects the tree
    atomicMax(depth, d);
}

}  // namespace

bool mesh_build_gpu(MeshBvh& m, const std::vector<double>& verts, const std::vector<int>& faces, hipStream_t st) {
    check_mesh(verts, faces);
    const double scale = coordinate_scale(verts);
    const int F = (int)(faces.size() / 3);
    if (F < 2) return false;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < verts.size(); i += 3)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], verts[i + d]);
            hi[d] = std::max(hi[d], verts[i + d]);
        }
    double sc[3];
    for (int d = 0; d < 3; ++d) sc[d] = hi[d] > lo[d] ? 1024.0 / (hi[d] - lo[d]) : 0.0;
    Buffer v, f, nodes, prim;
    upload(v, verts.data(), verts.size() * sizeof(double));
    upload(f, faces.data(), faces.size() * sizeof(int));
    const int NN = 2 * F - 1;
    nodes = Buffer(sizeof(BvhNode) * (size_t)NN);
    prim = Buffer(sizeof(int) * (size_t)F);
    Buffer keys(sizeof(unsigned long long) * (size_t)F * 2), parent(sizeof(int) * (size_t)NN),
        arrived(sizeof(int) * (size_t)F + 64);
    auto* k_in = keys.as<unsigned long long>();
    auto* k_out = k_in + F;
    int* depth = arrived.as<int>() + F;
    const unsigned nb = (unsigned)((F + 255) / 256);
    hipLaunchKernelGGL(k_tri_keys, dim3(nb), dim3(256), 0, st, v.as<double>(), f.as<int>(), F,
                       make_double3(lo[0], lo[1], lo[2]), make_double3(sc[0], sc[1], sc[2]), k_in);
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, k_in, k_out, F, 0, 62, st));
    Buffer tmp(std::max<size_t>(tmp_bytes, 16));
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, k_in, k_out, F, 0, 62, st));
    HIPCHK(hipMemsetAsync(arrived.p, 0, arrived.bytes, st));
    HIPCHK(hipMemsetAsync(parent.p, 0xff, parent.bytes, st));
    hipLaunchKernelGGL(k_lbvh_nodes, dim3(nb), dim3(256), 0, st, k_out, F, nodes.as<BvhNode>(), parent.as<int>(),
                       prim.as<int>());
    hipLaunchKernelGGL(k_lbvh_boxes, dim3(nb), dim3(256), 0, st, v.as<double>(), f.as<int>(), prim.as<int>(), F,
                       parent.as<int>(), arrived.as<int>(), nodes.as<BvhNode>());
    hipLaunchKernelGGL(k_lbvh_depth, dim3(nb), dim3(256), 0, st, parent.as<int>(), F, depth);
    HIPCHK(hipGetLastError());
    int hd = 0;
    HIPCHK(hipMemcpyAsync(&hd, depth, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hd >= kBvhStack - 2) return false;
    m.v = std::move(v);
    m.f = std::move(f);
    m.nodes = std::move(nodes);
    m.prim = std::move(prim);
    m.ntri = F;
    m.nnodes = NN;
    m.depth = hd;
    m.slack = 1e-12 * (1.0 + scale);
    return true;
}

void segment_build(MeshBvh& m, const std::vector<double>& verts) {
    if (verts.size() % 3) throw ArgError("polyline vertices must hold 3 values each");
    const int64_t nv = (int64_t)verts.size() / 3;
    if (nv < 2) throw ArgError("polyline needs >= 2 vertices");
    if (nv > (1 << 28)) throw ArgError("too many polyline vertices");
    const double scale = coordinate_scale(verts);
    const int F = (int)(nv - 1);   // segment k joins vertices k and k + 1
    std::vector<Box> tb((size_t)F);
    for (int t = 0; t < F; ++t) {
        const double* a = &verts[(size_t)3 * t];
        tb[(size_t)t].grow(a, a);
        tb[(size_t)t].grow(a + 3, a + 3);
    }
    std::vector<int> prim;
    std::vector<BvhNode> nodes;
    m.depth = build_nodes(tb, prim, nodes);
    upload(m.v, verts.data(), verts.size() * sizeof(double));
    m.f = Buffer();
    upload(m.nodes, nodes.data(), nodes.size() * sizeof(BvhNode));
    upload(m.prim, prim.data(), prim.size() * sizeof(int));
    m.ntri = F;
    m.nnodes = (int)nodes.size();
    m.slack = 1e-12 * (1.0 + scale);
}

// ---------------------------------------------------------------- queries
#pragma clang fp contract(off)

__device__ __forceinline__ double dot3(const double* a, const double* b) {
    double s = 0.0;
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}

// Closest point on triangle abc to p: Ericson, "Real-Time Collision
// Detection", section 5.1.5 (Voronoi regions of the vertices, edges, face).
__device__ void cp_triangle(const double* p, const double* a, const double* b, const double* c, double* out) {
    double ab[3], ac[3], ap[3];
    for (int d = 0; d < 3; ++d) { ab[d] = b[d] - a[d]; ac[d] = c[d] - a[d]; ap[d] = p[d] - a[d]; }
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0 && d2 <= 0) { for (int d = 0; d < 3; ++d) out[d] = a[d]; return; }
    double bp[3];
    for (int d = 0; d < 3; ++d) bp[d] = p[d] - b[d];
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) { for (int d = 0; d < 3; ++d) out[d] = b[d]; return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        for (int d = 0; d < 3; ++d) out[d] = a[d] + v * ab[d];
        return;
    }
    double cpv[3];
    for (int d = 0; d < 3; ++d) cpv[d] = p[d] - c[d];
    const double d5 = dot3(ab, cpv), d6 = dot3(ac, cpv);
    if (d6 >= 0 && d5 <= d6) { for (int d = 0; d < 3; ++d) out[d] = c[d]; return; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        for (int d = 0; d < 3; ++d) out[d] = a[d] + w * ac[d];
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int d = 0; d < 3; ++d) out[d] = b[d] + w * (c[d] - b[d]);
        return;
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    for (int d = 0; d < 3; ++d) out[d] = a[d] + ab[d] * v + ac[d] * w;
}

__device__ __forceinline__ double dist3(const double* a, const double* b) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s += (a[d] - b[d]) * (a[d] - b[d]);
    return sqrt(s);
}

// Closest point on segment ab to x (the projection clamped to the segment)
// and its distance, each sum in the order of the oracle's restatement.
__device__ double cp_segment(const double* x, const double* a, const double* b, double* q) {
    double ab[3], ax[3];
    for (int d = 0; d < 3; ++d) { ab[d] = b[d] - a[d]; ax[d] = x[d] - a[d]; }
    const double L2 = dot3(ab, ab);
    double t = L2 > 0 ? dot3(ax, ab) / L2 : 0.0;
    t = fmin(1.0, fmax(0.0, t));
    for (int d = 0; d < 3; ++d) q[d] = a[d] + t * ab[d];
    return dist3(x, q);
}

#pragma clang fp contract(on)

// Squared distance from x to the node's box.
__device__ __forceinline__ double box_d2(const BvhNode* nd, const double* x) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double e = fmax(fmax(nd->lo[d] - x[d], x[d] - nd->hi[d]), 0.0);
        s += e * e;
    }
    return s;
}

// Depth-first traversal, nearer child first.  A box is skipped only when it is
// farther than the best distance plus the slack, so every primitive whose
// computed distance could tie or beat the best is visited; ties go to the
// lower primitive index.  SEG: polyline segments (vertices t, t + 1), else
// triangles.
template <bool SEG>
__device__ void cp_bvh(const MeshDev& m, const double* x, double* c, double& dist) {
    int stack[kBvhStack];
    int sp = 0;
    double best = INFINITY;
    int bt = 0x7fffffff;
    c[0] = c[1] = c[2] = NAN;
    int node = 0;
    for (;;) {
        const BvhNode* nd = m.nodes + node;
        if (nd->count > 0) {
            for (int k = 0; k < nd->count; ++k) {
                const int t = m.prim[nd->first + k];
                double q[3], dd;
                if (SEG) {
                    dd = cp_segment(x, m.v + 3 * (int64_t)t, m.v + 3 * (int64_t)t + 3, q);
                } else {
                    const int* fv = m.f + 3 * (int64_t)t;
                    cp_triangle(x, m.v + 3 * (int64_t)fv[0], m.v + 3 * (int64_t)fv[1], m.v + 3 * (int64_t)fv[2], q);
                    dd = dist3(x, q);
                }
                if (dd < best || (dd == best && t < bt)) {
                    best = dd;
                    bt = t;
                    c[0] = q[0]; c[1] = q[1]; c[2] = q[2];
                }
            }
        } else {
            const double lim = (best + m.slack) * (best + m.slack);
            const int l = nd->first, r = nd->second;
            const double dl = box_d2(m.nodes + l, x), dr = box_d2(m.nodes + r, x);
            const bool vl = dl <= lim, vr = dr <= lim;
            if (vl && vr) {
                if (sp == kBvhStack) { dist = NAN; return; }   // unreachable: the builders bound the depth
                const bool lfirst = dl <= dr;
                stack[sp++] = lfirst ? r : l;
                node = lfirst ? l : r;
                continue;
            }
            if (vl || vr) {
                node = vl ? l : r;
                continue;
            }
        }
        const double lim = (best + m.slack) * (best + m.slack);
        bool more = false;
        while (sp > 0) {
            const int cand = stack[--sp];
            if (box_d2(m.nodes + cand, x) <= lim) { node = cand; more = true; break; }
        }
        if (!more) break;
    }
    dist = best;
}

__device__ void cp_mesh(const MeshDev& m, const double* x, double* c, double& dist) { cp_bvh<false>(m, x, c, dist); }
__device__ void cp_polyline(const MeshDev& m, const double* x, double* c, double& dist) { cp_bvh<true>(m, x, c, dist); }

}  // namespace cpmgpu
