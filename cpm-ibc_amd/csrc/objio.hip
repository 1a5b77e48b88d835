// Wavefront OBJ input of meshes and curves, as the reference reads them:
// SimplePolygonMesh::readMeshFromObjFile (src/SimplePolygonMesh.cpp:139-213)
// and the surface kind SurfaceSpecifier derives from the faces
// (src/SurfaceSpecifier.cpp:298-363).  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpm_gpu.h"
#include "ops.h"
#include "tube.h"

using namespace cpmgpu;

namespace cpmgpu {
int status_of_current_exception();
}

namespace {

struct ObjData {
    std::vector<double> v;     // 3 per vertex
    std::vector<int32_t> f;    // face_size per face
    int face_size = 0;
};

// First integer of an index token "p", "p/t", "p//n" or "p/t/n" (1-based).
int64_t obj_index(const std::string& tok, const std::string& path, int64_t lineno) {
    const std::string head = tok.substr(0, tok.find('/'));
    char* end = nullptr;
    const long long k = std::strtoll(head.c_str(), &end, 10);
    if (head.empty() || *end != '\0')
        throw ArgError(path + ":" + std::to_string(lineno) + ": bad index '" + tok + "'");
    if (k <= 0)   // the reference has no meaning for relative (negative) indices either
        throw ArgError(path + ":" + std::to_string(lineno) + ": index " + head + " is not a positive 1-based index");
    return k - 1;
}

ObjData obj_parse(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ArgError(std::string("couldn't open file ") + path);
    ObjData d;
    std::string line, tok;
    int64_t lineno = 0;
    auto add_face = [&](const std::vector<int64_t>& idx) {
        const int fs = (int)idx.size();
        if (d.face_size && fs != d.face_size)
            throw ArgError(std::string(path) + ":" + std::to_string(lineno) +
                           ": a file must hold only triangles or only polyline edges");
        d.face_size = fs;
        for (int64_t k : idx) d.f.push_back((int32_t)k);
    };
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();   // DOS line ends
        std::istringstream ss(line);
        if (!(ss >> tok)) continue;
        if (tok == "v") {
            std::vector<double> pos;
            double x;
            while (ss >> x) pos.push_back(x);
            if (pos.size() != 3)
                throw ArgError(std::string(path) + ":" + std::to_string(lineno) +
                               ": vertices must have 3 coordinates (the GPU tube is 3-D)");
            d.v.insert(d.v.end(), pos.begin(), pos.end());
        } else if (tok == "f") {
            std::vector<int64_t> idx;
            while (ss >> tok) idx.push_back(obj_index(tok, path, lineno));
            if (idx.size() != 3)
                throw ArgError(std::string(path) + ":" + std::to_string(lineno) + ": faces must be triangles");
            add_face(idx);
        } else if (tok == "l") {   // a polyline of n vertices is n - 1 edges
            std::vector<int64_t> idx;
            while (ss >> tok) idx.push_back(obj_index(tok, path, lineno));
            if (idx.size() < 2)
                throw ArgError(std::string(path) + ":" + std::to_string(lineno) + ": a polyline needs 2 vertices");
            for (size_t k = 0; k + 1 < idx.size(); ++k) add_face({idx[k], idx[k + 1]});
        }   // vt, vn, comments, groups, materials: ignored
    }
    const int64_t nv = (int64_t)d.v.size() / 3;
    if (nv == 0 || d.f.empty()) throw ArgError(std::string(path) + ": no vertices or no faces/edges");
    for (int32_t k : d.f)
        if (k >= nv) throw ArgError(std::string(path) + ": face index " + std::to_string(k + 1) + " > vertex count");
    return d;
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return CPM_OK;
    } catch (...) {
        return status_of_current_exception();
    }
}

// Consecutive edges -> vertex chain; a closed curve repeats its first vertex.
std::vector<double> edge_chain(const double* v, int64_t nv, const int32_t* e, int64_t ne) {
    if (ne < 1) throw ArgError("a polyline needs at least one edge");
    for (int64_t k = 0; k < 2 * ne; ++k)
        if (e[k] < 0 || e[k] >= nv) throw ArgError("edge index out of range");
    for (int64_t k = 1; k < ne; ++k)
        if (e[2 * k] != e[2 * k - 1])
            throw ArgError("polyline edges must be consecutive (edge " + std::to_string(k) +
                           " does not start where the previous one ends)");
    std::vector<double> c;
    auto push = [&](int32_t i) { c.insert(c.end(), v + 3 * (int64_t)i, v + 3 * (int64_t)i + 3); };
    push(e[0]);
    for (int64_t k = 0; k < ne; ++k) push(e[2 * k + 1]);
    // SurfaceSpecifier.cpp:334-341: closed when the first and last vertices of
    // the file coincide or the last edge touches vertex 0
    const bool closed = std::memcmp(v, v + 3 * (nv - 1), 3 * sizeof(double)) == 0 || e[2 * ne - 2] == 0 ||
                        e[2 * ne - 1] == 0;
    const size_t n = c.size() / 3;
    const bool ends_equal = std::memcmp(c.data(), c.data() + 3 * (n - 1), 3 * sizeof(double)) == 0;
    if (closed && !ends_equal) push(e[0]);
    return c;
}

}  // namespace

extern "C" {

int cpm_obj_read(const char* path, double* verts, int64_t* nverts, int32_t* faces, int64_t* nfaces, int* face_size) {
    return guard([&] {
        if (!path || !nverts || !nfaces || !face_size) throw ArgError("null argument");
        const ObjData d = obj_parse(path);
        const int64_t nv = (int64_t)d.v.size() / 3, nf = (int64_t)d.f.size() / d.face_size;
        if (verts || faces) {
            if (*nverts < nv || *nfaces < nf) throw ArgError("output arrays smaller than the file's counts");
            if (verts) std::copy(d.v.begin(), d.v.end(), verts);
            if (faces) std::copy(d.f.begin(), d.f.end(), faces);
        }
        *nverts = nv;
        *nfaces = nf;
        *face_size = d.face_size;
    });
}

int cpm_polyline_chain(const double* verts, int64_t nverts, const int32_t* edges, int64_t nedges, double* chain,
                       int64_t* nchain) {
    return guard([&] {
        if (!verts || !edges || !nchain) throw ArgError("null argument");
        const std::vector<double> c = edge_chain(verts, nverts, edges, nedges);
        const int64_t n = (int64_t)c.size() / 3;
        if (chain) {
            if (*nchain < n) throw ArgError("chain array smaller than the vertex count");
            std::copy(c.begin(), c.end(), chain);
        }
        *nchain = n;
    });
}

int cpm_tube_create_obj(cpm_ctx* ctx, const char* path, double bbox_lo, double bbox_hi, int p, double dx,
                        cpm_tube** out) {
    return guard([&] {
        if (!ctx || !path || !out) throw ArgError("null argument");
        const ObjData d = obj_parse(path);
        const int64_t nv = (int64_t)d.v.size() / 3;
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        HIPCHK(hipSetDevice(c->device));
        SurfaceHost s;
        if (d.face_size == 3) {
            s = surface_mesh(d.v.data(), nv, d.f.data(), (int64_t)d.f.size() / 3, bbox_lo, bbox_hi);
        } else {
            if (!(bbox_lo < bbox_hi)) {   // SurfaceSpecifier::SetSurface() (:346-363)
                bbox_lo = 0.0;
                bbox_hi = 0.0;
                for (double x : d.v) {
                    bbox_lo = std::min(bbox_lo, x);
                    bbox_hi = std::max(bbox_hi, x);
                }
            }
            const std::vector<double> ch = edge_chain(d.v.data(), nv, d.f.data(), (int64_t)d.f.size() / 2);
            s = surface_from_desc("Polyline", nullptr, 0, bbox_lo, bbox_hi, 3, ch.data(), (int64_t)ch.size() / 3);
        }
        *out = reinterpret_cast<cpm_tube*>(tube_build(c, s, p, dx));
    });
}

}  // extern "C"
