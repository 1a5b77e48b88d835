// C++ host side of the MI355X engine: the reference's cpm::Solver interface
// (src/Solver.h:19-53) with identical member signatures, forwarding to the C
// ABI of libcpmgpu.so (include/cpm_gpu.h).  Eigen is not used: SpMat is a
// row-major CSR with int32 indices (the layout of Eigen::SparseMatrix<double,
// RowMajor>, which the reference's custom solver uses, src/Scalar.h:30-31) and
// VectorX is a std::vector<double>.
//
// Differences a caller can see:
//  * failures throw cpm::Error (the reference reports none: Solver::Init returns
//    void and Solve ignores non-convergence, src/Solver.cpp:69-114);
//  * LastInfo() returns what the reference drops: status, iterations, restarts,
//    relative residual;
//  * SetPrecision selects FP64 (the reference's arithmetic, the default) or FP32
//    Krylov iterations refined in FP64.
#pragma once
#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cpm_gpu.h"

namespace cpm {

typedef double Scalar;
typedef std::vector<Scalar> VectorX;

// Preconditioner, src/Defines.h:13-18.
enum Preconditioner { identity = CPM_PRECOND_IDENTITY, diagonal = CPM_PRECOND_DIAGONAL, jacobi = CPM_PRECOND_JACOBI };

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(int status, const char* what) {
    if (status != CPM_OK) throw Error(status, std::string(what) + ": " + cpm_last_error());
}

struct Triplet {   // Eigen::Triplet<Scalar>
    int64_t row, col;
    Scalar value;
};

// Row-major CSR, int32 indices, ascending columns; duplicates summed
// (Eigen::SparseMatrix::setFromTriplets semantics).
class SpMat {
  public:
    SpMat() = default;
    SpMat(int64_t rows, int64_t cols) : rows_(rows), cols_(cols), ptr_((size_t)rows + 1, 0) {}
    int64_t rows() const { return rows_; }
    int64_t cols() const { return cols_; }
    int64_t nonZeros() const { return (int64_t)val_.size(); }
    const int32_t* outerIndexPtr() const { return ptr_.data(); }
    const int32_t* innerIndexPtr() const { return col_.data(); }
    const Scalar* valuePtr() const { return val_.data(); }
    template <typename It>
    void setFromTriplets(It first, It last) {
        std::vector<int64_t> cnt((size_t)rows_ + 1, 0);
        for (It t = first; t != last; ++t) ++cnt[(size_t)t->row + 1];
        for (int64_t r = 0; r < rows_; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
        std::vector<std::pair<int64_t, Scalar>> ent((size_t)cnt[(size_t)rows_]);
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (It t = first; t != last; ++t) ent[(size_t)fill[(size_t)t->row]++] = {t->col, t->value};
        ptr_.assign((size_t)rows_ + 1, 0);
        col_.clear();
        val_.clear();
        for (int64_t r = 0; r < rows_; ++r) {
            auto b = ent.begin() + cnt[(size_t)r], e = ent.begin() + cnt[(size_t)r + 1];
            std::stable_sort(b, e, [](const std::pair<int64_t, Scalar>& a, const std::pair<int64_t, Scalar>& c) {
                return a.first < c.first;
            });
            for (auto it = b; it != e; ++it) {
                if (!col_.empty() && (int64_t)col_.size() > ptr_[(size_t)r] && col_.back() == it->first) {
                    val_.back() += it->second;
                } else {
                    col_.push_back((int32_t)it->first);
                    val_.push_back(it->second);
                }
            }
            ptr_[(size_t)r + 1] = (int32_t)col_.size();
        }
    }
    // Adopt CSR arrays (row pointer of rows+1 entries).
    void assign(int64_t rows, int64_t cols, std::vector<int32_t> ptr, std::vector<int32_t> col, std::vector<Scalar> val) {
        rows_ = rows;
        cols_ = cols;
        ptr_ = std::move(ptr);
        col_ = std::move(col);
        val_ = std::move(val);
    }

  private:
    int64_t rows_ = 0, cols_ = 0;
    std::vector<int32_t> ptr_{0}, col_;
    std::vector<Scalar> val_;
};

// One HIP context (device + stream) shared by the objects of a thread.
class Context {
  public:
    explicit Context(int device = 0) { check(cpm_ctx_create(device, &ctx_), "cpm_ctx_create"); }
    ~Context() { cpm_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cpm_ctx* get() const { return ctx_; }
    static std::shared_ptr<Context> shared(int device = 0) {
        static thread_local std::shared_ptr<Context> c;
        if (!c) c = std::make_shared<Context>(device);
        return c;
    }

  private:
    cpm_ctx* ctx_ = nullptr;
};

// cpm::Solver (src/Solver.h:19-53).
class Solver {
  public:
    Solver() : ctx_(Context::shared()) { cpm_solve_opts_default(&opts_); }
    explicit Solver(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) { cpm_solve_opts_default(&opts_); }
    ~Solver() { release(); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // src/Solver.cpp:8-68.  L and E are copied to the device once; the
    // reference keeps two host copies (by-value parameter and member).
    void Init(Scalar m, Scalar n, SpMat L_input, SpMat E_input, Preconditioner precond_input,
              const std::vector<bool>& identity_rows_input, int nD = 0) {
        const int64_t D = L_input.rows();
        if (E_input.rows() != D || (int64_t)identity_rows_input.size() != D)
            throw Error(CPM_ERR_ARG, "Solver::Init: L, E and identity_rows must have the same number of rows");
        std::vector<uint8_t> ident((size_t)D);
        for (int64_t i = 0; i < D; ++i) ident[(size_t)i] = identity_rows_input[(size_t)i] ? 1 : 0;
        release();
        check(cpm_op_create_csr(ctx_->get(), D, nD, L_input.outerIndexPtr(), L_input.innerIndexPtr(),
                                L_input.valuePtr(), E_input.outerIndexPtr(), E_input.innerIndexPtr(),
                                E_input.valuePtr(), ident.data(), m, n, (int)precond_input, precond_iters_, &op_),
              "Solver::Init");
        precond_ = precond_input;
    }

    // Adopt an operator built on the GPU (cpm_op_create_from_tube / _from_problem).
    void Init(cpm_op* op) {
        release();
        op_ = op;
    }

    // src/Solver.cpp:69-114: x0 = 0, tol 1e-10, max 100000 iterations.
    VectorX Solve(VectorX& rhs) {
        need_op();
        VectorX x(rhs.size(), 0.0);
        if ((int64_t)rhs.size() != cpm_op_rows(op_)) throw Error(CPM_ERR_ARG, "Solver::Solve: rhs size");
        check(cpm_solve(op_, rhs.data(), x.data(), &opts_, &info_), "Solver::Solve");
        return x;
    }

    // Several right-hand sides of this operator in one batched solve.
    std::vector<VectorX> Solve(const std::vector<VectorX>& rhs) {
        need_op();
        const int k = (int)rhs.size();
        std::vector<VectorX> xs(rhs.size());
        std::vector<cpm_op*> ops((size_t)k, op_);
        std::vector<const double*> b((size_t)k);
        std::vector<double*> x((size_t)k);
        for (int i = 0; i < k; ++i) {
            if ((int64_t)rhs[(size_t)i].size() != cpm_op_rows(op_)) throw Error(CPM_ERR_ARG, "Solver::Solve: rhs size");
            xs[(size_t)i].assign(rhs[(size_t)i].size(), 0.0);
            b[(size_t)i] = rhs[(size_t)i].data();
            x[(size_t)i] = xs[(size_t)i].data();
        }
        batch_info_.assign((size_t)k, cpm_solve_info{});
        check(cpm_solve_batch(ctx_->get(), k, ops.data(), b.data(), x.data(), &opts_, batch_info_.data()),
              "Solver::Solve(batch)");
        if (k) info_ = batch_info_[0];
        return xs;
    }

    // src/Solver.cpp:115-120 -> GeneralizedBiCGStabSystem::SetPrecond.
    void SetPrecond(int precond_iters) {
        precond_iters_ = precond_iters;
        if (op_) check(cpm_op_set_precond(op_, (int)precond_, precond_iters), "Solver::SetPrecond");
    }

    // src/Solver.cpp:121-138 (custom-solver branch): b = A(m, n) x with the
    // operator given to Init; L and E are not read, as in the reference.
    void Multiply(VectorX& x, VectorX& b, Scalar m, Scalar n, SpMat /*L*/, SpMat /*E*/) {
        need_op();
        b.assign(x.size(), 0.0);
        check(cpm_multiply(op_, x.data(), b.data(), m, n), "Solver::Multiply");
    }

    // Time stepping re-uses the operator with new coefficients (problems/HeatWithIBC.cpp:166-167).
    void SetMN(Scalar m, Scalar n) {
        need_op();
        check(cpm_op_set_mn(op_, m, n), "Solver::SetMN");
    }

    void SetPrecision(int precision, double tol = 1e-10) {
        opts_.precision = precision;
        opts_.tol = tol;
    }
    const cpm_solve_info& LastInfo() const { return info_; }
    const std::vector<cpm_solve_info>& LastBatchInfo() const { return batch_info_; }
    cpm_op* op() const { return op_; }

  private:
    void need_op() const {
        if (!op_) throw Error(CPM_ERR_STATE, "Solver: Init was not called");
    }
    void release() {
        if (op_) cpm_op_destroy(op_);
        op_ = nullptr;
    }
    std::shared_ptr<Context> ctx_;
    cpm_op* op_ = nullptr;
    Preconditioner precond_ = diagonal;
    int precond_iters_ = 0;
    cpm_solve_opts opts_{};
    cpm_solve_info info_{};
    std::vector<cpm_solve_info> batch_info_;
};

// Mesh or curve of a Wavefront OBJ file (SimplePolygonMesh::readMeshFromObjFile,
// src/SimplePolygonMesh.cpp:139-213): face_size 3 = triangles, 2 = polyline edges.
struct ObjMesh {
    std::vector<double> verts;    // 3 per vertex
    std::vector<int32_t> faces;   // face_size per face, 0-based
    int face_size = 0;
};

inline ObjMesh ReadObj(const std::string& path) {
    int64_t nv = 0, nf = 0;
    ObjMesh m;
    check(cpm_obj_read(path.c_str(), nullptr, &nv, nullptr, &nf, &m.face_size), "ReadObj");
    m.verts.resize((size_t)nv * 3);
    m.faces.resize((size_t)nf * m.face_size);
    check(cpm_obj_read(path.c_str(), m.verts.data(), &nv, m.faces.data(), &nf, &m.face_size), "ReadObj");
    return m;
}

// Consecutive polyline edges as the vertex chain of a "Polyline" surface or
// IBC (closed curves repeat their first vertex, src/SurfaceSpecifier.cpp:334-341).
inline std::vector<double> PolylineChain(const ObjMesh& m) {
    if (m.face_size != 2) throw Error(CPM_ERR_ARG, "PolylineChain: the mesh holds no polyline edges");
    int64_t n = 0;
    const int64_t nv = (int64_t)m.verts.size() / 3, ne = (int64_t)m.faces.size() / 2;
    check(cpm_polyline_chain(m.verts.data(), nv, m.faces.data(), ne, nullptr, &n), "PolylineChain");
    std::vector<double> c((size_t)n * 3);
    check(cpm_polyline_chain(m.verts.data(), nv, m.faces.data(), ne, c.data(), &n), "PolylineChain");
    return c;
}

struct ObjFile {
    std::string path;
};

// IBCMetaData (src/IBCMetaData.h:9-24): the fields the GPU path reads.
struct IBCMetaData {
    size_t boundary_type[2] = {0, 0};   // 0 Dirichlet, 1 Neumann
    size_t boundary_order = 1;
    Scalar on_ibc_tolerance = 0.0;      // uninitialised in the reference; every caller sets it
    Scalar ibc_DOFSubset_radius_scale = 1.0;
};

// SurfaceSpecifier (src/SurfaceSpecifier.h, .cpp:65-381): a surface or curve by
// type name and parameters, a polyline, a triangle mesh or an OBJ file.
class SurfaceSpecifier {
  public:
    SurfaceSpecifier() = default;
    // SetSurface(string): the reference's default parameters for the type
    explicit SurfaceSpecifier(const std::string& surf) { SetSurface(surf); }
    SurfaceSpecifier(const std::string& surf, const std::vector<Scalar>& bounding_box,
                     const std::vector<Scalar>& params, size_t embed_dim) {
        SetSurface(surf, bounding_box, params, embed_dim);
    }
    void SetSurface(const std::string& surf) {
        if (surf.size() > 4 && surf.compare(surf.size() - 4, 4, ".obj") == 0) {
            type_ = "obj";
            file_ = surf;
        } else {
            type_ = surf;
        }
        params_.clear();
        bbox_[0] = bbox_[1] = 0.0;
    }
    void SetSurface(const std::string& surf, const std::vector<Scalar>& bounding_box, const std::vector<Scalar>& params,
                    size_t embed_dim) {
        if (embed_dim != 3) throw Error(CPM_ERR_ARG, "SurfaceSpecifier: the GPU path takes 3-D embeddings");
        type_ = surf;
        params_ = params;
        bbox_[0] = bounding_box.size() > 0 ? bounding_box[0] : 0.0;
        bbox_[1] = bounding_box.size() > 1 ? bounding_box[1] : 0.0;
    }
    // A polyline through the vertex chain (3 values per vertex; closed when its
    // last vertex repeats the first, SurfaceSpecifier.cpp:334-341).
    static SurfaceSpecifier Polyline(const std::vector<Scalar>& chain, const std::vector<Scalar>& bounding_box = {}) {
        SurfaceSpecifier s("Polyline", bounding_box, {}, 3);
        s.verts_ = chain;
        return s;
    }
    // A closed triangle mesh ("Triangulation"; the box from the vertices, :346-363).
    static SurfaceSpecifier Mesh(const std::vector<Scalar>& verts, const std::vector<int32_t>& faces) {
        SurfaceSpecifier s;
        s.type_ = "Triangulation";
        s.verts_ = verts;
        s.faces_ = faces;
        return s;
    }
    void SetIBCMetaData(const IBCMetaData& meta) { meta_ = meta; }
    const IBCMetaData& IBCMeta() const { return meta_; }
    const std::string& SurfaceType() const { return type_; }
    const std::vector<Scalar>& Params() const { return params_; }
    const std::vector<Scalar>& Verts() const { return verts_; }
    const std::vector<int32_t>& Faces() const { return faces_; }
    const std::string& File() const { return file_; }
    Scalar BBoxLo() const { return bbox_[0]; }
    Scalar BBoxHi() const { return bbox_[1]; }
    // The curve as an interior boundary of the C ABI (pointers into this object).
    cpm_ibc_desc IBCDesc() const {
        cpm_ibc_desc d{};
        d.type = type_.c_str();
        d.params = params_.empty() ? nullptr : params_.data();
        d.nparams = (int)params_.size();
        d.verts = verts_.empty() ? nullptr : verts_.data();
        d.nverts = (int64_t)verts_.size() / 3;
        d.boundary_order = (int)meta_.boundary_order;
        d.boundary_type[0] = (int)meta_.boundary_type[0];
        d.boundary_type[1] = (int)meta_.boundary_type[1];
        d.on_ibc_tolerance = meta_.on_ibc_tolerance;
        d.dof_radius_scale = meta_.ibc_DOFSubset_radius_scale;
        return d;
    }

  private:
    std::string type_, file_;
    std::vector<Scalar> params_, verts_;
    std::vector<int32_t> faces_;
    Scalar bbox_[2] = {0.0, 0.0};
    IBCMetaData meta_;
};

// GPU tube of an analytic surface (Tube(SurfaceSpecifier&, p, dx), src/Tube.cpp:88-461).
class Tube {
  public:
    // Tube(SurfaceSpecifier&, interp_deg, dx) (src/Tube.h:45): any surface the GPU path takes.
    Tube(const SurfaceSpecifier& s, int p, Scalar dx, std::shared_ptr<Context> ctx = Context::shared())
        : ctx_(std::move(ctx)) {
        const auto& P = s.Params();
        if (s.SurfaceType() == "obj") {
            check(cpm_tube_create_obj(ctx_->get(), s.File().c_str(), s.BBoxLo(), s.BBoxHi(), p, dx, &t_), "Tube");
        } else if (s.SurfaceType() == "Triangulation") {
            check(cpm_tube_create_mesh(ctx_->get(), s.Verts().data(), (int64_t)s.Verts().size() / 3, s.Faces().data(),
                                       (int64_t)s.Faces().size() / 3, s.BBoxLo(), s.BBoxHi(), p, dx, &t_),
                  "Tube");
        } else {
            check(cpm_tube_create(ctx_->get(), s.SurfaceType().c_str(), P.empty() ? nullptr : P.data(), (int)P.size(),
                                  s.BBoxLo(), s.BBoxHi(), 3, s.Verts().empty() ? nullptr : s.Verts().data(),
                                  (int64_t)s.Verts().size() / 3, p, dx, &t_),
                  "Tube");
        }
        check(cpm_tube_get_info(t_, &info_), "Tube");
    }
    Tube(const std::string& type, Scalar dx, const std::vector<double>& params = {}, double bbox_lo = 0,
         double bbox_hi = 0, int p = 3, const std::vector<double>& verts = {},
         std::shared_ptr<Context> ctx = Context::shared())
        : ctx_(std::move(ctx)) {
        check(cpm_tube_create(ctx_->get(), type.c_str(), params.empty() ? nullptr : params.data(), (int)params.size(),
                              bbox_lo, bbox_hi, 3, verts.empty() ? nullptr : verts.data(), (int64_t)verts.size() / 3, p,
                              dx, &t_),
              "Tube");
        check(cpm_tube_get_info(t_, &info_), "Tube");
    }
    // Triangle mesh ("Triangulation", SurfaceSpecifier(".obj")): vertices 3 per
    // vertex, faces 3 indices per triangle; bbox_lo >= bbox_hi takes the
    // reference's box from the vertices.
    Tube(const std::vector<double>& verts, const std::vector<int32_t>& faces, Scalar dx, double bbox_lo = 0,
         double bbox_hi = 0, int p = 3, std::shared_ptr<Context> ctx = Context::shared())
        : ctx_(std::move(ctx)) {
        check(cpm_tube_create_mesh(ctx_->get(), verts.data(), (int64_t)verts.size() / 3, faces.data(),
                                   (int64_t)faces.size() / 3, bbox_lo, bbox_hi, p, dx, &t_),
              "Tube");
        check(cpm_tube_get_info(t_, &info_), "Tube");
    }
    // The surface of an OBJ file, SurfaceSpecifier(file) (src/SurfaceSpecifier.cpp:298-363):
    // triangles or a polyline.
    Tube(const ObjFile& file, Scalar dx, double bbox_lo = 0, double bbox_hi = 0, int p = 3,
         std::shared_ptr<Context> ctx = Context::shared())
        : ctx_(std::move(ctx)) {
        check(cpm_tube_create_obj(ctx_->get(), file.path.c_str(), bbox_lo, bbox_hi, p, dx, &t_), "Tube");
        check(cpm_tube_get_info(t_, &info_), "Tube");
    }
    ~Tube() { cpm_tube_destroy(t_); }
    Tube(const Tube&) = delete;
    Tube& operator=(const Tube&) = delete;
    int64_t NumTubeNodes() const { return info_.nodes; }
    const cpm_tube_info& info() const { return info_; }
    cpm_tube* get() const { return t_; }
    // Closest points of the tube nodes, 3 per node (Tube::GetCP).
    std::vector<double> CP() const {
        std::vector<double> cp((size_t)info_.nodes * 3);
        check(cpm_tube_export(t_, nullptr, nullptr, cp.data(), nullptr, nullptr, nullptr, nullptr), "Tube::CP");
        return cp;
    }
    // Operator m I + n (Ld + Loff E) built on the GPU (Poisson::Initialize + Solver::Init).
    cpm_op* Operator(Scalar m, Scalar n, Preconditioner p = diagonal) const {
        cpm_op* op = nullptr;
        check(cpm_op_create_from_tube(t_, m, n, (int)p, 0, &op), "Tube::Operator");
        return op;
    }

  private:
    std::shared_ptr<Context> ctx_;
    cpm_tube* t_ = nullptr;
    cpm_tube_info info_{};
};

typedef std::function<Scalar(const std::vector<Scalar>&)> ScalarFunction;
// ibc_exact(curve, which side, x) of the reference's GetRHS / TimeStep (side 0 for unoriented curves)
typedef std::function<Scalar(size_t&, const size_t&, const std::vector<Scalar>&)> IBCFunction;

// Device vector of doubles on a context.
class DeviceVec {
  public:
    DeviceVec(std::shared_ptr<Context> ctx, size_t n) : ctx_(std::move(ctx)), n_(n) {
        check(cpm_device_alloc(ctx_->get(), (int64_t)(8 * std::max<size_t>(n, 1)), &p_), "DeviceVec");
    }
    ~DeviceVec() { cpm_device_free(ctx_->get(), p_); }
    DeviceVec(const DeviceVec&) = delete;
    DeviceVec& operator=(const DeviceVec&) = delete;
    void upload(const VectorX& v) {
        if (v.size() != n_) throw Error(CPM_ERR_ARG, "DeviceVec: size");
        check(cpm_memcpy_h2d(ctx_->get(), p_, v.data(), (int64_t)(8 * n_)), "DeviceVec::upload");
    }
    void download(VectorX& v) const {
        v.resize(n_);
        check(cpm_memcpy_d2h(ctx_->get(), v.data(), p_, (int64_t)(8 * n_)), "DeviceVec::download");
    }
    double* get() const { return static_cast<double*>(p_); }

  private:
    std::shared_ptr<Context> ctx_;
    size_t n_;
    void* p_ = nullptr;
};

// y = A x for a host CSR matrix (plotting rows times a solution).
inline VectorX Apply(const SpMat& A, const VectorX& x) {
    VectorX y((size_t)A.rows(), 0.0);
    for (int64_t r = 0; r < A.rows(); ++r) {
        double acc = 0.0;
        for (int32_t e = A.outerIndexPtr()[r]; e < A.outerIndexPtr()[r + 1]; ++e)
            acc += A.valuePtr()[e] * x[(size_t)A.innerIndexPtr()[e]];
        y[(size_t)r] = acc;
    }
    return y;
}

inline std::vector<double> Flatten(const std::vector<std::vector<Scalar>>& pts) {
    std::vector<double> q;
    q.reserve(pts.size() * 3);
    for (const auto& x : pts) {
        if (x.size() != 3) throw Error(CPM_ERR_ARG, "points need 3 coordinates");
        q.insert(q.end(), x.begin(), x.end());
    }
    return q;
}

// Poisson without boundaries on the GPU tube (problems/Poisson.h:14-112, Poisson.cpp:5-238):
// E and L of Helpers::SetupMatrices, no identity rows (closed surfaces).
class Poisson {
  public:
    Poisson(SurfaceSpecifier& surface_specs, Scalar dx, int interp_deg = 3,
            std::shared_ptr<Context> ctx = Context::shared())
        : tube_(new Tube(surface_specs, interp_deg, dx, std::move(ctx))) {
        const auto cp = tube_->CP();
        cpx_.resize(cp.size() / 3);
        for (size_t i = 0; i < cpx_.size(); ++i) cpx_[i] = {cp[3 * i], cp[3 * i + 1], cp[3 * i + 2]};
        identity_rows_.assign(cpx_.size(), false);
    }
    // GetRHS (Poisson.cpp): f at the closest point of every tube node; exact is read only on
    // boundary rows of open surfaces, which the GPU path does not build.
    VectorX GetRHS(ScalarFunction f, ScalarFunction /*exact*/ = nullptr) const {
        VectorX rhs(cpx_.size());
        for (size_t i = 0; i < cpx_.size(); ++i) rhs[i] = f(cpx_[i]);
        return rhs;
    }
    SpMat GetLfull() const {
        const int64_t N = tube_->NumTubeNodes(), nnz = tube_->info().nnz_L;
        std::vector<int32_t> p((size_t)N + 1), c((size_t)nnz);
        std::vector<double> v((size_t)nnz);
        check(cpm_tube_export_L(tube_->get(), p.data(), c.data(), v.data()), "Poisson::GetLfull");
        SpMat L;
        L.assign(N, N, std::move(p), std::move(c), std::move(v));
        return L;
    }
    SpMat GetEfull() const {
        const int64_t N = tube_->NumTubeNodes(), W = StencilSize(), nnz = N * W;
        std::vector<int32_t> p((size_t)N + 1), c((size_t)nnz);
        std::vector<double> v((size_t)nnz);
        check(cpm_tube_export_E(tube_->get(), p.data(), c.data(), v.data()), "Poisson::GetEfull");
        SpMat E;
        E.assign(N, N, std::move(p), std::move(c), std::move(v));
        return E;
    }
    const std::vector<bool>& GetIdentityRows() const { return identity_rows_; }
    // Interpolation rows at points on the surface (Helpers::GetPlotInterpMatrix, src/Helpers.cpp:97-103).
    SpMat GetInterpMatrix(const std::vector<std::vector<Scalar>>& xq) const {
        const auto q = Flatten(xq);
        const int64_t nq = (int64_t)xq.size(), W = StencilSize();
        std::vector<int32_t> p((size_t)nq + 1), c((size_t)(nq * W));
        std::vector<double> v((size_t)(nq * W));
        check(cpm_tube_interp_matrix(tube_->get(), q.data(), nq, p.data(), c.data(), v.data()), "Poisson::GetInterpMatrix");
        SpMat A;
        A.assign(nq, tube_->NumTubeNodes(), std::move(p), std::move(c), std::move(v));
        return A;
    }
    // The operator m I + n (Ld + Loff E) built on the GPU, for Solver::Init(cpm_op*).
    cpm_op* Operator(Scalar m, Scalar n, Preconditioner p = diagonal) const { return tube_->Operator(m, n, p); }
    const std::vector<std::vector<Scalar>>& cpx() const { return cpx_; }
    Tube& tube() { return *tube_; }

  private:
    int64_t StencilSize() const {
        const int64_t s = tube_->info().p + 1;
        return s * s * s;
    }
    std::unique_ptr<Tube> tube_;
    std::vector<std::vector<Scalar>> cpx_;
    std::vector<bool> identity_rows_;
};

// PoissonWithIBC on the GPU (problems/PoissonWithIBC.h:15-141, .cpp:8-192): rows [N | G | K].
class PoissonWithIBC {
  public:
    // PoissonWithIBC(surface_specs, ibc_specs, dx, interp_deg) (problems/PoissonWithIBC.cpp:8-47).
    PoissonWithIBC(SurfaceSpecifier& surface_specs, std::vector<SurfaceSpecifier>& ibc_specs, Scalar dx,
                   size_t interp_deg = 3, std::shared_ptr<Context> ctx = Context::shared())
        : own_(new Tube(surface_specs, (int)interp_deg, dx, std::move(ctx))) {
        std::vector<cpm_ibc_desc> d;
        for (const auto& s : ibc_specs) d.push_back(s.IBCDesc());
        create(*own_, d);
    }
    // The curves on a resident tube (several problems can share it).
    PoissonWithIBC(Tube& tube, const std::vector<cpm_ibc_desc>& ibcs) { create(tube, ibcs); }
    ~PoissonWithIBC() { cpm_problem_destroy(p_); }
    PoissonWithIBC(const PoissonWithIBC&) = delete;
    PoissonWithIBC& operator=(const PoissonWithIBC&) = delete;
    const cpm_problem_info& info() const { return info_; }
    cpm_problem* get() const { return p_; }
    size_t NumIBCs() const { return (size_t)info_.n_ibc; }
    int64_t Rows() const { return info_.rows; }
    // Evaluation points and kinds of GetRHS (problems/PoissonWithIBC.cpp:50-90):
    // 0 f at cp, 1 f at the curve's closest point, 2 the boundary value there.
    void RHSPoints(std::vector<double>& pts, std::vector<int32_t>& kind) const {
        pts.resize((size_t)info_.rows * 3);
        kind.resize((size_t)info_.rows);
        check(cpm_problem_rhs_points(p_, pts.data(), kind.data()), "PoissonWithIBC::RHSPoints");
    }
    // GetRHS(f, surface_exact) (problems/PoissonWithIBC.cpp:50-91).
    VectorX GetRHS(ScalarFunction f, ScalarFunction surface_exact) const {
        return rhs(f, [&](size_t&, const std::vector<Scalar>& x) { return surface_exact(x); });
    }
    // GetRHS(f, ibc_exact(curve, side, x)) (:93-133); side 0 (curves are not oriented).
    VectorX GetRHS(ScalarFunction f, IBCFunction ibc_exact) const {
        return rhs(f, [&](size_t& c, const std::vector<Scalar>& x) {
            const size_t side = 0;
            return ibc_exact(c, side, x);
        });
    }
    // Dirichlet values of every row (0 on the others) from ibc_exact at the curves' closest points:
    // the identity rows' values (HeatWithIBC::SetDirichletValue, problems/HeatWithIBC.cpp:50-78).
    VectorX DirichletValues(IBCFunction ibc_exact) const {
        VectorX g((size_t)info_.rows, 0.0);
        std::vector<double> pts;
        std::vector<int32_t> kind;
        RHSPoints(pts, kind);
        for (size_t c = 0; c < NumIBCs(); ++c) {
            int64_t rows[4];
            check(cpm_problem_curve_rows(p_, (int)c, rows), "PoissonWithIBC::DirichletValues");
            for (int64_t r = rows[2]; r >= 0 && r < rows[2] + rows[3]; ++r) {
                const size_t side = 0;
                g[(size_t)r] = ibc_exact(c, side, {pts[3 * r], pts[3 * r + 1], pts[3 * r + 2]});
            }
        }
        return g;
    }
    // GetLfull / GetEfull / GetIdentityRows (problems/PoissonWithIBC.h:102-115).
    void Matrices(SpMat& L, SpMat& E, std::vector<bool>& identity_rows) const {
        int64_t nL = 0, nE = 0;
        check(cpm_problem_export(p_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &nL, &nE),
              "PoissonWithIBC::Matrices");
        const int64_t D = info_.rows;
        std::vector<int32_t> lp((size_t)D + 1), lc((size_t)nL), ep((size_t)D + 1), ec((size_t)nE);
        std::vector<double> lv((size_t)nL), ev((size_t)nE);
        std::vector<uint8_t> id((size_t)D);
        check(cpm_problem_export(p_, lp.data(), lc.data(), lv.data(), ep.data(), ec.data(), ev.data(), id.data(),
                                 &nL, &nE),
              "PoissonWithIBC::Matrices");
        L.assign(D, D, std::move(lp), std::move(lc), std::move(lv));
        E.assign(D, D, std::move(ep), std::move(ec), std::move(ev));
        identity_rows.assign(id.begin(), id.end());
    }
    SpMat GetLfull() const {
        SpMat L, E;
        std::vector<bool> id;
        Matrices(L, E, id);
        return L;
    }
    SpMat GetEfull() const {
        SpMat L, E;
        std::vector<bool> id;
        Matrices(L, E, id);
        return E;
    }
    std::vector<bool> GetIdentityRows() const {
        SpMat L, E;
        std::vector<bool> id;
        Matrices(L, E, id);
        return id;
    }
    // GetPlotInterpMatrix at the points xp (problems/PoissonWithIBC.cpp:135-143): the tube's
    // interpolation rows with the stencil entries across a curve moved to its ghost columns.
    SpMat GetPlotInterpMatrix(const std::vector<std::vector<Scalar>>& xp) const {
        const auto q = Flatten(xp);
        const int64_t nq = (int64_t)xp.size();
        const int64_t W = 64;   // (p + 1)^3 with p <= 3
        std::vector<int32_t> p((size_t)nq + 1), c((size_t)(nq * W));
        std::vector<double> v((size_t)(nq * W));
        check(cpm_problem_plot_matrix(p_, q.data(), nq, p.data(), c.data(), v.data()),
              "PoissonWithIBC::GetPlotInterpMatrix");
        const int32_t nnz = p[(size_t)nq];
        c.resize((size_t)nnz);
        v.resize((size_t)nnz);
        SpMat A;
        A.assign(nq, info_.rows, std::move(p), std::move(c), std::move(v));
        return A;
    }
    // IBCParameter(ibc_index, x) (problems/PoissonWithIBC.h:57-60 -> Surface::IBCParameter).
    Scalar IBCParameter(size_t ibc_index, const std::vector<Scalar>& x) const {
        return IBCParameters(ibc_index, {x})[0];
    }
    // The same for many points; vertex_param: one value per polyline vertex (the nearest
    // vertex's value, the reference's rule for parameterised curves).
    std::vector<Scalar> IBCParameters(size_t ibc_index, const std::vector<std::vector<Scalar>>& xs,
                                      const std::vector<Scalar>& vertex_param = {}) const {
        const auto q = Flatten(xs);
        std::vector<Scalar> out(xs.size());
        check(cpm_problem_ibc_parameter(p_, (int)ibc_index, q.data(), (int64_t)xs.size(),
                                        vertex_param.empty() ? nullptr : vertex_param.data(), out.data()),
              "PoissonWithIBC::IBCParameter");
        return out;
    }
    cpm_op* Operator(Scalar m, Scalar n, Preconditioner p = diagonal, int precond_iters = 0) const {
        cpm_op* op = nullptr;
        check(cpm_op_create_from_problem(p_, m, n, (int)p, precond_iters, &op), "PoissonWithIBC::Operator");
        return op;
    }

  private:
    void create(Tube& tube, const std::vector<cpm_ibc_desc>& ibcs) {
        check(cpm_problem_create(tube.get(), (int)ibcs.size(), ibcs.data(), &p_), "PoissonWithIBC");
        check(cpm_problem_get_info(p_, &info_), "PoissonWithIBC");
    }
    template <typename G>
    VectorX rhs(const ScalarFunction& f, G boundary) const {
        std::vector<double> pts;
        std::vector<int32_t> kind;
        RHSPoints(pts, kind);
        VectorX b((size_t)info_.rows, 0.0);
        std::vector<size_t> curve((size_t)info_.rows, 0);
        for (size_t c = 0; c < NumIBCs(); ++c) {
            int64_t rows[4];
            check(cpm_problem_curve_rows(p_, (int)c, rows), "PoissonWithIBC::GetRHS");
            for (int64_t r = rows[2]; r >= 0 && r < rows[2] + rows[3]; ++r) curve[(size_t)r] = c;
        }
        for (int64_t r = 0; r < info_.rows; ++r) {
            const std::vector<Scalar> x{pts[3 * r], pts[3 * r + 1], pts[3 * r + 2]};
            if (kind[(size_t)r] == 0 || kind[(size_t)r] == 1) b[(size_t)r] = f(x);
            else if (kind[(size_t)r] == 2) b[(size_t)r] = boundary(curve[(size_t)r], x);
        }
        return b;
    }
    std::unique_ptr<Tube> own_;
    cpm_problem* p_ = nullptr;
    cpm_problem_info info_{};
};

// HeatWithIBC (problems/HeatWithIBC.h:17-140): time_stepping_method 0 explicit,
// 1 backward Euler, 2 Crank-Nicolson (.cpp:80-103, 160-166), on the GPU.
class HeatWithIBC {
  public:
    HeatWithIBC(SurfaceSpecifier& surface_specs, std::vector<SurfaceSpecifier>& ibc_specs, Scalar dx, Scalar dt,
                size_t time_stepping_method, size_t interp_deg = 3, std::shared_ptr<Context> ctx = Context::shared())
        : ctx_(ctx), problem_(surface_specs, ibc_specs, dx, interp_deg, ctx), dt_(dt),
          method_((int)time_stepping_method), u_(ctx, (size_t)problem_.Rows()), g_(ctx, (size_t)problem_.Rows()) {
        if (method_ < 0 || method_ > 2) throw Error(CPM_ERR_ARG, "HeatWithIBC: time_stepping_method 0, 1 or 2");
        const Scalar r = method_ == 1 ? 1.0 : 0.5;   // Initialize (:160-166)
        op_ = problem_.Operator(1.0, -r * dt);
        cpm_solve_opts_default(&opts_);
    }
    ~HeatWithIBC() { cpm_op_destroy(op_); }
    HeatWithIBC(const HeatWithIBC&) = delete;
    HeatWithIBC& operator=(const HeatWithIBC&) = delete;
    // TimeStep(rhsfull, ibc_exact) (:80-103): the next state from u^n, the Dirichlet rows set to
    // ibc_exact at the curve points (the caller's lambda reads its own current time).
    VectorX TimeStep(VectorX& rhsfull, IBCFunction ibc_exact) {
        u_.upload(rhsfull);
        g_.upload(problem_.DirichletValues(ibc_exact));
        check(cpm_heat_step_device(op_, method_, dt_, g_.get(), u_.get(), &opts_, &info_), "HeatWithIBC::TimeStep");
        VectorX out;
        u_.download(out);
        return out;
    }
    SpMat GetPlotInterpMatrix(const std::vector<std::vector<Scalar>>& xp) const { return problem_.GetPlotInterpMatrix(xp); }
    PoissonWithIBC& problem() { return problem_; }
    const cpm_solve_info& LastInfo() const { return info_; }
    void SetPrecision(int precision, double tol = 1e-10) {
        opts_.precision = precision;
        opts_.tol = tol;
    }

  private:
    std::shared_ptr<Context> ctx_;
    PoissonWithIBC problem_;
    Scalar dt_;
    int method_;
    DeviceVec u_, g_;
    cpm_op* op_ = nullptr;
    cpm_solve_opts opts_{};
    cpm_solve_info info_{};
};

// ReactionDiffusionWithIBC (problems/ReactionDiffusionWithIBC.h:17-126): Gray-Scott with the
// reference's constants (.cpp:81-84) and the given dt; method 0 is its explicit TimeStep
// (:79-104), method 1 treats the diffusion implicitly (no counterpart in the reference).
class ReactionDiffusionWithIBC {
  public:
    ReactionDiffusionWithIBC(SurfaceSpecifier& surface_specs, std::vector<SurfaceSpecifier>& ibc_specs, Scalar dx,
                             Scalar dt, size_t interp_deg = 3, std::shared_ptr<Context> ctx = Context::shared())
        : ctx_(ctx), problem_(surface_specs, ibc_specs, dx, interp_deg, ctx), u_(ctx, (size_t)problem_.Rows()),
          v_(ctx, (size_t)problem_.Rows()), gu_(ctx, (size_t)problem_.Rows()), gv_(ctx, (size_t)problem_.Rows()) {
        cpm_rd_params_default(dx, &prm_);
        prm_.dt = dt;
        op_ = problem_.Operator(1.0, -1.0);
        cpm_solve_opts_default(&opts_);
    }
    ~ReactionDiffusionWithIBC() { cpm_op_destroy(op_); }
    ReactionDiffusionWithIBC(const ReactionDiffusionWithIBC&) = delete;
    ReactionDiffusionWithIBC& operator=(const ReactionDiffusionWithIBC&) = delete;
    // TimeStep(ufull, vfull, ibc_exact_u, ibc_exact_v) (:79-104): both fields in place.
    void TimeStep(VectorX& ufull, VectorX& vfull, IBCFunction ibc_exact_u, IBCFunction ibc_exact_v) {
        u_.upload(ufull);
        v_.upload(vfull);
        gu_.upload(problem_.DirichletValues(ibc_exact_u));
        gv_.upload(problem_.DirichletValues(ibc_exact_v));
        double* du = u_.get();
        double* dv = v_.get();
        check(cpm_rd_step_device(op_, &prm_, 1, &du, &dv, gu_.get(), gv_.get(), &opts_, info_),
              "ReactionDiffusionWithIBC::TimeStep");
        u_.download(ufull);
        v_.download(vfull);
    }
    cpm_rd_params& Params() { return prm_; }   // F, k, nu_u, nu_v, dt, method
    PoissonWithIBC& problem() { return problem_; }

  private:
    std::shared_ptr<Context> ctx_;
    PoissonWithIBC problem_;
    DeviceVec u_, v_, gu_, gv_;
    cpm_rd_params prm_{};
    cpm_op* op_ = nullptr;
    cpm_solve_opts opts_{};
    cpm_solve_info info_[2]{};
};

// GeodesicDistance (problems/GeodesicDistance.h:20-66): the heat method from the problem's
// curves on the GPU (ComputeDistance, .cpp:11-67).
class GeodesicDistance {
  public:
    // ComputeDistance(Tube&, IBCSubsets&, dt): the distance at the points xp (the reference
    // returns it at its surface's parameterisation points, :62-66) through the problem's
    // plotting rows; jacobi_sweeps as the reference's 2000 (:112-115).
    VectorX ComputeDistance(PoissonWithIBC& problem, Scalar dt, const std::vector<std::vector<Scalar>>& xp,
                            int jacobi_sweeps = 2000, std::shared_ptr<Context> ctx = Context::shared()) {
        DeviceVec phi(ctx, (size_t)problem.Rows());
        cpm_solve_opts o;
        cpm_solve_opts_default(&o);
        infos_.assign(4, cpm_solve_info{});
        check(cpm_geodesic_distance(problem.get(), dt, jacobi_sweeps, &o, phi.get(), infos_.data()),
              "GeodesicDistance::ComputeDistance");
        phi.download(phi_);
        return Apply(problem.GetPlotInterpMatrix(xp), phi_);
    }
    const VectorX& Phi() const { return phi_; }   // on the problem's D rows
    const std::vector<cpm_solve_info>& Infos() const { return infos_; }

  private:
    VectorX phi_;
    std::vector<cpm_solve_info> infos_;
};

}  // namespace cpm
