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

// GPU tube of an analytic surface (Tube(SurfaceSpecifier&, p, dx), src/Tube.cpp:88-461).
class Tube {
  public:
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

// PoissonWithIBC on the GPU (problems/PoissonWithIBC.cpp:8-47): rows [N | G | K].
class PoissonWithIBC {
  public:
    PoissonWithIBC(Tube& tube, const std::vector<cpm_ibc_desc>& ibcs) {
        check(cpm_problem_create(tube.get(), (int)ibcs.size(), ibcs.data(), &p_), "PoissonWithIBC");
        check(cpm_problem_get_info(p_, &info_), "PoissonWithIBC");
    }
    ~PoissonWithIBC() { cpm_problem_destroy(p_); }
    PoissonWithIBC(const PoissonWithIBC&) = delete;
    PoissonWithIBC& operator=(const PoissonWithIBC&) = delete;
    const cpm_problem_info& info() const { return info_; }
    // Evaluation points and kinds of GetRHS (problems/PoissonWithIBC.cpp:50-90).
    void RHSPoints(std::vector<double>& pts, std::vector<int32_t>& kind) const {
        pts.resize((size_t)info_.rows * 3);
        kind.resize((size_t)info_.rows);
        check(cpm_problem_rhs_points(p_, pts.data(), kind.data()), "PoissonWithIBC::RHSPoints");
    }
    // GetLfull / GetEfull / GetIdentityRows (problems/PoissonWithIBC.h).
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
    cpm_op* Operator(Scalar m, Scalar n, Preconditioner p = diagonal) const {
        cpm_op* op = nullptr;
        check(cpm_op_create_from_problem(p_, m, n, (int)p, 0, &op), "PoissonWithIBC::Operator");
        return op;
    }

  private:
    cpm_problem* p_ = nullptr;
    cpm_problem_info info_{};
};

}  // namespace cpm
