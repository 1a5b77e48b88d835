// Host-side handles of libcpmgpu: context and operator.
#pragma once
#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cpm_gpu.h"
#include "engine.h"
#include "route.h"

namespace cpmgpu {

struct ArgError : std::runtime_error { using std::runtime_error::runtime_error; };
struct NoMem : std::runtime_error { using std::runtime_error::runtime_error; };
struct GeomError : std::runtime_error { using std::runtime_error::runtime_error; };
struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };

// Fixed device memory carved front to back: the workspaces of a solve that
// may not allocate (cpm_solve_device_noalloc) come from here instead of the
// context's grow-only buffers.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void* take(size_t bytes, const char* what);
};

struct Ctx {
    static constexpr int kHostStates = 8192;
    int device = 0;
    hipStream_t stream = nullptr;
    KState* host_states = nullptr;                            // pinned mirror of device states
    std::map<std::string, std::unique_ptr<Buffer>> ws;        // grow-only named workspaces
    std::map<void*, int64_t> user_bytes;
    int64_t op_bytes = 0;
    Arena* arena = nullptr;   // set: workspace() carves from it and never allocates
    bool no_builds = false;   // set: solves build no shared-stencil tiles (they would allocate)
    // FP32+refine history of tensor-product operators by family (the same m, n, L_off and L_ii:
    // one equation on one grid spacing), which FP64 routing consults for an operator with no
    // history of its own (a served request's operator is built for it and solved once)
    std::map<std::array<double, 4>, Fp32Record> fp32_family;
    explicit Ctx(int dev);
    ~Ctx();
    double* workspace(const char* name, size_t bytes);
};

// Solves on the context build no tiles while this object lives.
struct NoBuilds {
    Ctx* c;
    bool old;
    explicit NoBuilds(Ctx* c_) : c(c_), old(c_->no_builds) { c->no_builds = true; }
    ~NoBuilds() { c->no_builds = old; }
    NoBuilds(const NoBuilds&) = delete;
    NoBuilds& operator=(const NoBuilds&) = delete;
};

// Operator in device memory, in one of two execution formats (see OpView):
// slot-major ELL with int32 columns shared by the FP64 and FP32 value arrays,
// or the tensor-product format of tube-built operators.
struct Op {
    Ctx* ctx = nullptr;
    Fp32Record fp32;            // FP32+refine history of this operator's systems (FP64 routing)
    int64_t D = 0;
    int nD = 0;
    int ew = 0, lw = 0;
    double m = 0, n = 0;
    int precond = 1, precond_iters = 0;
    Buffer ecol, eval64, eval32, lcol, lval64, lval32, ldiag64, ldiag32, flag;
    // OP_TP
    int fmt = OP_ELL;
    Buffer perm;              // int32 [D]: internal row -> caller row (empty = identity)
    Buffer tpc, tpw64, tpw32, exc, ecx, xcol, xval64, xval32;
    int64_t nexc = 0, nmoved = 0;
    double loff = 0, ldu = 0;
    Buffer tln, tlo, tlr, tpl, tlx;   // LDS tiles (OP_TP)
    Buffer gofs, grow, gpl, gw64, gw32, gln, glr, glx;   // shared-stencil tiles (OP_TP)
    int ngt = 0, nitem = 0;
    int64_t nslot = 0;
    bool grp_mode = false, grp_tried = false;   // shared-stencil tiles wanted / build attempted
    // gathered (non-identity) rows: [0] plain and [1] exception rows of LDS
    // tiles, [2] plain and [3] exception rows read directly; rows in all lists;
    // kind-0 items of shared-stencil tiles
    int64_t tile_rows[4] = {0, 0, 0, 0};
    int64_t tile_list = 0;
    int64_t grp_items = 0, grp_items0 = 0, grp_list = 0;
    int64_t grp_rows0 = 0, grp_exc = 0, grp_ident = 0, grp_ident_items = 0;
    // Lfull rows of the rows that solves treat as identity rows (internal
    // order, hw slots each, padded with weight 0): explicit time stepping
    // applies Lfull to every row (problems/HeatWithIBC.cpp:84-88).  Entries
    // whose row is not an identity row are skipped.
    Buffer hrow, hcol, hval;  // int32 [nh], int32 [hw][nh], double [hw][nh]
    int64_t nh = 0;
    int hw = 0;
    Op() = default;
    Op(Ctx* c, int64_t D_, int nD_);
    void upload_ell(const std::vector<int>& ec, const std::vector<double>& ev, int ew,
                    const std::vector<int>& lc, const std::vector<double>& lv, int lw,
                    const std::vector<double>& ld, const std::vector<uint8_t>& fl);
    template <typename T> OpView<T> view() const;
    size_t device_bytes() const;
};

// Builds a tensor-product operator's shared-stencil tiles on the first call
// (when its build asked for them); later calls return at once.
void op_ensure_group_tiles(Op& op);
// The same for every operator of a launch of T x nr values, if that shape
// uses shared-stencil tiles.
void ensure_group_tiles_for(const std::vector<Op*>& ops, int tsize, int nr);

// Batched BiCGStab over instances (ops[i], rhs[i] -> xs[i], caller-order
// device vectors); mn, when given, overrides each instance's (m, n).
void solve_batch(Ctx* ctx, const std::vector<Op*>& ops, const std::vector<const double*>& rhs,
                 const std::vector<double*>& xs, const cpm_solve_opts& o, std::vector<cpm_solve_info>& infos,
                 const std::vector<std::pair<double, double>>* mn = nullptr);
// y = E x for caller-order device vectors of the operator's D rows (identity
// rows copy x); x and y may not alias.
void apply_E(Op* op, const double* x, double* y);
cpm_solve_opts resolve_opts(const cpm_solve_opts* o);

Op* op_from_csr(Ctx* ctx, int64_t D, int nD, const int32_t* Lp, const int32_t* Lc, const double* Lv,
                const int32_t* Ep, const int32_t* Ec, const double* Ev, const uint8_t* ident, double m, double n,
                int precond, int precond_iters);

}  // namespace cpmgpu
