// C ABI of libcpmgpu.so: contexts, operators, solves.  See include/cpm_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpm_gpu.h"
#include "engine.h"
#include "krylov.h"
#include "ops.h"
#include "pool.h"

namespace cpmgpu {

static thread_local std::string g_err;

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

Buffer::Buffer(size_t n) : bytes(n) {
    if (n == 0) return;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        throw NoMem("hipMalloc of " + std::to_string(n) + " bytes failed");
    }
}
Buffer::~Buffer() {
    if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------- packing
// Move up to NR caller vectors into a workspace vector and back: interleaved
// [row][rhs] (nr = NR, ls = 1) or lane-major [rhs][row] (nr = 1, ls = D).
// The operator's internal row r holds the caller's row perm[r] (tube-built
// operators run in lexicographic order).
struct PackEntry {
    const double* src;   // caller vector (pack) or nullptr
    double* dst;         // caller vector (unpack) or nullptr
    const int* perm;     // internal row -> caller row, nullptr = identity
    int64_t D;
    int64_t off;
    int lane, nr;        // lane, row stride
    int block0;
    int64_t ls = 1;      // lane stride
};

__global__ void k_pack(const PackEntry* ent, const int* bent, double* ws64, float* ws32, int to_ws) {
    const PackEntry& e = ent[bent[blockIdx.x]];
    const int64_t row = (int64_t)(blockIdx.x - e.block0) * 256 + threadIdx.x;
    if (row >= e.D) return;
    const size_t i = (size_t)(e.off + row * e.nr + e.lane * e.ls);
    const int64_t cr = e.perm ? e.perm[row] : row;
    if (to_ws) {
        const double v = e.src ? e.src[cr] : 0.0;
        if (ws64) ws64[i] = v;
        if (ws32) ws32[i] = (float)v;
    } else if (e.dst) {   // padding lanes of a group have no caller vector
        e.dst[cr] = ws64 ? ws64[i] : (double)ws32[i];
    }
}

// ---------------------------------------------------------------- no-alloc scratch
// Device memory of cpm_solve_device_noalloc: a file-scope array that every
// GPU running this library holds until the process ends (its own copy at its
// own address).  Host code never uses the array's name as an address: the
// address on each device comes from hipGetSymbolAddress, taken once per
// device when the first context on it is created.
static __device__ __align__(256) char scratch[CPM_NOALLOC_SCRATCH_BYTES];

namespace {
struct ScratchAddr {
    void* ptr = nullptr;
    std::string err;   // why the address could not be taken (ptr stays null)
};
std::mutex g_scratch_mu;
std::map<int, ScratchAddr> g_scratch;   // device -> address of `scratch` there
}  // namespace

// With `dev` current: record the device address of `scratch` on it.
static void scratch_setup(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch.count(dev)) return;
    ScratchAddr a;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        a.err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    } else {
        a.ptr = p;
    }
    g_scratch[dev] = a;
}

static void* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(dev);
    if (it == g_scratch.end()) throw std::logic_error("no-alloc scratch not set up on this device");
    if (!it->second.ptr) throw HipError(it->second.err);
    return it->second.ptr;
}

void* Arena::take(size_t bytes, const char* what) {
    const size_t n = (bytes + 255) & ~size_t(255);
    if (n > cap - used)
        throw ArgError(std::string("no-alloc scratch of ") + std::to_string(cap) + " bytes too small: " + what +
                       " needs " + std::to_string(n) + " more bytes after " + std::to_string(used));
    void* p = base + used;
    used += n;
    return p;
}

Ctx::Ctx(int dev) : device(dev) {
    HIPCHK(hipSetDevice(dev));
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc(&host_states, sizeof(KState) * kHostStates));
    scratch_setup(dev);
}
Ctx::~Ctx() {
    (void)hipStreamSynchronize(stream);
    ws.clear();
    (void)hipHostFree(host_states);
    (void)hipStreamDestroy(stream);
}

double* Ctx::workspace(const char* name, size_t bytes) {
    if (arena) return static_cast<double*>(arena->take(std::max(bytes, (size_t)256), name));
    auto& b = ws[name];
    if (!b || b->bytes < bytes) {
        b.reset();
        b.reset(new Buffer(std::max(bytes, (size_t)256)));
    }
    return b->as<double>();
}

// ---------------------------------------------------------------- operator
Op::Op(Ctx* c, int64_t D_, int nD_) : ctx(c), D(D_), nD(nD_) {}

void Op::upload_ell(const std::vector<int>& ec, const std::vector<double>& ev, int ew_,
                    const std::vector<int>& lc, const std::vector<double>& lv, int lw_,
                    const std::vector<double>& ld, const std::vector<uint8_t>& fl) {
    ew = ew_;
    lw = lw_;
    auto up = [&](Buffer& b, const void* src, size_t bytes) {
        b.~Buffer();
        new (&b) Buffer(bytes);
        if (bytes) HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    std::vector<float> ev32(ev.begin(), ev.end()), lv32(lv.begin(), lv.end()), ld32(ld.begin(), ld.end());
    up(ecol, ec.data(), ec.size() * sizeof(int));
    up(eval64, ev.data(), ev.size() * sizeof(double));
    up(eval32, ev32.data(), ev32.size() * sizeof(float));
    up(lcol, lc.data(), lc.size() * sizeof(int));
    up(lval64, lv.data(), lv.size() * sizeof(double));
    up(lval32, lv32.data(), lv32.size() * sizeof(float));
    up(ldiag64, ld.data(), ld.size() * sizeof(double));
    up(ldiag32, ld32.data(), ld32.size() * sizeof(float));
    up(flag, fl.data(), fl.size());
}

template <typename T>
static OpView<T> make_view(const Op& o, const Buffer& ev, const Buffer& lv, const Buffer& ld, const Buffer& tw,
                           const Buffer& xv, const Buffer& gw) {
    OpView<T> v{};
    v.D = o.D;
    v.nD = o.nD;
    v.ew = o.ew;
    v.lw = o.lw;
    v.ecol = o.ecol.as<int>();
    v.eval = ev.as<T>();
    v.lcol = o.lcol.as<int>();
    v.lval = lv.as<T>();
    v.ldiag = ld.as<T>();
    v.flag = o.flag.as<uint8_t>();
    v.fmt = o.fmt;
    v.tpc = o.tpc.as<int>();
    v.tpw = tw.as<T>();
    v.exc = o.exc.as<int>();
    v.ecx = o.ecx.as<int>();
    v.xcol = o.xcol.as<int>();
    v.xval = xv.as<T>();
    v.loff = (T)o.loff;
    v.ldu = (T)o.ldu;
    v.tln = o.tln.as<int>();
    v.tlo = o.tlo.as<int>();
    v.tlr = o.tlr.as<int>();
    v.tpl = o.tpl.as<int16_t>();
    v.tlx = o.tlx.as<int16_t>();
    v.ngt = o.ngt;
    v.nitem = o.nitem;
    v.nslot = o.nslot;
    v.gofs = o.gofs.as<int>();
    v.grow = o.grow.as<int>();
    v.gpl = o.gpl.as<int>();
    v.gw = gw.as<T>();
    v.gln = o.gln.as<int>();
    v.glr = o.glr.as<int>();
    v.glx = o.glx.as<int16_t>();
    return v;
}
template <>
OpView<double> Op::view<double>() const { return make_view<double>(*this, eval64, lval64, ldiag64, tpw64, xval64, gw64); }
template <>
OpView<float> Op::view<float>() const { return make_view<float>(*this, eval32, lval32, ldiag32, tpw32, xval32, gw32); }

size_t Op::device_bytes() const {
    return ecol.bytes + eval64.bytes + eval32.bytes + lcol.bytes + lval64.bytes + lval32.bytes + ldiag64.bytes +
           ldiag32.bytes + flag.bytes + perm.bytes + tpc.bytes + tpw64.bytes + tpw32.bytes + exc.bytes + ecx.bytes +
           xcol.bytes + xval64.bytes + xval32.bytes + hrow.bytes + hcol.bytes + hval.bytes + tln.bytes + tlo.bytes +
           tlr.bytes + tpl.bytes + tlx.bytes + gofs.bytes + grow.bytes + gpl.bytes + gw64.bytes + gw32.bytes + gln.bytes +
           glr.bytes + glx.bytes;
}

Op* op_from_csr(Ctx* ctx, int64_t D, int nD, const int32_t* Lp, const int32_t* Lc, const double* Lv,
                const int32_t* Ep, const int32_t* Ec, const double* Ev, const uint8_t* ident, double m, double n,
                int precond, int precond_iters) {
    if (D <= 0 || nD < 0 || nD > D) throw ArgError("invalid D / nD");
    if (D > (int64_t)INT32_MAX) throw ArgError("D exceeds int32 column range");
    std::vector<uint8_t> fl((size_t)D, 0);
    for (int64_t r = 0; r < D; ++r) fl[(size_t)r] = ident[r] ? ROW_IDENTITY : 0;
    // E -> ELL (identity rows are served by the flag)
    int ew = 0;
    for (int64_t r = 0; r < D; ++r)
        if (!(fl[(size_t)r] & ROW_IDENTITY)) ew = std::max(ew, Ep[r + 1] - Ep[r]);
    std::vector<int> ec((size_t)ew * D);
    std::vector<double> ev((size_t)ew * D, 0.0);
    for (int64_t r = 0; r < D; ++r) {
        int s = 0;
        if (!(fl[(size_t)r] & ROW_IDENTITY))
            for (int32_t k = Ep[r]; k < Ep[r + 1]; ++k, ++s) {
                if (Ec[k] < 0 || Ec[k] >= D) throw ArgError("E column out of range");
                ec[(size_t)s * D + r] = Ec[k];
                ev[(size_t)s * D + r] = Ev[k];
            }
        for (; s < ew; ++s) ec[(size_t)s * D + r] = (int)r;   // padding reads the row's own entry with weight 0
    }
    // L -> diagonal + off-diagonal ELL
    int lw = 0;
    for (int64_t r = 0; r < D; ++r) {
        int c = 0;
        for (int32_t k = Lp[r]; k < Lp[r + 1]; ++k) c += (Lc[k] != r);
        lw = std::max(lw, c);
    }
    std::vector<int> lc((size_t)lw * D);
    std::vector<double> lv((size_t)lw * D, 0.0), ld((size_t)D, 0.0);
    for (int64_t r = 0; r < D; ++r) {
        int s = 0;
        bool has_diag = false;
        for (int32_t k = Lp[r]; k < Lp[r + 1]; ++k) {
            if (Lc[k] < 0 || Lc[k] >= D) throw ArgError("L column out of range");
            if (Lc[k] == r) {
                if (!has_diag) ld[(size_t)r] = Lv[k];   // first diagonal entry (GeneralizedBiCGStabSystem.h:321-326)
                has_diag = true;
                continue;
            }
            lc[(size_t)s * D + r] = Lc[k];
            lv[(size_t)s * D + r] = Lv[k];
            ++s;
        }
        for (; s < lw; ++s) lc[(size_t)s * D + r] = (int)r;
        if (!has_diag) fl[(size_t)r] |= ROW_NODIAG;
    }
    std::unique_ptr<Op> op(new Op(ctx, D, nD));
    op->upload_ell(ec, ev, ew, lc, lv, lw, ld, fl);
    // full L rows (diagonal included) of the identity rows
    std::vector<int> hr;
    int hw = 1;
    for (int64_t r = 0; r < D; ++r)
        if (fl[(size_t)r] & ROW_IDENTITY) {
            hr.push_back((int)r);
            hw = std::max(hw, Lp[r + 1] - Lp[r]);
        }
    if (!hr.empty()) {
        const size_t nh = hr.size();
        std::vector<int> hc((size_t)hw * nh);
        std::vector<double> hv((size_t)hw * nh, 0.0);
        for (size_t k = 0; k < nh; ++k) {
            const int r = hr[k];
            int s = 0;
            for (int32_t q = Lp[r]; q < Lp[r + 1]; ++q, ++s) {
                hc[(size_t)s * nh + k] = Lc[q];
                hv[(size_t)s * nh + k] = Lv[q];
            }
            for (; s < hw; ++s) hc[(size_t)s * nh + k] = r;
        }
        auto up = [](Buffer& b, const void* src, size_t bytes) {
            b.~Buffer();
            new (&b) Buffer(bytes);
            HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
        };
        up(op->hrow, hr.data(), nh * sizeof(int));
        up(op->hcol, hc.data(), hc.size() * sizeof(int));
        up(op->hval, hv.data(), hv.size() * sizeof(double));
        op->nh = (int64_t)nh;
        op->hw = hw;
    }
    op->m = m;
    op->n = n;
    op->precond = precond;
    op->precond_iters = precond_iters;
    return op.release();
}

// ---------------------------------------------------------------- batch solve
struct Group {
    Op* op;
    double m, n;             // operator coefficients of this group's solves
    std::vector<int> inst;   // instance indices (<= NR)
};

template <typename T>
static void fill_seg(Seg<T>& s, const Op* op, int64_t off, int nblocks, int nr, KState* st, double* part) {
    s.op = op->view<T>();
    s.op64 = op->view<double>();
    s.m = op->m;
    s.n = op->n;
    s.off = off;
    s.nblocks = nblocks;
    s.nr = nr;
    s.precond = op->precond;
    s.precond_iters = op->precond_iters;
    // FP32 P^-1 multiplies by these (set again if m or n change: fill_seg
    // callers that override m, n use FP64 segments, which divide)
    const float m32 = (float)op->m, n32 = (float)op->n;
    s.dinv = (T)(1.0f / (m32 + (float)op->ldu * n32));
    s.tinv = (T)(1.0f / (float)(op->m - op->n));
    s.st = st;
    s.part = part;
}

// Block maps of a launch over segments of nb[g] blocks each (entries: segment
// | block within it << kSegBits), one after the other: segment by segment,
// which the vector kernels run (a workgroup's rows stream from one region of
// each vector), then round-robin (block b of every segment, then block b +
// 1), which the E gathers run: segments that share an operator (the groups of
// a streamed solve) gather through each tile together, its metadata read
// into an L2 once for all of them rather than once per segment.
static std::vector<int> block_map(const std::vector<int>& nb) {
    if (nb.size() > (size_t)kMaxSegs) throw ArgError("too many segments in one launch");
    int most = 0;
    size_t total = 0;
    for (int v : nb) {
        most = std::max(most, v);
        total += (size_t)v;
    }
    if (most > (1 << (31 - kSegBits))) throw ArgError("operator too large for one launch's block map");
    std::vector<int> m;
    m.reserve(2 * total);
    for (size_t g = 0; g < nb.size(); ++g)   // segment by segment: the vector kernels
        for (int b = 0; b < nb[g]; ++b) m.push_back((int)g | (b << kSegBits));
    for (int b = 0; b < most; ++b)           // round-robin: the gathers
        for (size_t g = 0; g < nb.size(); ++g)
            if (b < nb[g]) m.push_back((int)g | (b << kSegBits));
    return m;
}

// The launches of a run of iterations between two host checks (16 of a
// streamed solve, 4 at a time of a batch solve's chunks) captured once into a
// HIP graph and replayed (CPMGPU_GRAPH=1): the iteration kernels take the same
// arguments every time (segment tables, vectors and states stay in place for
// the whole solve) and record no memset, so every replay ends where the eager
// launches end, bit for bit.
// Replays measured faster than eager launches only for FP32+refine batch
// solves of one group of lanes (one right-hand side on the torus-knot system:
// 28.8 against 30.0 ms per solve); FP64 one-RHS solves were 2% slower and
// streamed solves of 8 and 32 lanes 0.2-0.9% slower (their launches are
// queued 16 iterations ahead), so those replay only with CPMGPU_GRAPH=1.
// CPMGPU_GRAPH=0 turns replays off everywhere.
static bool use_graph(bool by_default = false) {   // read per solve call (not per launch)
    const char* e = std::getenv("CPMGPU_GRAPH");
    if (e && std::string(e) == "1") return true;
    if (e && std::string(e) == "0") return false;
    return by_default;
}

struct IterGraph {
    hipGraphExec_t exec = nullptr;
    IterGraph() = default;
    IterGraph(const IterGraph&) = delete;
    IterGraph& operator=(const IterGraph&) = delete;
    ~IterGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
    template <class F>
    void run(hipStream_t st, F&& launches) {
        if (!exec) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launches();   // kernel launches only: nothing here can throw mid-capture
            HIPCHK(hipStreamEndCapture(st, &g));
            const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(ie);
        }
        HIPCHK(hipGraphLaunch(exec, st));
    }
};

// Workspace bytes run_batch<NR> asks the context for, over G groups of
// `blocks` row blocks and `total` vector elements in all (each vector), and
// their sum as Arena::take rounds them.
struct BatchBytes {
    size_t tables, vec64, vec32, pack;
    size_t arena() const {
        auto r256 = [](size_t n) { return (n + 255) & ~size_t(255); };
        return r256(tables) + r256(vec64) + (vec32 ? r256(vec32) : 0) + r256(pack);
    }
};
static BatchBytes batch_bytes(int G, int blocks, int64_t total, int NR, bool refine) {
    auto r256 = [](size_t n) { return (n + 255) & ~size_t(255); };
    BatchBytes b;
    b.tables = r256(sizeof(Seg<double>) * G) + r256(sizeof(Seg<float>) * G) + r256(sizeof(int) * 2 * (size_t)blocks) +
               r256(sizeof(KState) * (size_t)G * kMaxNR) + r256(sizeof(double) * 2 * NR * (size_t)blocks) + 256;
    b.vec64 = sizeof(double) * (size_t)total * (refine ? 5 : 11) + 1024;
    b.vec32 = refine ? sizeof(float) * (size_t)total * 10 + 1024 : 0;
    b.pack = sizeof(PackEntry) * (size_t)G * NR + sizeof(int) * (size_t)blocks * NR + 512;
    return b;
}
// One operator of D rows with NR right-hand sides.
static BatchBytes batch_bytes1(int64_t D, int NR, bool refine) {
    return batch_bytes(1, (int)((D + kBlock - 1) / kBlock), (D * NR + 63) & ~int64_t(63), NR, refine);
}

// Grow the context's workspaces to what a batch of one operator of D rows
// with NR right-hand sides needs, so that such a batch allocates nothing.
static void reserve_batch_workspaces(Ctx* ctx, int64_t D, int NR, bool refine) {
    const BatchBytes b = batch_bytes1(D, NR, refine);
    ctx->workspace("tables", b.tables);
    ctx->workspace("vec64", b.vec64);
    if (refine) ctx->workspace("vec32", b.vec32);
    ctx->workspace("pack", b.pack);
}

template <int NR>
static void run_batch(Ctx* ctx, const std::vector<Group>& groups, const std::vector<const double*>& rhs,
                      const std::vector<double*>& xs, const cpm_solve_opts& o, std::vector<cpm_solve_info>& infos) {
    const hipStream_t st = ctx->stream;
    const int G = (int)groups.size();
    const bool refine = o.precision == CPM_FP32_REFINE;
    if (ctx->arena) {
        // no-alloc solve: every workspace of this batch is carved from the
        // scratch (an earlier batch of the same call, e.g. the FP64 re-solve's
        // first attempt, has already written its results out); shared-stencil
        // tiles are used only if an earlier call built them
        ctx->arena->used = 0;
    } else if (!ctx->no_builds) {
        std::vector<Op*> ops;
        for (auto& g : groups) ops.push_back(g.op);
        ensure_group_tiles_for(ops, refine ? 4 : 8, NR);   // before the views are taken
    }
    // layout
    std::vector<int64_t> off(G);
    std::vector<int> b0(G), nb(G);
    int64_t total = 0;
    int blocks = 0;
    for (int g = 0; g < G; ++g) {
        off[g] = total;
        total += groups[g].op->D * NR;
        total = (total + 63) & ~int64_t(63);
        b0[g] = blocks;
        nb[g] = (int)((groups[g].op->D + kBlock - 1) / kBlock);
        blocks += nb[g];
    }
    // device tables: segments (one per precision), block map, states, partials.
    // The FP64 and FP32 segments of a group share one state record per rhs.
    const int nstate = G * kMaxNR;
    if (nstate > Ctx::kHostStates) throw ArgError("batch too large");
    const size_t part_stride = 2 * NR;
    auto r256 = [](size_t n) { return (n + 255) & ~size_t(255); };
    const BatchBytes bb = batch_bytes(G, blocks, total, NR, refine);
    const size_t bytes_tables = bb.tables;
    char* tb = reinterpret_cast<char*>(ctx->workspace("tables", bytes_tables));
    char* const tb_end = tb + bytes_tables;
    auto carve = [&](size_t n) {
        char* p = tb;
        tb += r256(n);
        if (tb > tb_end) throw std::logic_error("table workspace overflow");
        return p;
    };
    auto* segs64 = reinterpret_cast<Seg<double>*>(carve(sizeof(Seg<double>) * G));
    auto* segs32 = reinterpret_cast<Seg<float>*>(carve(sizeof(Seg<float>) * G));
    int* bseg = reinterpret_cast<int*>(carve(sizeof(int) * 2 * (size_t)blocks));   // two block maps
    KState* sts = reinterpret_cast<KState*>(carve(sizeof(KState) * nstate));
    double* part = reinterpret_cast<double*>(carve(sizeof(double) * part_stride * blocks));

    std::vector<Seg<double>> h64(G);
    std::vector<Seg<float>> h32(G);
    for (int g = 0; g < G; ++g) {
        const int nr = (int)groups[g].inst.size();
        fill_seg(h64[g], groups[g].op, off[g], nb[g], nr, sts + g * kMaxNR, part + part_stride * b0[g]);
        fill_seg(h32[g], groups[g].op, off[g], nb[g], nr, sts + g * kMaxNR, part + part_stride * b0[g]);
        h64[g].m = h32[g].m = groups[g].m;
        h64[g].n = h32[g].n = groups[g].n;
    }
    const std::vector<int> hb = block_map(nb);
    HIPCHK(hipMemcpyAsync(segs64, h64.data(), sizeof(Seg<double>) * G, hipMemcpyHostToDevice, st));
    if (refine) HIPCHK(hipMemcpyAsync(segs32, h32.data(), sizeof(Seg<float>) * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(bseg, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice, st));

    // workspace vectors, each `total` entries: FP64 b x r r0 p v s t y z tmp;
    // mixed: FP64 b x r t tmp and FP32 x r r0 p v s t y z tmp.
    Vecs<double> V64{};
    Vecs<float> V32{};
    {
        std::vector<double**> f64;
        if (refine) f64 = {&V64.b, &V64.x, &V64.r, &V64.t, &V64.tmp};
        else f64 = {&V64.b, &V64.x, &V64.r, &V64.r0, &V64.p, &V64.v, &V64.s, &V64.t, &V64.y, &V64.z, &V64.tmp};
        double* q = ctx->workspace("vec64", bb.vec64);
        for (auto* pp : f64) { *pp = q; q += total; }
        if (refine) {
            float** f32[] = {&V32.x, &V32.r, &V32.r0, &V32.p, &V32.v, &V32.s, &V32.t, &V32.y, &V32.z, &V32.tmp};
            float* q32 = reinterpret_cast<float*>(ctx->workspace("vec32", bb.vec32));
            for (auto* pp : f32) { *pp = q32; q32 += total; }
        }
    }

    // pack right-hand sides
    std::vector<PackEntry> pe;
    std::vector<int> pb;
    for (int g = 0; g < G; ++g)
        for (int k = 0; k < NR; ++k) {
            PackEntry e{};
            e.src = k < (int)groups[g].inst.size() ? rhs[groups[g].inst[k]] : nullptr;
            e.dst = k < (int)groups[g].inst.size() ? xs[groups[g].inst[k]] : nullptr;
            e.perm = groups[g].op->perm.as<int>();
            e.D = groups[g].op->D;
            e.off = off[g];
            e.lane = k;
            e.nr = refine ? 1 : NR;   // mixed: FP64 vectors lane-major
            e.ls = refine ? e.D : 1;
            e.block0 = (int)pb.size();
            for (int b = 0; b < nb[g]; ++b) pb.push_back((int)pe.size());
            pe.push_back(e);
        }
    if (pe.size() != (size_t)G * NR || pb.size() != (size_t)blocks * NR) throw std::logic_error("pack table size");
    PackEntry* dpe = reinterpret_cast<PackEntry*>(ctx->workspace("pack", bb.pack));
    int* dpb = reinterpret_cast<int*>(reinterpret_cast<char*>(dpe) + ((sizeof(PackEntry) * pe.size() + 255) & ~size_t(255)));
    HIPCHK(hipMemcpyAsync(dpe, pe.data(), sizeof(PackEntry) * pe.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dpb, pb.data(), sizeof(int) * pb.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pack, dim3((unsigned)pb.size()), dim3(256), 0, st, dpe, dpb, V64.b, (float*)nullptr, 1);

    int sweeps = 0;
    for (auto& g : groups)
        if (g.op->precond == CPM_PRECOND_JACOBI) sweeps = std::max(sweeps, g.op->precond_iters);
    int grp = groups.empty() ? 0 : INT_MAX;
    for (auto& g : groups) grp = std::min(grp, g.op->ngt);
    const LaunchCfg<double> c64{segs64, bseg, blocks, G, st, sweeps, grp};
    const LaunchCfg<float> c32{segs32, bseg, blocks, G, st, sweeps, grp};
    KState* hs = ctx->host_states;
    const double delta = o.update_delta > 0 ? std::min(o.update_delta, 0.999) : kDefaultUpdateDelta;
    launch_reset_states(sts, nstate, o.tol * o.tol, o.max_iters, delta * delta, st);

    // Iterations are queued in growing chunks; every kernel skips instances
    // that have stopped, so the host only looks at the states between chunks.
    auto poll = [&](bool& run, bool& restart) {
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hs, sts, sizeof(KState) * nstate, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        run = restart = false;
        for (int g = 0; g < G; ++g)
            for (size_t k = 0; k < groups[g].inst.size(); ++k) {
                const int s = hs[g * kMaxNR + k].status;
                run |= s == ST_RUN || (s >= ST_UPDATE && s <= ST_READY);
                restart |= s == ST_RESTART;
            }
    };
    int chunk = 4;
    bool run = false, restart = false;
    // graph replays (use_graph): the launches of 4 iterations captured once
    // into a HIP graph and replayed chunk / 4 times per chunk (chunks are
    // multiples of 4)
    // (never inside the allocation-free entry point: instantiating a graph allocates)
    const bool graph = !ctx->arena && use_graph(refine && G == 1);
    IterGraph ig;
    if (!refine) {
        Launch<double, NR>::init(c64, V64.b, V64.x, V64.r, V64.r0, V64.p, V64.v);
        auto four = [&] {
            for (int c = 0; c < 4; ++c) Launch<double, NR>::iteration(c64, V64, false, MixVec{});
        };
        for (;;) {
            for (int c = 0; c < chunk; c += 4) {
                if (graph) ig.run(st, four);
                else four();
            }
            poll(run, restart);
            if (restart) {   // rare: handled between chunks, the other instances wait
                Launch<double, NR>::restart(c64, V64);
                continue;
            }
            if (!run) break;
            chunk = std::min(chunk * 2, 64);
        }
    } else {
        double* x64 = V64.x;
        double* r64 = V64.r;
        MixedLaunch<NR>::start(c64, V64.b, x64, r64);
        MixedLaunch<NR>::down(c64, r64, V32);
        const MixVec mv{x64, r64, V64.b, V64.tmp};
        auto four = [&] {
            for (int c = 0; c < 4; ++c) Launch<float, NR>::iteration(c32, V32, true, mv);
        };
        for (;;) {
            for (int c = 0; c < chunk; c += 4) {
                if (graph) ig.run(st, four);
                else four();
            }
            poll(run, restart);
            if (!run) break;
            chunk = std::min(chunk * 2, 64);
        }
    }
    hipLaunchKernelGGL(k_pack, dim3((unsigned)pb.size()), dim3(256), 0, st, dpe, dpb, V64.x, (float*)nullptr, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    for (int g = 0; g < G; ++g)
        for (size_t k = 0; k < groups[g].inst.size(); ++k) {
            const KState& s = hs[g * kMaxNR + k];
            cpm_solve_info& inf = infos[groups[g].inst[k]];
            inf.status = s.status;
            inf.iters = s.iters;
            inf.restarts = s.restarts;
            inf.updates = refine ? s.updates : 0;
            const double rr = refine ? s.rr_true : s.rr;
            inf.rel_residual = s.bsq > 0 ? std::sqrt(rr / s.bsq) : 0.0;
        }
}

// Continuous batching (FP32 + refinement): n systems of one operator through
// G x NR interleaved lanes (G segments of NR lanes of the same operator in
// every launch).  At every poll a lane whose system stopped hands its solution
// out and takes the next right-hand side: b64 <- new b, x64 = x32 = 0 and a
// fresh state in the refinement status, whose step then computes
// ||b - A 0||^2 = ||b||^2 and restarts the lane's Krylov space (mx_update).
// Lanes never wait for the slowest system of a batch.
template <int NR>
static void run_stream(Ctx* ctx, Op* op, const std::vector<const double*>& rhs, const std::vector<double*>& xs,
                       int G, const cpm_solve_opts& o, std::vector<cpm_solve_info>& infos) {
    const hipStream_t st = ctx->stream;
    const int n = (int)rhs.size();
    G = std::max(1, std::min(G, (n + NR - 1) / NR));   // no segment without a first system
    const int64_t D = op->D;
    const int nb = (int)((D + kBlock - 1) / kBlock);
    const int blocks = G * nb;
    const int nstate = G * kMaxNR;
    if (nstate > Ctx::kHostStates) throw ArgError("too many lanes");
    ensure_group_tiles_for({op}, o.precision == CPM_FP32_REFINE ? 4 : 8, NR);
    const int64_t seg_total = (D * NR + 63) & ~int64_t(63);   // vector entries of one segment
    const int64_t total = seg_total * G;
    auto r256 = [](size_t k) { return (k + 255) & ~size_t(255); };
    const size_t part_stride = 2 * NR;
    const size_t bytes = r256(sizeof(Seg<double>) * G) + r256(sizeof(Seg<float>) * G) + 3 * r256(sizeof(int) * blocks) +
                         r256(sizeof(KState) * nstate) + r256(sizeof(double) * part_stride * blocks) +
                         r256(sizeof(PackEntry) * 4) + 256;
    char* tb = reinterpret_cast<char*>(ctx->workspace("stream_tables", bytes));
    auto carve = [&](size_t k) { char* p = tb; tb += r256(k); return p; };
    auto* seg64 = reinterpret_cast<Seg<double>*>(carve(sizeof(Seg<double>) * G));
    auto* seg32 = reinterpret_cast<Seg<float>*>(carve(sizeof(Seg<float>) * G));
    int* bseg = reinterpret_cast<int*>(carve(sizeof(int) * 2 * (size_t)blocks));   // two block maps
    int* bzero = reinterpret_cast<int*>(carve(sizeof(int) * blocks));   // all zero: one pack entry
    KState* sts = reinterpret_cast<KState*>(carve(sizeof(KState) * nstate));
    double* part = reinterpret_cast<double*>(carve(sizeof(double) * part_stride * blocks));
    PackEntry* dpe = reinterpret_cast<PackEntry*>(carve(sizeof(PackEntry) * 4));
    std::vector<int> nr0(G);
    std::vector<Seg<double>> h64(G);
    std::vector<Seg<float>> h32(G);
    std::vector<int> hz(blocks, 0);
    for (int g = 0; g < G; ++g) {
        nr0[g] = std::min(NR, n - g * NR);
        fill_seg(h64[g], op, g * seg_total, nb, nr0[g], sts + g * kMaxNR, part + part_stride * g * nb);
        fill_seg(h32[g], op, g * seg_total, nb, nr0[g], sts + g * kMaxNR, part + part_stride * g * nb);
    }
    const std::vector<int> hb = block_map(std::vector<int>((size_t)G, nb));
    HIPCHK(hipMemcpyAsync(seg64, h64.data(), sizeof(Seg<double>) * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(seg32, h32.data(), sizeof(Seg<float>) * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(bseg, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(bzero, hz.data(), sizeof(int) * blocks, hipMemcpyHostToDevice, st));
    Vecs<double> V64{};
    Vecs<float> V32{};
    {
        double* q = ctx->workspace("vec64", sizeof(double) * total * 5 + 1024);
        for (double** pp : {&V64.b, &V64.x, &V64.r, &V64.t, &V64.tmp}) { *pp = q; q += total; }
        float* q32 = reinterpret_cast<float*>(ctx->workspace("vec32", sizeof(float) * total * 10 + 1024));
        for (float** pp : {&V32.x, &V32.r, &V32.r0, &V32.p, &V32.v, &V32.s, &V32.t, &V32.y, &V32.z, &V32.tmp}) {
            *pp = q32;
            q32 += total;
        }
    }
    // one lane of one vector in or out: a single pack entry with block0 = 0;
    // FP64 vectors are lane-major, FP32 ones interleaved, within the lane's segment
    auto lane_pack = [&](int g, int k, const double* src, double* dst, double* ws64, float* ws32, int to_ws) {
        PackEntry e{};
        e.src = src;
        e.dst = dst;
        e.perm = op->perm.as<int>();
        e.D = D;
        e.off = g * seg_total;
        e.lane = k;
        e.nr = ws64 ? 1 : NR;
        e.ls = ws64 ? D : 1;
        e.block0 = 0;
        HIPCHK(hipMemcpyAsync(dpe, &e, sizeof(e), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, st, dpe, bzero, ws64, ws32, to_ws);
        HIPCHK(hipStreamSynchronize(st));   // the entry is rewritten by the next call
    };
    // all lanes start from zero vectors (padding lanes stay zero and idle)
    HIPCHK(hipMemsetAsync(V64.b, 0, sizeof(double) * total * 5, st));
    HIPCHK(hipMemsetAsync(V32.x, 0, sizeof(float) * total * 10, st));
    const double delta = o.update_delta > 0 ? std::min(o.update_delta, 0.999) : kDefaultUpdateDelta;
    KState fresh{};
    fresh.tol2 = o.tol * o.tol;
    fresh.max_it = o.max_iters;
    fresh.delta2 = delta * delta;
    fresh.status = ST_UPDATE;
    fresh.bsq = -1.0;       // learnt by the first refinement step
    fresh.updates = -1;     // that step is not a refinement
    fresh.scale = 1.0;
    KState idle{};
    idle.status = ST_CONVERGED;
    std::vector<KState> hst(nstate, idle);
    std::vector<int> cur(nstate, -1);   // system of lane (g, k) at g * kMaxNR + k
    int next = 0;
    for (int g = 0; g < G; ++g)
        for (int k = 0; k < nr0[g]; ++k) {
            lane_pack(g, k, rhs[next], nullptr, V64.b, nullptr, 1);
            hst[g * kMaxNR + k] = fresh;
            cur[g * kMaxNR + k] = next++;
        }
    HIPCHK(hipMemcpyAsync(sts, hst.data(), sizeof(KState) * nstate, hipMemcpyHostToDevice, st));
    int sweeps = op->precond == CPM_PRECOND_JACOBI ? op->precond_iters : 0;
    const LaunchCfg<float> c32{seg32, bseg, blocks, G, st, sweeps, op->ngt};
    const MixVec mv{V64.x, V64.r, V64.b, V64.tmp};
    KState* hs = ctx->host_states;
    int done = 0;
    const bool graph = use_graph();
    IterGraph ig;
    auto iterations = [&] {
        for (int c = 0; c < 16; ++c) Launch<float, NR>::iteration(c32, V32, true, mv);
    };
    while (done < n) {
        if (graph) ig.run(st, iterations);
        else iterations();
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hs, sts, sizeof(KState) * nstate, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        bool changed = false;
        for (int g = 0; g < G; ++g)
            for (int k = 0; k < nr0[g]; ++k) {
                const int L = g * kMaxNR + k;
                if (cur[L] < 0) continue;
                const KState& s = hs[L];
                const bool running = s.status == ST_RUN || (s.status >= ST_UPDATE && s.status <= ST_READY);
                if (running) continue;
                cpm_solve_info& inf = infos[(size_t)cur[L]];
                inf.status = s.status;
                inf.iters = s.iters;
                inf.restarts = s.restarts;
                inf.updates = s.updates;
                inf.rel_residual = s.bsq > 0 ? std::sqrt(s.rr_true / s.bsq) : 0.0;
                lane_pack(g, k, nullptr, xs[(size_t)cur[L]], V64.x, nullptr, 0);
                ++done;
                cur[L] = -1;
                KState ns = idle;
                if (next < n) {
                    lane_pack(g, k, rhs[(size_t)next], nullptr, V64.b, nullptr, 1);
                    lane_pack(g, k, nullptr, nullptr, V64.x, nullptr, 1);   // x64 = 0
                    lane_pack(g, k, nullptr, nullptr, nullptr, V32.x, 1);   // x32 = 0
                    ns = fresh;
                    cur[L] = next++;
                }
                HIPCHK(hipMemcpyAsync(sts + L, &ns, sizeof(KState), hipMemcpyHostToDevice, st));
                changed = true;
            }
        if (changed) HIPCHK(hipStreamSynchronize(st));
    }
}

// FP32_REFINE systems whose refinement stopped reducing the true residual
// (ST_FALLBACK) are solved again in FP64 from zero, the reference algorithm,
// in one batch; their infos keep the FP32 counts and add fallback_iters.
// FP64 routing (Fp32Record::prefer_fp64, route.h): FP32+refine solves of an
// operator start in FP64 when its record (or, for a tensor-product operator
// without one, its family's) says FP64 is cheaper (CPMGPU_FP64_ROUTING=0:
// always try FP32 first).
static bool fp64_routing() {   // read per solve call (not per launch)
    const char* e = std::getenv("CPMGPU_FP64_ROUTING");
    return !(e && std::string(e) == "0");
}

static std::array<double, 4> fp32_family_key(const Op* op) { return {op->m, op->n, op->loff, op->ldu}; }

// Whether FP32+refine solves of op on ctx start in FP64: by the operator's own
// record once it has one, else (the tensor-product format) by its family's.
static bool prefer_fp64(Ctx* ctx, const Op* op) {
    if (op->fp32.solved > 0 || op->fmt != OP_TP) return op->fp32.prefer_fp64();
    const auto it = ctx->fp32_family.find(fp32_family_key(op));
    return it != ctx->fp32_family.end() && it->second.prefer_fp64();
}

// Adds the outcome of FP32+refine systems solved on ctx (infos after the FP64
// re-solves) to their operators' records and their families'.
static void record_fp32(Ctx* ctx, const std::vector<Op*>& ops, const std::vector<cpm_solve_info>& infos) {
    for (size_t i = 0; i < infos.size(); ++i) {
        ops[i]->fp32.add(infos[i].iters, infos[i].fallback_iters);
        if (ops[i]->fmt == OP_TP) ctx->fp32_family[fp32_family_key(ops[i])].add(infos[i].iters, infos[i].fallback_iters);
    }
}

static void fp64_fallback(Ctx* ctx, const std::vector<Op*>& ops, const std::vector<const double*>& rhs,
                          const std::vector<double*>& xs, const cpm_solve_opts& o,
                          std::vector<cpm_solve_info>& infos, const std::vector<std::pair<double, double>>* mn) {
    std::vector<int> idx;
    for (size_t i = 0; i < infos.size(); ++i)
        if (infos[i].status == ST_FALLBACK) idx.push_back((int)i);
    if (idx.empty()) return;
    std::vector<Op*> fo;
    std::vector<const double*> fr;
    std::vector<double*> fx;
    std::vector<std::pair<double, double>> fmn;
    for (int i : idx) {
        fo.push_back(ops[(size_t)i]);
        fr.push_back(rhs[(size_t)i]);
        fx.push_back(xs[(size_t)i]);
        fmn.push_back(mn ? (*mn)[(size_t)i] : std::make_pair(ops[(size_t)i]->m, ops[(size_t)i]->n));
    }
    cpm_solve_opts o64 = o;
    o64.precision = CPM_FP64;
    std::vector<cpm_solve_info> fi;
    solve_batch(ctx, fo, fr, fx, o64, fi, &fmn);
    for (size_t j = 0; j < idx.size(); ++j) {
        cpm_solve_info& inf = infos[(size_t)idx[j]];
        inf.status = fi[j].status;
        inf.restarts += fi[j].restarts;
        inf.fallback_iters = fi[j].iters;
        inf.rel_residual = fi[j].rel_residual;
    }
}

void solve_batch(Ctx* ctx, const std::vector<Op*>& ops, const std::vector<const double*>& rhs,
                 const std::vector<double*>& xs, const cpm_solve_opts& o, std::vector<cpm_solve_info>& infos,
                 const std::vector<std::pair<double, double>>* mn) {
    const int n = (int)ops.size();
    infos.assign(n, cpm_solve_info{});
    if (n == 0) return;
    if (o.precision != CPM_FP64 && o.precision != CPM_FP32_REFINE) throw ArgError("unknown precision");
    for (Op* op : ops) {
        // operators of another context on the same device are complete (their builds end with a
        // stream synchronisation) and are solved on this context's stream
        if (!op || op->ctx->device != ctx->device) throw ArgError("operator lives on another device");
        if (op->precond == CPM_PRECOND_JACOBI && op->precond_iters > 100000) throw ArgError("too many Jacobi sweeps");
    }
    if (o.precision == CPM_FP32_REFINE && fp64_routing()) {
        // operators whose FP32+refine solves fell back to FP64 before: FP64 from the start
        std::vector<int> r64, r32;
        for (int i = 0; i < n; ++i) (prefer_fp64(ctx, ops[i]) ? r64 : r32).push_back(i);
        if (!r64.empty()) {
            cpm_solve_opts o64 = o;
            o64.precision = CPM_FP64;
            for (int pass = 0; pass < 2; ++pass) {
                const std::vector<int>& sel = pass == 0 ? r64 : r32;
                if (sel.empty()) continue;
                std::vector<Op*> so;
                std::vector<const double*> sr;
                std::vector<double*> sx;
                std::vector<std::pair<double, double>> smn;
                for (int i : sel) {
                    so.push_back(ops[i]);
                    sr.push_back(rhs[(size_t)i]);
                    sx.push_back(xs[(size_t)i]);
                    smn.push_back(mn ? (*mn)[(size_t)i] : std::make_pair(ops[i]->m, ops[i]->n));
                }
                std::vector<cpm_solve_info> si;
                solve_batch(ctx, so, sr, sx, pass == 0 ? o64 : o, si, &smn);
                for (size_t j = 0; j < sel.size(); ++j) {
                    cpm_solve_info& inf = infos[(size_t)sel[j]];
                    inf = si[j];
                    if (pass == 0) {   // reported as an FP64 re-solve without FP32 iterations
                        inf.fallback_iters = si[j].iters;
                        inf.iters = 0;
                    }
                }
            }
            return;
        }
    }
    // group instances that share an operator and its (m, n): at most kMaxNR per group, at most 4
    // in FP64 where the operator has LDS tiles for FP64 x 4 (FP64 x 8 rows do not fit them)
    std::vector<Group> groups;
    std::map<std::tuple<Op*, double, double>, int> open;
    for (int i = 0; i < n; ++i) {
        const double m = mn ? (*mn)[(size_t)i].first : ops[i]->m;
        const double nn = mn ? (*mn)[(size_t)i].second : ops[i]->n;
        const auto key = std::make_tuple(ops[i], m, nn);
        auto it = open.find(key);
        const size_t cap = (o.precision == CPM_FP64 && ops[i]->fmt == OP_TP && (ops[i]->ngt > 0 || ops[i]->tln.p))
                               ? 4 : (size_t)kMaxNR;
        if (it == open.end() || groups[it->second].inst.size() == cap) {
            groups.push_back(Group{ops[i], m, nn, {}});
            open[key] = (int)groups.size() - 1;
            it = open.find(key);
        }
        groups[it->second].inst.push_back(i);
    }
    size_t nr = 1;
    for (auto& g : groups) nr = std::max(nr, g.inst.size());
    if (nr == 1) run_batch<1>(ctx, groups, rhs, xs, o, infos);
    else if (nr == 2) run_batch<2>(ctx, groups, rhs, xs, o, infos);
    else if (nr <= 4) run_batch<4>(ctx, groups, rhs, xs, o, infos);
    else run_batch<8>(ctx, groups, rhs, xs, o, infos);
    if (o.precision == CPM_FP32_REFINE) {
        fp64_fallback(ctx, ops, rhs, xs, o, infos, mn);
        record_fp32(ctx, ops, infos);
    }
}

static void solve_stream(Op* op, const std::vector<const double*>& rhs, const std::vector<double*>& xs, int lanes,
                         const cpm_solve_opts& o, std::vector<cpm_solve_info>& infos) {
    infos.assign(rhs.size(), cpm_solve_info{});
    if (rhs.empty()) return;
    if (o.precision != CPM_FP32_REFINE) throw ArgError("streamed solves run in FP32_REFINE precision");
    if (op->nD != 0) throw ArgError("streamed solves need an operator without trailing nD rows");
    if (op->precond == CPM_PRECOND_JACOBI && op->precond_iters > 100000) throw ArgError("too many Jacobi sweeps");
    if (prefer_fp64(op->ctx, op) && fp64_routing()) {   // see solve_batch
        solve_batch(op->ctx, std::vector<Op*>(rhs.size(), op), rhs, xs, o, infos, nullptr);
        return;
    }
    if (lanes == 1) run_stream<1>(op->ctx, op, rhs, xs, 1, o, infos);
    else if (lanes == 2) run_stream<2>(op->ctx, op, rhs, xs, 1, o, infos);
    else if (lanes == 4) run_stream<4>(op->ctx, op, rhs, xs, 1, o, infos);
    else if (lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64) run_stream<8>(op->ctx, op, rhs, xs, lanes / 8, o, infos);
    else throw ArgError("lanes must be 1, 2, 4, 8, 16, 32 or 64");
    const std::vector<Op*> ops(rhs.size(), op);
    fp64_fallback(op->ctx, ops, rhs, xs, o, infos, nullptr);
    record_fp32(op->ctx, ops, infos);
}

void multiply(Op* op, const double* dx, double* dy, double m, double n) {
    Ctx* ctx = op->ctx;
    const hipStream_t st = ctx->stream;
    const int nblocks = (int)((op->D + kBlock - 1) / kBlock);
    char* tb = reinterpret_cast<char*>(ctx->workspace("mul_tables", 8192 + sizeof(int) * nblocks));
    Seg<double> h{};
    fill_seg(h, op, 0, nblocks, 1, nullptr, nullptr);
    h.m = m;
    h.n = n;
    std::vector<int> hb(nblocks, 0);
    PackEntry pe[2]{};
    pe[0].src = dx;
    pe[1].dst = dy;
    for (auto& e : pe) {
        e.perm = op->perm.as<int>();
        e.D = op->D;
        e.nr = 1;
    }
    auto* dseg = reinterpret_cast<Seg<double>*>(tb);
    auto* dpe = reinterpret_cast<PackEntry*>(tb + 4096);
    int* dbseg = reinterpret_cast<int*>(tb + 8192);
    HIPCHK(hipMemcpyAsync(dseg, &h, sizeof(h), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dpe, pe, sizeof(pe), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dbseg, hb.data(), sizeof(int) * nblocks, hipMemcpyHostToDevice, st));
    double* w = ctx->workspace("mul_tmp", sizeof(double) * op->D * 3);
    double *xi = w, *yi = w + op->D, *tmp = w + 2 * op->D;
    LaunchCfg<double> c{dseg, dbseg, nblocks, 1, st, 0, op->ngt};
    if (op->perm.p) {   // caller order -> internal order and back
        hipLaunchKernelGGL(k_pack, dim3(nblocks), dim3(256), 0, st, dpe, dbseg, xi, (float*)nullptr, 1);
        Launch<double, 1>::apply(c, xi, tmp, yi);
        hipLaunchKernelGGL(k_pack, dim3(nblocks), dim3(256), 0, st, dpe + 1, dbseg, yi, (float*)nullptr, 0);
    } else {
        Launch<double, 1>::apply(c, dx, tmp, dy);
    }
    HIPCHK(hipGetLastError());
}

template <typename T, int NR>
static void profile_egather_t(Op* op, int G, int reps, double* avg_ms, double* bytes) {
    ensure_group_tiles_for({op}, sizeof(T), NR);
    Ctx* ctx = op->ctx;
    const hipStream_t st = ctx->stream;
    const int nblocks = (int)((op->D + kBlock - 1) / kBlock);
    const int64_t seg_total = (op->D * NR + 63) & ~int64_t(63);
    std::vector<Seg<T>> h(G);
    std::vector<int> hb((size_t)G * nblocks);
    for (int g = 0; g < G; ++g) {   // G copies of the operator in one launch, as streamed solves of 16, 32 lanes
        fill_seg(h[g], op, g * seg_total, nblocks, NR, nullptr, nullptr);
    }
    hb = block_map(std::vector<int>((size_t)G, nblocks));   // both maps (the gather runs the second)
    const size_t seg_bytes = (sizeof(Seg<T>) * G + 255) & ~size_t(255);
    char* tb = reinterpret_cast<char*>(ctx->workspace("prof_tables", seg_bytes + sizeof(int) * hb.size()));
    auto* dseg = reinterpret_cast<Seg<T>*>(tb);
    int* dbseg = reinterpret_cast<int*>(tb + seg_bytes);
    HIPCHK(hipMemcpyAsync(dseg, h.data(), sizeof(Seg<T>) * G, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dbseg, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice, st));
    // x and out rotate over enough copies that a launch finds none of its vectors in the 256 MB
    // last-level cache (as inside a solve, where every iteration streams ~10 such vectors)
    const size_t vec_bytes = sizeof(T) * (size_t)seg_total * G;
    const int nbuf = (int)std::max<size_t>(1, std::min<size_t>(8, (size_t(640) << 20) / (2 * vec_bytes) + 1));
    T* xs = reinterpret_cast<T*>(ctx->workspace("prof_x", vec_bytes * nbuf));
    T* ys = reinterpret_cast<T*>(ctx->workspace("prof_y", vec_bytes * nbuf));
    HIPCHK(hipMemsetAsync(xs, 0, vec_bytes * nbuf, st));
    LaunchCfg<T> c{dseg, dbseg, nblocks * G, G, st, 0, op->ngt};
    auto buf = [&](T* b, int r) { return b + (size_t)(r % nbuf) * seg_total * G; };
    // the gather the iterations launch: of P^-1 x unless Jacobi sweeps follow the first
    const bool fly = !(op->precond == CPM_PRECOND_JACOBI && op->precond_iters > 1);
    launch_egather<T, NR>(c, buf(xs, 0), buf(ys, 0), fly);   // warm-up
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch_egather<T, NR>(c, buf(xs, r + 1), buf(ys, r + 1), fly);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<uint8_t> fl((size_t)op->D);
    HIPCHK(hipMemcpy(fl.data(), op->flag.p, fl.size(), hipMemcpyDeviceToHost));
    int64_t gathered = 0;
    for (uint8_t f : fl) gathered += !(f & ROW_IDENTITY);
    *avg_ms = ms / reps;
    // algorithmic bytes: operator data of the gathered rows, x read once, out written once, flags
    // (tensor-product rows: 12 weights, the exception index, and either 16
    // int16 line positions in the tile's list (LDS tiles) or 16 int32 line
    // starts; exception rows 72 int16 positions or 64 int32 columns, plus the
    // extra entry; LDS tiles also read their lists, 4 bytes per row)
    double opb;
    const double sT = sizeof(T);
    if (op->fmt == OP_TP && op->ngt > 0 && grp_gather(sizeof(T), NR, op->ngt) && GrpTile<T, NR>::kRows > 0) {
        // shared-stencil tiles: per plain item 16 int16 positions and its
        // slot offset, per plain row 12 weights and its row index; exception
        // rows as in the row tiles plus slot offset and row index; identity
        // items their slot offset, identity rows their row index; 4 bytes per
        // list row and 12 per tile
        opb = (double)op->grp_items0 * (16 * 2.0 + 4) + (double)op->grp_rows0 * (12 * sT + 4) +
              (double)op->grp_exc * (72 * 2.0 + 12 * sT + 4 + sT + 8) +
              (double)op->grp_ident_items * 4.0 + (double)op->grp_ident * 4.0 + 4.0 * (double)op->grp_list +
              12.0 * (double)op->ngt;
        *avg_ms = ms / reps;
        *bytes = G * (opb + 2.0 * op->D * NR * sizeof(T));
        return;
    }
    if (op->fmt == OP_TP && op->tln.p) {
        const int64_t* n = op->tile_rows;
        opb = (double)n[0] * (16 * 2.0 + 12 * sT + 4) + (double)n[1] * (72 * 2.0 + 12 * sT + 4 + sT) +
              (double)n[2] * (16 * 4.0 + 12 * sT + 4) + (double)n[3] * (64 * 4.0 + 12 * sT + 8 + sT) +
              4.0 * (double)op->tile_list + 8.0 * (double)((op->D + kBlock - 1) / kBlock);
    } else if (op->fmt == OP_TP) {
        opb = (double)gathered * (16 * 4.0 + 12.0 * sizeof(T) + 4.0) + (double)op->nexc * (64 * 4.0 + 4 + sizeof(T));
    } else {
        opb = (double)gathered * op->ew * (4.0 + sizeof(T));
    }
    *bytes = G * (opb + 2.0 * op->D * NR * sizeof(T) + (double)op->D);
}

// Identity rows of a caller-order vector y take the values of g.
__global__ void k_set_identity_rows(const uint8_t* flag, const int* perm, int64_t D, const double* g, double* y) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= D || !(flag[r] & ROW_IDENTITY)) return;
    const int64_t cr = perm ? perm[r] : r;
    y[cr] = g[cr];
}

// z = x + dt (Lfull x) on the identity rows, from the operator's held Lfull rows.
__global__ void k_held_apply(int64_t nh, int hw, const int* hrow, const int* hcol, const double* hval,
                             const uint8_t* flag, double dt, const double* x, double* z) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nh) return;
    const int r = hrow[k];
    if (!(flag[r] & ROW_IDENTITY)) return;
    double acc = 0.0;
    for (int s = 0; s < hw; ++s) acc += hval[(int64_t)s * nh + k] * x[hcol[(int64_t)s * nh + k]];
    z[r] = x[r] + dt * acc;
}

// Explicit stepping in the operator's internal order: one segment per
// diffusion coefficient c (apply with m = 1, n = c) and pack entries moving
// caller vectors in and out, uploaded together.
struct StepTables {
    Seg<double>* seg = nullptr;
    PackEntry* pe = nullptr;
    int* bseg = nullptr;
    int nblocks = 0;
};

static StepTables step_tables(Op* op, const char* name, const std::vector<double>& coef,
                              std::vector<PackEntry> pes) {
    Ctx* ctx = op->ctx;
    const hipStream_t st = ctx->stream;
    StepTables t;
    t.nblocks = (int)((op->D + kBlock - 1) / kBlock);
    const size_t seg_bytes = (sizeof(Seg<double>) * coef.size() + 255) / 256 * 256;
    const size_t pe_bytes = (sizeof(PackEntry) * pes.size() + 255) / 256 * 256;
    char* tb = reinterpret_cast<char*>(ctx->workspace(name, seg_bytes + pe_bytes + sizeof(int) * t.nblocks));
    std::vector<Seg<double>> h(coef.size());
    for (size_t k = 0; k < coef.size(); ++k) {
        fill_seg(h[k], op, 0, t.nblocks, 1, nullptr, nullptr);
        h[k].m = 1.0;
        h[k].n = coef[k];
    }
    for (auto& e : pes) {
        e.perm = op->perm.as<int>();
        e.D = op->D;
        e.nr = 1;
    }
    std::vector<int> hb(t.nblocks, 0);
    t.seg = reinterpret_cast<Seg<double>*>(tb);
    t.pe = reinterpret_cast<PackEntry*>(tb + seg_bytes);
    t.bseg = reinterpret_cast<int*>(tb + seg_bytes + pe_bytes);
    HIPCHK(hipMemcpyAsync(t.seg, h.data(), sizeof(Seg<double>) * h.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(t.pe, pes.data(), sizeof(PackEntry) * pes.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(t.bseg, hb.data(), sizeof(int) * t.nblocks, hipMemcpyHostToDevice, st));
    return t;
}

static PackEntry pack_in(const double* src) { PackEntry e{}; e.src = src; return e; }
static PackEntry pack_out(double* dst) { PackEntry e{}; e.dst = dst; return e; }

// z = x + c (Lfull x), internal order, on every row: the tube rows through the
// apply kernel with tmp = x, the identity rows from the held Lfull rows.
static void explicit_diffusion(Op* op, const StepTables& t, int k, double c, const double* x, double* z) {
    const hipStream_t st = op->ctx->stream;
    LaunchCfg<double> cfg{t.seg + k, t.bseg, t.nblocks, 1, st, 0};
    launch_lfinish_plain<double>(cfg, x, x, z);
    if (op->nh > 0)
        hipLaunchKernelGGL(k_held_apply, dim3((unsigned)((op->nh + 255) / 256)), dim3(256), 0, st, op->nh, op->hw,
                           op->hrow.as<int>(), op->hcol.as<int>(), op->hval.as<double>(), op->flag.as<uint8_t>(), c,
                           x, z);
}

static void extension(Op* op, const StepTables& t, const double* z, double* y) {
    LaunchCfg<double> cfg{t.seg, t.bseg, t.nblocks, 1, op->ctx->stream, 0};
    launch_egather<double, 1>(cfg, z, y);
}

void apply_E(Op* op, const double* x, double* y) {
    const hipStream_t st = op->ctx->stream;
    const StepTables t = step_tables(op, "applyE_tables", {0.0}, {pack_in(x), pack_out(y)});
    double* w = op->ctx->workspace("applyE_vecs", sizeof(double) * op->D * 2);
    hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe, t.bseg, w, (float*)nullptr, 1);
    extension(op, t, w, w + op->D);
    hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe + 1, t.bseg, w + op->D, (float*)nullptr, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));   // the tables are rewritten by the next call
}

static void set_dirichlet(Op* op, const double* g, double* u) {
    hipLaunchKernelGGL(k_set_identity_rows, dim3((unsigned)((op->D + 255) / 256)), dim3(256), 0, op->ctx->stream,
                       op->flag.as<uint8_t>(), op->perm.as<int>(), op->D, g, u);
}

// HeatWithIBC::TimeStep (problems/HeatWithIBC.cpp:80-103).
static void heat_step(Op* op, int method, double dt, const double* d_g, double* d_u, const cpm_solve_opts& o,
                      cpm_solve_info* info) {
    Ctx* ctx = op->ctx;
    const hipStream_t st = ctx->stream;
    const int64_t D = op->D;
    if (method == 0) {
        // internal order: x = P u ; z = x + dt Lfull x ; y = E z ; u = P^T y ; Dirichlet rows
        const StepTables t = step_tables(op, "heat_tables", {dt}, {pack_in(d_u), pack_out(d_u)});
        double* w = ctx->workspace("heat_vecs", sizeof(double) * D * 3);
        double *xi = w, *zi = w + D, *yi = w + 2 * D;
        hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe, t.bseg, xi, (float*)nullptr, 1);
        explicit_diffusion(op, t, 0, dt, xi, zi);
        extension(op, t, zi, yi);
        hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe + 1, t.bseg, yi, (float*)nullptr, 0);
        set_dirichlet(op, d_g, d_u);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        if (info) *info = cpm_solve_info{CPM_SOLVE_CONVERGED, 0, 0, 0, 0.0};
        return;
    }
    double* rhs = ctx->workspace("heat_rhs", sizeof(double) * D);
    if (method == 2) multiply(op, d_u, rhs, 1.0, 0.5 * dt);
    else HIPCHK(hipMemcpyAsync(rhs, d_u, sizeof(double) * D, hipMemcpyDeviceToDevice, st));
    set_dirichlet(op, d_g, rhs);
    HIPCHK(hipGetLastError());
    std::vector<cpm_solve_info> infos;
    solve_batch(ctx, {op}, {rhs}, {d_u}, o, infos);
    if (info) *info = infos[0];
}

// Gray-Scott reaction added to the diffused fields (ReactionDiffusionWithIBC.cpp:89-93):
// zu += dt (-u v^2 + F (1 - u)), zv += dt (u v^2 - (F + k) v), from the old u, v.
__global__ void k_gray_scott(int64_t D, double F, double k, double dt, const double* u, const double* v, double* zu,
                             double* zv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const double a = u[i], b = v[i];
    zu[i] += dt * (-a * b * b + F * (1.0 - a));
    zv[i] += dt * (a * b * b - (F + k) * b);
}

// IMEX right-hand sides: ru = u + dt Ru(u, v), rv = v + dt Rv(u, v).
__global__ void k_gray_scott_rhs(int64_t D, double F, double k, double dt, const double* u, const double* v,
                                 double* ru, double* rv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const double a = u[i], b = v[i];
    ru[i] = a + dt * (-a * b * b + F * (1.0 - a));
    rv[i] = b + dt * (a * b * b - (F + k) * b);
}

// ReactionDiffusionWithIBC::TimeStep (problems/ReactionDiffusionWithIBC.cpp:79-104)
// for nf fields (u, v pairs); method 1 treats diffusion implicitly instead:
// (1 - dt nu L) u' = u + dt Ru(u, v), one batched solve for all 2 nf fields.
static void rd_step(Op* op, const cpm_rd_params& p, int nf, double* const* d_u, double* const* d_v,
                    const double* d_gu, const double* d_gv, const cpm_solve_opts& o, cpm_solve_info* infos) {
    Ctx* ctx = op->ctx;
    const hipStream_t st = ctx->stream;
    const int64_t D = op->D;
    const unsigned nb = (unsigned)((D + 255) / 256);
    if (p.method == 0) {
        double* w = ctx->workspace("rd_vecs", sizeof(double) * D * 4);
        double *xu = w, *xv = w + D, *zu = w + 2 * D, *zv = w + 3 * D;
        for (int f = 0; f < nf; ++f) {
            const StepTables t = step_tables(op, "rd_tables", {p.nu_u * p.dt, p.nu_v * p.dt},
                                             {pack_in(d_u[f]), pack_in(d_v[f]), pack_out(d_u[f]), pack_out(d_v[f])});
            hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe, t.bseg, xu, (float*)nullptr, 1);
            hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe + 1, t.bseg, xv, (float*)nullptr, 1);
            explicit_diffusion(op, t, 0, p.nu_u * p.dt, xu, zu);
            explicit_diffusion(op, t, 1, p.nu_v * p.dt, xv, zv);
            hipLaunchKernelGGL(k_gray_scott, dim3(nb), dim3(256), 0, st, D, p.F, p.k, p.dt, xu, xv, zu, zv);
            extension(op, t, zu, xu);   // the old fields are no longer needed
            extension(op, t, zv, xv);
            hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe + 2, t.bseg, xu, (float*)nullptr, 0);
            hipLaunchKernelGGL(k_pack, dim3(t.nblocks), dim3(256), 0, st, t.pe + 3, t.bseg, xv, (float*)nullptr, 0);
            set_dirichlet(op, d_gu, d_u[f]);
            set_dirichlet(op, d_gv, d_v[f]);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(st));   // the tables are rewritten for the next field
        }
        if (infos)
            for (int i = 0; i < 2 * nf; ++i) infos[i] = cpm_solve_info{CPM_SOLVE_CONVERGED, 0, 0, 0, 0.0};
        return;
    }
    double* w = ctx->workspace("rd_rhs", sizeof(double) * D * 2 * nf);
    std::vector<Op*> ops;
    std::vector<const double*> rhs;
    std::vector<double*> xs;
    std::vector<std::pair<double, double>> mn;
    for (int f = 0; f < nf; ++f) {
        double *ru = w + 2 * f * D, *rv = ru + D;
        hipLaunchKernelGGL(k_gray_scott_rhs, dim3(nb), dim3(256), 0, st, D, p.F, p.k, p.dt, d_u[f], d_v[f], ru, rv);
        set_dirichlet(op, d_gu, ru);
        set_dirichlet(op, d_gv, rv);
        ops.insert(ops.end(), {op, op});
        rhs.insert(rhs.end(), {ru, rv});
        xs.insert(xs.end(), {d_u[f], d_v[f]});
        mn.push_back({1.0, -p.dt * p.nu_u});
        mn.push_back({1.0, -p.dt * p.nu_v});
    }
    HIPCHK(hipGetLastError());
    std::vector<cpm_solve_info> inf;
    solve_batch(ctx, ops, rhs, xs, o, inf, &mn);
    if (infos) std::copy(inf.begin(), inf.end(), infos);
}

// A plain copy, one 16-byte element per thread over a grid that covers the
// buffer: the form that moved the most bytes per second among those measured
// on MI355X (a grid-stride loop of 8192 workgroups: 5.0 TB/s; this: 6.2 TB/s).
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

static void profile_copy(Ctx* ctx, int64_t bytes, int reps, double* avg_ms) {
    const hipStream_t st = ctx->stream;
    const int64_t n = bytes / 16;
    char* b = reinterpret_cast<char*>(ctx->workspace("copy_bufs", (size_t)bytes * 2));
    const uint4* src = reinterpret_cast<const uint4*>(b);
    uint4* dst = reinterpret_cast<uint4*>(b + bytes);
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, src, dst, n);   // warm-up
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, src, dst, n);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
}

// ---------------------------------------------------------------- serving
// Continuous batching across operators: every admitted request (an operator
// and up to NR right-hand sides) is one segment of the FP32+refine iteration;
// requests join and leave between chunks of iterations.  Workspace vectors
// live in pools of `cap` elements per vector, handed out as ranges (first
// fit); a request that does not fit waits (admission by device memory).  A
// lane's arithmetic and its reductions depend only on its own segment, so a
// request's results are the same alone and in any mix.
struct ServerBase {
    virtual ~ServerBase() = default;
    virtual bool submit(Op* op, int nrhs, const double* const* rhs, double* const* xs, int64_t tag) = 0;
    virtual void run(int iters) = 0;
    virtual int active() const = 0;
    struct Done {
        int64_t tag;
        std::vector<cpm_solve_info> infos;
    };
    std::deque<Done> done;
    int lanes = 0;
};

template <int NR>
struct Server : ServerBase {
    struct Req {
        int64_t tag;
        Op* op;
        int nr;
        int64_t off, len;
        int slot;
        std::vector<const double*> rhs;
        std::vector<double*> xs;
    };
    Ctx* ctx;
    cpm_solve_opts o;
    int64_t cap;
    Buffer p32, p64, states, tables;
    int slots;
    std::vector<int> free_slots;
    RangePool ranges;   // vector ranges of the pools
    std::vector<Req> reqs;
    bool dirty = true;
    int nblocks = 0;
    Seg<float>* d_segs = nullptr;
    int* d_bseg = nullptr;
    double* d_part = nullptr;
    PackEntry* d_pack = nullptr;
    int* d_zmap = nullptr;   // all-zero block map of single-entry pack launches
    size_t tables_bytes = 0;

    Server(Ctx* c, int64_t pool_bytes, const cpm_solve_opts& opts) : ctx(c), o(opts) {
        lanes = NR;
        const int64_t per_elem = 10 * 4 + 5 * 8;   // FP32 x r r0 p v s t y z tmp, FP64 b x r t tmp
        cap = (pool_bytes / per_elem) & ~int64_t(63);
        if (cap < 64) throw ArgError("server pool too small");
        p32 = Buffer(sizeof(float) * (size_t)cap * 10);
        p64 = Buffer(sizeof(double) * (size_t)cap * 5);
        slots = Ctx::kHostStates / kMaxNR;
        states = Buffer(sizeof(KState) * (size_t)slots * kMaxNR);
        for (int i = slots - 1; i >= 0; --i) free_slots.push_back(i);
        ranges.reset(cap);
    }
    float* v32(int k) const { return p32.template as<float>() + (size_t)k * cap; }
    double* v64(int k) const { return p64.template as<double>() + (size_t)k * cap; }
    int active() const override { return (int)reqs.size(); }

    // one pack launch per lane: b64 (lane-major) from the caller, or x64 to the caller
    void lane_pack(const Req& q, int k, double* ws64, int to_ws) {
        const hipStream_t st = ctx->stream;
        PackEntry e{};
        e.src = to_ws ? (k < q.nr ? q.rhs[(size_t)k] : nullptr) : nullptr;
        e.dst = to_ws ? nullptr : (k < q.nr ? q.xs[(size_t)k] : nullptr);
        e.perm = q.op->perm.template as<int>();
        e.D = q.op->D;
        e.off = q.off;
        e.lane = k;
        e.nr = 1;
        e.ls = q.op->D;
        e.block0 = 0;
        ensure_tables();
        HIPCHK(hipMemcpyAsync(d_pack, &e, sizeof(e), hipMemcpyHostToDevice, st));
        const int nb = (int)((q.op->D + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, st, d_pack, d_zmap, ws64, (float*)nullptr, to_ws);
        HIPCHK(hipStreamSynchronize(st));   // the entry is rewritten by the next call
    }
    // Device tables sized for the current requests: segment table, block map,
    // partial sums, one pack entry and an all-zero block map for packing.
    // Growing them reallocates (and marks the segment table for rebuilding).
    void ensure_tables() {
        int blocks = 0;
        for (auto& q : reqs) blocks += (int)((q.op->D + kBlock - 1) / kBlock);
        const int zn = (int)((cap + kBlock - 1) / kBlock);   // no request has more blocks
        auto r256 = [](size_t k) { return (k + 255) & ~size_t(255); };
        const size_t sseg = r256(sizeof(Seg<float>) * std::max<size_t>(reqs.size(), 1));
        const size_t smap = r256(sizeof(int) * 2 * std::max(blocks, 1));   // two block maps
        const size_t spart = r256(sizeof(double) * 2 * NR * std::max(blocks, 1));
        const size_t spack = r256(sizeof(PackEntry));
        const size_t szero = r256(sizeof(int) * zn);
        const size_t need = sseg + smap + spart + spack + szero;
        char* t;
        if (need > tables_bytes) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            tables = Buffer(need * 2);   // room to grow
            tables_bytes = need * 2;
            dirty = true;
            t = tables.template as<char>();
            HIPCHK(hipMemsetAsync(t, 0, szero, ctx->stream));
        }
        t = tables.template as<char>();   // fixed-size parts first: zero map, pack entry
        d_zmap = reinterpret_cast<int*>(t);
        t += szero;
        d_pack = reinterpret_cast<PackEntry*>(t);
        t += spack;
        d_segs = reinterpret_cast<Seg<float>*>(t);
        t += sseg;
        d_bseg = reinterpret_cast<int*>(t);
        t += smap;
        d_part = reinterpret_cast<double*>(t);
    }

    bool submit(Op* op, int nrhs, const double* const* rhs, double* const* xs, int64_t tag) override {
        if (nrhs < 1 || nrhs > NR) throw ArgError("a request carries 1.." + std::to_string(NR) + " right-hand sides");
        // operators built on another context of the device (a build thread's) are complete: their
        // builds end with a stream synchronisation
        if (op->ctx->device != ctx->device) throw ArgError("operator lives on another device");
        if (op->nD != 0) throw ArgError("served operators have no trailing nD rows");
        if (op->precond == CPM_PRECOND_JACOBI) throw ArgError("served operators use the diagonal or no preconditioner");
        const int64_t len = (op->D * NR + 63) & ~int64_t(63);
        if (len > cap) throw ArgError("request larger than the server pool");
        // the FP64 re-solve of this request's lanes (when their refinement stalls) runs in the
        // context's workspaces: grow them now, so that retiring the request allocates nothing
        reserve_batch_workspaces(ctx, op->D, NR, false);
        if (fp64_routing() && prefer_fp64(ctx, op)) {
            // the operator (or its family) stalled in FP32 before: the request is solved in FP64
            // from zero now, as the re-solve of a stalled request is, without FP32 iterations
            Done d{tag, std::vector<cpm_solve_info>((size_t)nrhs)};
            for (auto& inf : d.infos) inf.status = ST_FALLBACK;
            NoBuilds nb(ctx);
            fp64_fallback(ctx, std::vector<Op*>((size_t)nrhs, op), std::vector<const double*>(rhs, rhs + nrhs),
                          std::vector<double*>(xs, xs + nrhs), o, d.infos, nullptr);
            done.push_back(std::move(d));
            return true;
        }
        int64_t off;
        if (free_slots.empty() || !ranges.take(len, off)) return false;   // wait for memory
        Req q{tag, op, nrhs, off, len, free_slots.back(), {}, {}};
        free_slots.pop_back();
        q.rhs.assign(rhs, rhs + nrhs);
        q.xs.assign(xs, xs + nrhs);
        const hipStream_t st = ctx->stream;
        // x64 = x32 = 0 over the range, b64 lanes from the caller (idle lanes 0)
        HIPCHK(hipMemsetAsync(v64(1) + off, 0, sizeof(double) * len, st));
        HIPCHK(hipMemsetAsync(v32(0) + off, 0, sizeof(float) * len, st));
        reqs.push_back(q);
        dirty = true;
        for (int k = 0; k < NR; ++k) lane_pack(reqs.back(), k, v64(0), 1);
        // fresh lanes enter the refinement step (||b - A 0||^2 = ||b||^2), idle lanes stay stopped
        const double delta = o.update_delta > 0 ? std::min(o.update_delta, 0.999) : kDefaultUpdateDelta;
        std::vector<KState> hs(kMaxNR);
        for (int k = 0; k < kMaxNR; ++k) {
            KState z{};
            z.tol2 = o.tol * o.tol;
            z.max_it = o.max_iters;
            z.delta2 = delta * delta;
            if (k < nrhs) {
                z.status = ST_UPDATE;
                z.bsq = -1.0;
                z.updates = -1;
                z.scale = 1.0;
            } else {
                z.status = ST_CONVERGED;
            }
            hs[(size_t)k] = z;
        }
        HIPCHK(hipMemcpyAsync(states.template as<KState>() + (size_t)q.slot * kMaxNR, hs.data(), sizeof(KState) * kMaxNR,
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return true;
    }

    void rebuild() {
        ensure_tables();
        {
            std::vector<Op*> ops;
            for (const Req& q : reqs) ops.push_back(q.op);
            ensure_group_tiles_for(ops, 4, NR);
        }
        std::vector<Seg<float>> h(reqs.size());
        std::vector<int> nbs;
        int b0 = 0;
        for (size_t g = 0; g < reqs.size(); ++g) {
            const Req& q = reqs[g];
            const int nb = (int)((q.op->D + kBlock - 1) / kBlock);
            fill_seg(h[g], q.op, q.off, nb, q.nr, states.template as<KState>() + (size_t)q.slot * kMaxNR,
                     d_part + (size_t)2 * NR * b0);
            nbs.push_back(nb);
            b0 += nb;
        }
        const std::vector<int> map = block_map(nbs);
        nblocks = b0;
        const hipStream_t st = ctx->stream;
        HIPCHK(hipMemcpyAsync(d_segs, h.data(), sizeof(Seg<float>) * h.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_bseg, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        dirty = false;
    }

    void run(int iters) override {
        if (reqs.empty()) return;
        if (dirty) rebuild();
        const hipStream_t st = ctx->stream;
        const Vecs<float> V32{nullptr, v32(0), v32(1), v32(2), v32(3), v32(4), v32(5), v32(6), v32(7), v32(8), v32(9)};
        const MixVec mv{v64(1), v64(2), v64(0), v64(4)};
        int grp = INT_MAX;
        for (const Req& q : reqs) grp = std::min(grp, q.op->ngt);
        const LaunchCfg<float> c{d_segs, d_bseg, nblocks, (int)reqs.size(), st, 0, grp};
        Vecs<float> w = V32;
        for (int i = 0; i < iters; ++i) Launch<float, NR>::iteration(c, w, true, mv);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(ctx->host_states, states.template as<KState>(), sizeof(KState) * (size_t)slots * kMaxNR,
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        // a private copy: FP64 re-solves below poll through the context's pinned state buffer
        const std::vector<KState> snap(ctx->host_states, ctx->host_states + (size_t)slots * kMaxNR);
        const KState* hs = snap.data();
        std::vector<Req> keep, leaving;
        for (Req& q : reqs) {
            const KState* s = hs + (size_t)q.slot * kMaxNR;
            bool running = false;
            for (int k = 0; k < q.nr; ++k)
                running |= s[k].status == ST_RUN || (s[k].status >= ST_UPDATE && s[k].status <= ST_READY);
            (running ? keep : leaving).push_back(q);
        }
        reqs.swap(keep);   // the leaving requests are out of the iteration before anything can throw
        if (leaving.empty()) return;
        dirty = true;
        auto finish = [&](const Req& q) {
            const KState* s = hs + (size_t)q.slot * kMaxNR;
            Done d{q.tag, std::vector<cpm_solve_info>((size_t)q.nr)};
            for (int k = 0; k < q.nr; ++k) {
                cpm_solve_info& inf = d.infos[(size_t)k];
                inf.status = s[k].status;
                inf.iters = s[k].iters;
                inf.restarts = s[k].restarts;
                inf.updates = s[k].updates;
                inf.rel_residual = s[k].bsq > 0 ? std::sqrt(s[k].rr_true / s[k].bsq) : 0.0;
            }
            try {
                for (int k = 0; k < q.nr; ++k) lane_pack(q, k, v64(1), 0);
                NoBuilds nb(ctx);   // the re-solve uses the workspaces reserved at submission
                const std::vector<Op*> ops((size_t)q.nr, q.op);
                fp64_fallback(ctx, ops, q.rhs, q.xs, o, d.infos, nullptr);
                record_fp32(ctx, ops, d.infos);
            } catch (...) {
                for (auto& inf : d.infos) inf.status = CPM_SOLVE_NOT_WRITTEN;
                done.push_back(std::move(d));
                throw;
            }
            done.push_back(std::move(d));
        };
        auto release = [&](const Req& q) {
            if (!ranges.give(q.off, q.len)) throw std::logic_error("server range given back twice");
            free_slots.push_back(q.slot);
        };
        retire_requests(leaving, finish, release);
    }
};

// The same request loop in FP64 (the reference's arithmetic, no refinement):
// each admitted request is one segment of the FP64 iteration.  A request's
// lanes start with the FP64 BiCGStab initialisation of its own segment
// (r = r0 = b, x = p = v = 0, rho = ||b||^2), launched for that segment alone
// when it is admitted; rho-restarts (BiCGStab.h:62-71) are taken between
// chunks of iterations.  Workspace: 11 FP64 vectors, 88 B per row and lane.
template <int NR>
struct Server64 : ServerBase {
    struct Req {
        int64_t tag;
        Op* op;
        int nr;
        int64_t off, len;
        int slot;
        std::vector<double*> xs;
    };
    Ctx* ctx;
    cpm_solve_opts o;
    int64_t cap;
    Buffer p64, states, tables, init_tables;
    int slots;
    std::vector<int> free_slots;
    RangePool ranges;
    std::vector<Req> reqs;
    bool dirty = true;
    int nblocks = 0;
    Seg<double>* d_segs = nullptr;
    int* d_bseg = nullptr;
    double* d_part = nullptr;
    PackEntry* d_pack = nullptr;
    int* d_zmap = nullptr;
    size_t tables_bytes = 0;

    Server64(Ctx* c, int64_t pool_bytes, const cpm_solve_opts& opts) : ctx(c), o(opts) {
        lanes = NR;
        cap = (pool_bytes / (11 * 8)) & ~int64_t(63);
        if (cap < 64) throw ArgError("server pool too small");
        p64 = Buffer(sizeof(double) * (size_t)cap * 11);
        slots = Ctx::kHostStates / kMaxNR;
        states = Buffer(sizeof(KState) * (size_t)slots * kMaxNR);
        for (int i = slots - 1; i >= 0; --i) free_slots.push_back(i);
        ranges.reset(cap);
        const int zn = (int)((cap + kBlock - 1) / kBlock);
        // single-segment tables of the admission launches: segment, partials, pack entry, zero block map
        init_tables = Buffer(4096 + sizeof(double) * 2 * NR * (size_t)zn + 4096 + sizeof(int) * (size_t)zn);
        HIPCHK(hipMemset(init_tables.p, 0, init_tables.bytes));
    }
    double* v64(int k) const { return p64.template as<double>() + (size_t)k * cap; }   // b x r r0 p v s t y z tmp
    int active() const override { return (int)reqs.size(); }
    Seg<double>* i_seg() const { return init_tables.template as<Seg<double>>(); }
    double* i_part() const { return reinterpret_cast<double*>(init_tables.template as<char>() + 4096); }
    PackEntry* i_pack() const {
        const int zn = (int)((cap + kBlock - 1) / kBlock);
        return reinterpret_cast<PackEntry*>(init_tables.template as<char>() + 4096 + sizeof(double) * 2 * NR * (size_t)zn);
    }
    int* i_zmap() const { return reinterpret_cast<int*>(reinterpret_cast<char*>(i_pack()) + 4096); }

    // one pack launch per lane of the interleaved FP64 vectors: b from the caller, or x to the caller
    void lane_pack(const Req& q, int k, const double* src, double* ws, int to_ws) {
        const hipStream_t st = ctx->stream;
        PackEntry e{};
        e.src = to_ws ? src : nullptr;
        e.dst = to_ws ? nullptr : (k < q.nr ? q.xs[(size_t)k] : nullptr);
        e.perm = q.op->perm.template as<int>();
        e.D = q.op->D;
        e.off = q.off;
        e.lane = k;
        e.nr = NR;
        e.ls = 1;
        e.block0 = 0;
        HIPCHK(hipMemcpyAsync(i_pack(), &e, sizeof(e), hipMemcpyHostToDevice, st));
        const int nb = (int)((q.op->D + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, st, i_pack(), i_zmap(), ws, (float*)nullptr, to_ws);
        HIPCHK(hipStreamSynchronize(st));   // the entry is rewritten by the next call
    }
    void ensure_tables() {
        int blocks = 0;
        for (auto& q : reqs) blocks += (int)((q.op->D + kBlock - 1) / kBlock);
        auto r256 = [](size_t k) { return (k + 255) & ~size_t(255); };
        const size_t sseg = r256(sizeof(Seg<double>) * std::max<size_t>(reqs.size(), 1));
        const size_t smap = r256(sizeof(int) * 2 * std::max(blocks, 1));   // two block maps
        const size_t spart = r256(sizeof(double) * 2 * NR * std::max(blocks, 1));
        const size_t need = sseg + smap + spart;
        if (need > tables_bytes) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            tables = Buffer(need * 2);
            tables_bytes = need * 2;
        }
        char* t = tables.template as<char>();
        d_segs = reinterpret_cast<Seg<double>*>(t);
        d_bseg = reinterpret_cast<int*>(t + sseg);
        d_part = reinterpret_cast<double*>(t + sseg + smap);
    }

    bool submit(Op* op, int nrhs, const double* const* rhs, double* const* xs, int64_t tag) override {
        if (nrhs < 1 || nrhs > NR) throw ArgError("a request carries 1.." + std::to_string(NR) + " right-hand sides");
        if (op->ctx->device != ctx->device) throw ArgError("operator lives on another device");
        if (op->nD != 0) throw ArgError("served operators have no trailing nD rows");
        if (op->precond == CPM_PRECOND_JACOBI) throw ArgError("served operators use the diagonal or no preconditioner");
        const int64_t len = (op->D * NR + 63) & ~int64_t(63);
        if (len > cap) throw ArgError("request larger than the server pool");
        ensure_group_tiles_for({op}, 8, NR);   // before its view is taken
        int64_t off;
        if (free_slots.empty() || !ranges.take(len, off)) return false;   // wait for memory
        Req q{tag, op, nrhs, off, len, free_slots.back(), {}};
        free_slots.pop_back();
        q.xs.assign(xs, xs + nrhs);
        const hipStream_t st = ctx->stream;
        for (int k = 0; k < NR; ++k) lane_pack(q, k, k < nrhs ? rhs[k] : nullptr, v64(0), 1);   // b, idle lanes 0
        // states of the request's lanes: fresh (FIN_INIT completes them) or stopped (idle lanes)
        std::vector<KState> hs(kMaxNR);
        for (int k = 0; k < kMaxNR; ++k) {
            KState z{};
            z.tol2 = o.tol * o.tol;
            z.max_it = o.max_iters;
            z.status = k < nrhs ? ST_RUN : ST_CONVERGED;
            hs[(size_t)k] = z;
        }
        KState* sts = states.template as<KState>() + (size_t)q.slot * kMaxNR;
        HIPCHK(hipMemcpyAsync(sts, hs.data(), sizeof(KState) * kMaxNR, hipMemcpyHostToDevice, st));
        // BiCGStab initialisation of this request's segment alone
        const int nb = (int)((op->D + kBlock - 1) / kBlock);
        Seg<double> h{};
        fill_seg(h, op, off, nb, nrhs, sts, i_part());
        HIPCHK(hipMemcpyAsync(i_seg(), &h, sizeof(h), hipMemcpyHostToDevice, st));
        const LaunchCfg<double> c1{i_seg(), nullptr, nb, 1, st, 0, op->ngt};
        Launch<double, NR>::init(c1, v64(0), v64(1), v64(2), v64(3), v64(4), v64(5));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        reqs.push_back(q);
        dirty = true;
        return true;
    }

    void rebuild() {
        ensure_tables();
        std::vector<Seg<double>> h(reqs.size());
        std::vector<int> nbs;
        int b0 = 0;
        for (size_t g = 0; g < reqs.size(); ++g) {
            const Req& q = reqs[g];
            const int nb = (int)((q.op->D + kBlock - 1) / kBlock);
            fill_seg(h[g], q.op, q.off, nb, q.nr, states.template as<KState>() + (size_t)q.slot * kMaxNR,
                     d_part + (size_t)2 * NR * b0);
            nbs.push_back(nb);
            b0 += nb;
        }
        const std::vector<int> map = block_map(nbs);
        nblocks = b0;
        const hipStream_t st = ctx->stream;
        HIPCHK(hipMemcpyAsync(d_segs, h.data(), sizeof(Seg<double>) * h.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_bseg, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        dirty = false;
    }

    void run(int iters) override {
        if (reqs.empty()) return;
        if (dirty) rebuild();
        const hipStream_t st = ctx->stream;
        Vecs<double> V{v64(0), v64(1), v64(2), v64(3), v64(4), v64(5), v64(6), v64(7), v64(8), v64(9), v64(10)};
        int grp = INT_MAX;
        for (const Req& q : reqs) grp = std::min(grp, q.op->ngt);
        const LaunchCfg<double> c{d_segs, d_bseg, nblocks, (int)reqs.size(), st, 0, grp};
        for (int i = 0; i < iters; ++i) Launch<double, NR>::iteration(c, V, false, MixVec{});
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(ctx->host_states, states.template as<KState>(), sizeof(KState) * (size_t)slots * kMaxNR,
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const std::vector<KState> snap(ctx->host_states, ctx->host_states + (size_t)slots * kMaxNR);
        const KState* hs = snap.data();
        bool restart = false;
        std::vector<Req> keep, leaving;
        for (Req& q : reqs) {
            const KState* s = hs + (size_t)q.slot * kMaxNR;
            bool running = false;
            for (int k = 0; k < q.nr; ++k) {
                running |= s[k].status == ST_RUN || s[k].status == ST_RESTART;
                restart |= s[k].status == ST_RESTART;
            }
            (running ? keep : leaving).push_back(q);
        }
        if (restart) {   // rare: the restart kernels act on the lanes in ST_RESTART only
            Launch<double, NR>::restart(c, V);
            HIPCHK(hipGetLastError());
        }
        reqs.swap(keep);
        if (leaving.empty()) return;
        dirty = true;
        auto finish = [&](const Req& q) {
            const KState* s = hs + (size_t)q.slot * kMaxNR;
            Done d{q.tag, std::vector<cpm_solve_info>((size_t)q.nr)};
            for (int k = 0; k < q.nr; ++k) {
                cpm_solve_info& inf = d.infos[(size_t)k];
                inf.status = s[k].status;
                inf.iters = s[k].iters;
                inf.restarts = s[k].restarts;
                inf.rel_residual = s[k].bsq > 0 ? std::sqrt(s[k].rr / s[k].bsq) : 0.0;
            }
            try {
                for (int k = 0; k < q.nr; ++k) lane_pack(q, k, nullptr, v64(1), 0);
            } catch (...) {
                for (auto& inf : d.infos) inf.status = CPM_SOLVE_NOT_WRITTEN;
                done.push_back(std::move(d));
                throw;
            }
            done.push_back(std::move(d));
        };
        auto release = [&](const Req& q) {
            if (!ranges.give(q.off, q.len)) throw std::logic_error("server range given back twice");
            free_slots.push_back(q.slot);
        };
        retire_requests(leaving, finish, release);
    }
};

}  // namespace cpmgpu

// ====================================================================== C ABI
using namespace cpmgpu;

// The opaque C handles are the engine objects themselves.
static inline Ctx* C(cpm_ctx* c) { return reinterpret_cast<Ctx*>(c); }
static inline const Ctx* C(const cpm_ctx* c) { return reinterpret_cast<const Ctx*>(c); }
static inline Op* O(cpm_op* o) { return reinterpret_cast<Op*>(o); }
static inline const Op* O(const cpm_op* o) { return reinterpret_cast<const Op*>(o); }

namespace cpmgpu {
void set_last_error(const std::string& s) { g_err = s; }
// Map the exception in flight to a status code and record its message.
int status_of_current_exception() {
    try {
        throw;
    } catch (const ArgError& e) {
        g_err = e.what();
        return CPM_ERR_ARG;
    } catch (const NoMem& e) {
        g_err = e.what();
        return CPM_ERR_NOMEM;
    } catch (const GeomError& e) {
        g_err = e.what();
        return CPM_ERR_GEOMETRY;
    } catch (const HipError& e) {
        g_err = e.what();
        return CPM_ERR_HIP;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CPM_ERR_STATE;
    } catch (...) {
        g_err = "unknown error";
        return CPM_ERR_STATE;
    }
}

cpm_solve_opts resolve_opts(const cpm_solve_opts* o) {
    cpm_solve_opts r;
    cpm_solve_opts_default(&r);
    if (o) r = *o;
    if (!(r.tol >= 0)) throw ArgError("tol must be >= 0");
    return r;
}
}  // namespace cpmgpu

template <typename F>
static int guard(F&& f) {
    try {
        f();
        return CPM_OK;
    } catch (...) {
        return status_of_current_exception();
    }
}

extern "C" {

void cpm_solve_opts_default(cpm_solve_opts* o) {
    if (!o) return;
    o->tol = 1e-10;
    o->max_iters = 100000;
    o->precision = CPM_FP64;
    o->update_delta = kDefaultUpdateDelta;
    o->reserved = 0;
}

const char* cpm_last_error(void) { return g_err.c_str(); }

int cpm_ctx_create(int device, cpm_ctx** out) {
    return guard([&] {
        if (!out) throw ArgError("out is null");
        int n = 0;
        HIPCHK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw ArgError("no such device");
        *out = reinterpret_cast<cpm_ctx*>(new Ctx(device));
    });
}

int cpm_ctx_destroy(cpm_ctx* ctx) {
    return guard([&] { delete C(ctx); });
}

int cpm_ctx_synchronize(cpm_ctx* ctx) {
    return guard([&] {
        if (!ctx) throw ArgError("null context");
        HIPCHK(hipStreamSynchronize(C(ctx)->stream));
    });
}

int64_t cpm_ctx_device_bytes(const cpm_ctx* ctx) {
    if (!ctx) return -1;
    int64_t b = 0;
    for (auto& kv : C(ctx)->ws) b += kv.second ? (int64_t)kv.second->bytes : 0;
    return b + C(ctx)->op_bytes;
}

int cpm_op_create_csr(cpm_ctx* ctx, int64_t D, int nD, const int32_t* Lp, const int32_t* Lc, const double* Lv,
                      const int32_t* Ep, const int32_t* Ec, const double* Ev, const uint8_t* ident, double m,
                      double n, int precond, int precond_iters, cpm_op** out) {
    return guard([&] {
        if (!ctx || !out || !Lp || !Lc || !Lv || !Ep || !Ec || !Ev || !ident) throw ArgError("null argument");
        if (precond < 0 || precond > 2) throw ArgError("unknown preconditioner");
        HIPCHK(hipSetDevice(C(ctx)->device));
        Op* op = op_from_csr(C(ctx), D, nD, Lp, Lc, Lv, Ep, Ec, Ev, ident, m, n, precond, precond_iters);
        C(ctx)->op_bytes += (int64_t)op->device_bytes();
        *out = reinterpret_cast<cpm_op*>(op);
    });
}

int cpm_op_destroy(cpm_op* op) {
    return guard([&] {
        if (!op) return;
        O(op)->ctx->op_bytes -= (int64_t)O(op)->device_bytes();
        delete O(op);
    });
}

int64_t cpm_op_rows(const cpm_op* op) { return op ? O(op)->D : -1; }

int cpm_op_format(const cpm_op* op) { return op ? O(op)->fmt : -1; }

int cpm_op_gather_kind(const cpm_op* op, int precision, int nrhs) {
    if (!op || (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8)) return -1;
    Op* o = const_cast<Op*>(O(op));   // the lazy tile build does not change what the operator computes
    const bool f32 = precision == CPM_FP32_REFINE;
    try {
        ensure_group_tiles_for({o}, f32 ? 4 : 8, nrhs);
    } catch (...) {
        return -1;
    }
    const bool fits = f32 ? (nrhs == 1 ? GrpTile<float, 1>::kRows : nrhs == 2 ? GrpTile<float, 2>::kRows
                             : nrhs == 4 ? GrpTile<float, 4>::kRows : GrpTile<float, 8>::kRows) > 0
                          : (nrhs == 1 ? GrpTile<double, 1>::kRows : nrhs == 2 ? GrpTile<double, 2>::kRows
                             : nrhs == 4 ? GrpTile<double, 4>::kRows : GrpTile<double, 8>::kRows) > 0;
    if (o->fmt == OP_TP && o->ngt > 0 && fits && grp_gather(f32 ? 4 : 8, nrhs, o->ngt)) return 2;
    const int bytes = (f32 ? 4 : 8) * nrhs;
    if (o->fmt == OP_TP && o->tln.p && kLdsRows * bytes <= kLdsBytes) return 1;
    return 0;
}

int cpm_op_tile_stats(const cpm_op* op, int64_t* out) {
    return guard([&] {
        if (!op || !out) throw ArgError("null operator or output");
        Op* o = const_cast<Op*>(O(op));   // the lazy tile build does not change what the operator computes
        op_ensure_group_tiles(*o);
        int64_t row_tiles = 0;
        if (o->tln.p) {
            std::vector<int> tln(o->tln.bytes / 4);
            HIPCHK(hipMemcpy(tln.data(), o->tln.p, o->tln.bytes, hipMemcpyDeviceToHost));
            for (int v : tln) row_tiles += v >= 0;
        }
        out[0] = o->ngt;
        out[1] = o->grp_items;
        out[2] = o->grp_items0;
        out[3] = o->grp_list;
        out[4] = row_tiles;
        out[5] = o->tile_list;
    });
}

int cpm_op_set_mn(cpm_op* op, double m, double n) {
    return guard([&] {
        if (!op) throw ArgError("null operator");
        O(op)->m = m;
        O(op)->n = n;
    });
}

int cpm_op_set_precond(cpm_op* op, int precond, int precond_iters) {
    return guard([&] {
        if (!op) throw ArgError("null operator");
        if (precond < 0 || precond > 2) throw ArgError("unknown preconditioner");
        O(op)->precond = precond;
        O(op)->precond_iters = precond_iters;
    });
}

int cpm_multiply_device(cpm_op* op, const double* dx, double* dy, double m, double n) {
    return guard([&] {
        if (!op || !dx || !dy) throw ArgError("null argument");
        multiply(O(op), dx, dy, m, n);
    });
}

int cpm_multiply(cpm_op* op, const double* x, double* y, double m, double n) {
    return guard([&] {
        if (!op || !x || !y) throw ArgError("null argument");
        Op* o = O(op);
        Ctx* ctx = o->ctx;
        double* d = ctx->workspace("mul_io", sizeof(double) * o->D * 2);
        HIPCHK(hipMemcpyAsync(d, x, sizeof(double) * o->D, hipMemcpyHostToDevice, ctx->stream));
        multiply(o, d, d + o->D, m, n);
        HIPCHK(hipMemcpyAsync(y, d + o->D, sizeof(double) * o->D, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int cpm_solve_stream_device(cpm_op* op, int n_rhs, const double* const* d_rhs, double* const* d_x, int lanes,
                            const cpm_solve_opts* opts, cpm_solve_info* infos) {
    return guard([&] {
        if (!op || n_rhs < 0 || (n_rhs > 0 && (!d_rhs || !d_x || !infos))) throw ArgError("null argument");
        for (int i = 0; i < n_rhs; ++i)
            if (!d_rhs[i] || !d_x[i]) throw ArgError("null vector");
        HIPCHK(hipSetDevice(O(op)->ctx->device));
        std::vector<const double*> r(d_rhs, d_rhs + n_rhs);
        std::vector<double*> x(d_x, d_x + n_rhs);
        std::vector<cpm_solve_info> inf;
        solve_stream(O(op), r, x, lanes, resolve_opts(opts), inf);
        std::copy(inf.begin(), inf.end(), infos);
    });
}

int cpm_solve_batch_device(cpm_ctx* ctx, int n_inst, cpm_op* const* ops, const double* const* d_rhs,
                           double* const* d_x, const cpm_solve_opts* opts, cpm_solve_info* infos) {
    return guard([&] {
        if (!ctx || n_inst < 0 || (n_inst > 0 && (!ops || !d_rhs || !d_x))) throw ArgError("null argument");
        cpm_solve_opts o = resolve_opts(opts);
        std::vector<Op*> vo(n_inst);
        for (int i = 0; i < n_inst; ++i) vo[i] = O(ops[i]);
        std::vector<const double*> vr(d_rhs, d_rhs + n_inst);
        std::vector<double*> vx(d_x, d_x + n_inst);
        std::vector<cpm_solve_info> inf;
        HIPCHK(hipSetDevice(C(ctx)->device));
        solve_batch(C(ctx), vo, vr, vx, o, inf);
        if (infos) std::copy(inf.begin(), inf.end(), infos);
    });
}

int cpm_solve_batch(cpm_ctx* ctx, int n_inst, cpm_op* const* ops, const double* const* rhs, double* const* x,
                    const cpm_solve_opts* opts, cpm_solve_info* infos) {
    return guard([&] {
        if (!ctx || n_inst < 0 || (n_inst > 0 && (!ops || !rhs || !x))) throw ArgError("null argument");
        cpm_solve_opts o = resolve_opts(opts);
        Ctx* cx = C(ctx);
        HIPCHK(hipSetDevice(cx->device));
        std::vector<Op*> vo(n_inst);
        int64_t tot = 0;
        for (int i = 0; i < n_inst; ++i) {
            vo[i] = O(ops[i]);
            if (!vo[i]) throw ArgError("null operator");
            tot += vo[i]->D;
        }
        double* d = cx->workspace("batch_io", sizeof(double) * tot * 2 + 256);
        std::vector<const double*> vr(n_inst);
        std::vector<double*> vx(n_inst);
        int64_t a = 0;
        for (int i = 0; i < n_inst; ++i) {
            vr[i] = d + a;
            vx[i] = d + tot + a;
            HIPCHK(hipMemcpyAsync(d + a, rhs[i], sizeof(double) * vo[i]->D, hipMemcpyHostToDevice, cx->stream));
            a += vo[i]->D;
        }
        std::vector<cpm_solve_info> inf;
        solve_batch(cx, vo, vr, vx, o, inf);
        for (int i = 0; i < n_inst; ++i)
            HIPCHK(hipMemcpyAsync(x[i], vx[i], sizeof(double) * vo[i]->D, hipMemcpyDeviceToHost, cx->stream));
        HIPCHK(hipStreamSynchronize(cx->stream));
        if (infos) std::copy(inf.begin(), inf.end(), infos);
    });
}

int cpm_solve_device(cpm_op* op, const double* d_rhs, double* d_x, const cpm_solve_opts* opts, cpm_solve_info* info) {
    if (!op) {
        g_err = "null operator";
        return CPM_ERR_ARG;
    }
    cpm_op* ops[1] = {op};
    const double* r[1] = {d_rhs};
    double* x[1] = {d_x};
    return cpm_solve_batch_device(reinterpret_cast<cpm_ctx*>(O(op)->ctx), 1, ops, r, x, opts, info);
}

int cpm_solve_device_noalloc(cpm_op* op, const double* d_rhs, double* d_x, const cpm_solve_opts* opts,
                             cpm_solve_info* info) {
    return guard([&] {
        if (!op || !d_rhs || !d_x) throw ArgError("null argument");
        Op* o = O(op);
        Ctx* cx = o->ctx;
        const cpm_solve_opts so = resolve_opts(opts);
        if (so.precision != CPM_FP64 && so.precision != CPM_FP32_REFINE) throw ArgError("unknown precision");
        const size_t need = std::max(batch_bytes1(o->D, 1, false).arena(), batch_bytes1(o->D, 1, true).arena());
        if (need > (size_t)CPM_NOALLOC_SCRATCH_BYTES)
            throw ArgError("operator of " + std::to_string(o->D) + " rows needs " + std::to_string(need) +
                           " bytes of workspace; the no-alloc scratch holds " +
                           std::to_string((long long)CPM_NOALLOC_SCRATCH_BYTES));
        HIPCHK(hipSetDevice(cx->device));
        Arena a;
        a.base = static_cast<char*>(scratch_ptr(cx->device));
        a.cap = (size_t)CPM_NOALLOC_SCRATCH_BYTES;
        struct Use {   // the context carves from the scratch only inside this call
            Ctx* c;
            Use(Ctx* c_, Arena* ar) : c(c_) { c->arena = ar; }
            ~Use() { c->arena = nullptr; }
        } use(cx, &a);
        grp_refresh_override();
        std::vector<cpm_solve_info> inf;
        solve_batch(cx, std::vector<Op*>{o}, std::vector<const double*>{d_rhs}, std::vector<double*>{d_x}, so, inf);
        if (info) *info = inf[0];
    });
}

int cpm_solve(cpm_op* op, const double* rhs, double* x, const cpm_solve_opts* opts, cpm_solve_info* info) {
    if (!op) {
        g_err = "null operator";
        return CPM_ERR_ARG;
    }
    cpm_op* ops[1] = {op};
    const double* r[1] = {rhs};
    double* xx[1] = {x};
    return cpm_solve_batch(reinterpret_cast<cpm_ctx*>(O(op)->ctx), 1, ops, r, xx, opts, info);
}

int cpm_profile_egather_groups(cpm_op* op, int precision, int nrhs, int groups, int reps, double* avg_ms,
                               double* bytes) {
    return guard([&] {
        if (!op || !avg_ms || !bytes || reps <= 0) throw ArgError("bad argument");
        if (groups < 1 || groups > 64) throw ArgError("groups must be in 1..64");
        Op* o = O(op);
        HIPCHK(hipSetDevice(o->ctx->device));
        const bool f32 = precision == CPM_FP32_REFINE;
        const int G = groups;
        switch (nrhs) {
        case 1: f32 ? profile_egather_t<float, 1>(o, G, reps, avg_ms, bytes) : profile_egather_t<double, 1>(o, G, reps, avg_ms, bytes); break;
        case 2: f32 ? profile_egather_t<float, 2>(o, G, reps, avg_ms, bytes) : profile_egather_t<double, 2>(o, G, reps, avg_ms, bytes); break;
        case 4: f32 ? profile_egather_t<float, 4>(o, G, reps, avg_ms, bytes) : profile_egather_t<double, 4>(o, G, reps, avg_ms, bytes); break;
        case 8: f32 ? profile_egather_t<float, 8>(o, G, reps, avg_ms, bytes) : profile_egather_t<double, 8>(o, G, reps, avg_ms, bytes); break;
        default: throw ArgError("nrhs must be 1, 2, 4 or 8");
        }
    });
}

int cpm_profile_egather(cpm_op* op, int precision, int nrhs, int reps, double* avg_ms, double* bytes) {
    return cpm_profile_egather_groups(op, precision, nrhs, 1, reps, avg_ms, bytes);
}

int cpm_heat_step_device(cpm_op* op, int method, double dt, const double* d_g, double* d_u,
                         const cpm_solve_opts* opts, cpm_solve_info* info) {
    return guard([&] {
        if (!op || !d_g || !d_u) throw ArgError("null argument");
        if (method < 0 || method > 2) throw ArgError("method must be 0 (explicit), 1 (backward Euler) or 2 (Crank-Nicolson)");
        if (!(dt > 0)) throw ArgError("dt must be positive");
        if (O(op)->nD != 0) throw ArgError("time stepping needs an operator without trailing nD rows");
        HIPCHK(hipSetDevice(O(op)->ctx->device));
        heat_step(O(op), method, dt, d_g, d_u, resolve_opts(opts), info);
    });
}

void cpm_rd_params_default(double dx, cpm_rd_params* p) {
    if (!p) return;
    p->F = 0.054;                          // ReactionDiffusionWithIBC.cpp:83-84
    p->k = 0.063;
    p->nu_u = 1.0 / ((3.0 / dx) * (3.0 / dx));   // :81-82
    p->nu_v = p->nu_u / 3.0;
    p->dt = 0.1 * (1.0 / std::max(p->nu_u, p->nu_v)) * dx * dx;   // examples/ReactionDiffusionFish/ReactionDiffusionFish.cpp:61-63
    p->method = 0;
    p->reserved = 0;
}

int cpm_rd_step_device(cpm_op* op, const cpm_rd_params* p, int n_fields, double* const* d_u, double* const* d_v,
                       const double* d_g_u, const double* d_g_v, const cpm_solve_opts* opts, cpm_solve_info* infos) {
    return guard([&] {
        if (!op || !p || !d_u || !d_v || !d_g_u || !d_g_v) throw ArgError("null argument");
        if (n_fields <= 0 || n_fields > 4096) throw ArgError("n_fields must be in 1..4096");
        for (int f = 0; f < n_fields; ++f)
            if (!d_u[f] || !d_v[f]) throw ArgError("null field pointer");
        if (p->method != 0 && p->method != 1) throw ArgError("method must be 0 (explicit) or 1 (implicit diffusion)");
        if (!(p->dt > 0) || !(p->nu_u >= 0) || !(p->nu_v >= 0)) throw ArgError("dt must be positive, nu non-negative");
        if (O(op)->nD != 0) throw ArgError("time stepping needs an operator without trailing nD rows");
        HIPCHK(hipSetDevice(O(op)->ctx->device));
        rd_step(O(op), *p, n_fields, d_u, d_v, d_g_u, d_g_v, resolve_opts(opts), infos);
    });
}

int cpm_profile_copy(cpm_ctx* ctx, int64_t bytes, int reps, double* avg_ms) {
    return guard([&] {
        if (!ctx || !avg_ms || reps <= 0 || bytes <= 0 || bytes % 16) throw ArgError("bad argument");
        HIPCHK(hipSetDevice(C(ctx)->device));
        profile_copy(C(ctx), bytes, reps, avg_ms);
    });
}

int cpm_device_alloc(cpm_ctx* ctx, int64_t bytes, void** d_ptr) {
    return guard([&] {
        if (!ctx || !d_ptr || bytes < 0) throw ArgError("bad argument");
        HIPCHK(hipSetDevice(C(ctx)->device));
        void* p = nullptr;
        if (hipMalloc(&p, (size_t)std::max<int64_t>(bytes, 1)) != hipSuccess) {
            (void)hipGetLastError();
            throw NoMem("hipMalloc failed");
        }
        C(ctx)->user_bytes[p] = bytes;
        *d_ptr = p;
    });
}

int cpm_device_free(cpm_ctx* ctx, void* p) {
    return guard([&] {
        if (!ctx) throw ArgError("null context");
        C(ctx)->user_bytes.erase(p);
        HIPCHK(hipFree(p));
    });
}

int cpm_memcpy_h2d(cpm_ctx* ctx, void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        if (!ctx) throw ArgError("null context");
        HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, C(ctx)->stream));
        HIPCHK(hipStreamSynchronize(C(ctx)->stream));
    });
}

int cpm_memcpy_d2h(cpm_ctx* ctx, void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        if (!ctx) throw ArgError("null context");
        HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, C(ctx)->stream));
        HIPCHK(hipStreamSynchronize(C(ctx)->stream));
    });
}

int cpm_server_create(cpm_ctx* ctx, int64_t pool_bytes, int lanes, const cpm_solve_opts* opts, cpm_server** out) {
    return guard([&] {
        if (!ctx || !out) throw ArgError("null argument");
        HIPCHK(hipSetDevice(C(ctx)->device));
        cpm_solve_opts o = resolve_opts(opts);
        if (!opts) o.precision = CPM_FP32_REFINE;   // the server's default precision
        ServerBase* s = nullptr;
        if (o.precision == CPM_FP64) {
            if (lanes == 1) s = new Server64<1>(C(ctx), pool_bytes, o);
            else if (lanes == 2) s = new Server64<2>(C(ctx), pool_bytes, o);
            else if (lanes == 4) s = new Server64<4>(C(ctx), pool_bytes, o);
            else if (lanes == 8) s = new Server64<8>(C(ctx), pool_bytes, o);
            else throw ArgError("lanes must be 1, 2, 4 or 8");
        } else {
            o.precision = CPM_FP32_REFINE;
            if (lanes == 1) s = new Server<1>(C(ctx), pool_bytes, o);
            else if (lanes == 2) s = new Server<2>(C(ctx), pool_bytes, o);
            else if (lanes == 4) s = new Server<4>(C(ctx), pool_bytes, o);
            else if (lanes == 8) s = new Server<8>(C(ctx), pool_bytes, o);
            else throw ArgError("lanes must be 1, 2, 4 or 8");
        }
        *out = reinterpret_cast<cpm_server*>(s);
    });
}

int cpm_server_destroy(cpm_server* s) {
    return guard([&] { delete reinterpret_cast<ServerBase*>(s); });
}

int cpm_server_submit(cpm_server* sv, cpm_op* op, int n_rhs, const double* const* d_rhs, double* const* d_x,
                      int64_t tag) {
    int admitted = 0;
    const int rc = guard([&] {
        if (!sv || !op || !d_rhs || !d_x) throw ArgError("null argument");
        for (int i = 0; i < n_rhs; ++i)
            if (!d_rhs[i] || !d_x[i]) throw ArgError("null vector");
        HIPCHK(hipSetDevice(O(op)->ctx->device));
        admitted = reinterpret_cast<ServerBase*>(sv)->submit(O(op), n_rhs, d_rhs, d_x, tag) ? 1 : 0;
    });
    return rc != CPM_OK ? rc : (admitted ? CPM_OK : CPM_BUSY);
}

int cpm_server_run(cpm_server* sv, int iterations, int max_done, int64_t* done_tags, cpm_solve_info* done_infos,
                   int* n_done, int* n_active) {
    return guard([&] {
        if (!sv || !n_done || iterations < 0 || max_done < 0 || (max_done > 0 && (!done_tags || !done_infos)))
            throw ArgError("null argument");
        ServerBase* s = reinterpret_cast<ServerBase*>(sv);
        s->run(iterations);
        int k = 0;
        while (k < max_done && !s->done.empty()) {
            const ServerBase::Done& d = s->done.front();
            done_tags[k] = d.tag;
            for (int j = 0; j < s->lanes; ++j)
                done_infos[(size_t)k * s->lanes + j] = j < (int)d.infos.size() ? d.infos[(size_t)j] : cpm_solve_info{};
            s->done.pop_front();
            ++k;
        }
        *n_done = k;
        if (n_active) *n_active = s->active() + (int)s->done.size();
    });
}

}  // extern "C"
