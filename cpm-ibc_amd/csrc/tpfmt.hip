// Tensor-product execution format of operators built on a tube.
//
// The reference stores E as a general sparse matrix: 64 (column, value)
// pairs per row for p = 3 (Interpolation.cpp:204-225), i.e. 768 bytes of
// FP64 per row, read twice per BiCGStab iteration.  Every E row is in fact
// the tensor product of three 1-D barycentric weight vectors over the 4x4x4
// grid nodes b + (i, j, k).  Once the tube's rows are renumbered in
// lexicographic grid order (z fastest), the nodes b + (i, j, 0..3) of a
// stencil line are four consecutive rows, so a row is described by 16 line
// starts and 12 weights (160 bytes in FP64, 112 in FP32).  Where an interior
// boundary moved some of a row's columns (PoissonWithIBC MoveToIBCColumns) or
// appended the Dirichlet identity column, an exception record lists the
// columns of all 64 slots and the extra entry.
//
// The builder derives the format from the reference-order ELL rows and the
// rows' query points, and accepts an operator only if every row is
// reproduced exactly: the same columns and bit-identical FP64 values
// (w_z w_y) w_x, the order of Interpolation.cpp:95-107.  Otherwise the caller
// keeps the ELL format.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <mutex>
#include <thread>
#include <cstdlib>
#include <string>
#include <vector>

#include "tube.h"
#include "krylov.h"

namespace cpmgpu {

namespace {

constexpr int kS = 4;   // p = 3

__global__ void k_lex_count(const int32_t* gidx, int64_t G, int* bcount) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int32_t v = i < G ? gidx[i] : -1;
    const int occ = (v >= 0 && v < kCand) ? 1 : 0;
    const unsigned long long m = __ballot(occ);
    __shared__ int ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Tube nodes in linear grid order: lex[pos] = tube index, rlex[tube index] = pos.
__global__ void k_lex_emit(const int32_t* gidx, int64_t G, const int* boff, int64_t N, int32_t* lex, int32_t* rlex,
                           int* err) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int32_t v = i < G ? gidx[i] : -1;
    const int occ = (v >= 0 && v < kCand) ? 1 : 0;
    const unsigned long long m = __ballot(occ);
    __shared__ int ws[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = __popcll(m);
    __syncthreads();
    if (!occ) return;
    int pos = boff[blockIdx.x];
    for (int k = 0; k < w; ++k) pos += ws[k];
    pos += __popcll(m & ((1ull << lane) - 1ull));
    if (pos >= N || v >= N) { atomicOr(err, 1); return; }
    lex[pos] = v;
    rlex[v] = pos;
}

struct TpArgs {
    int64_t D, N;
    const int* ecol;
    const double* eval;
    int ew;
    const int* lcol;
    const double* lval;
    int lw;
    const double* ldiag;
    const uint8_t* flag;
    const double* qpts;
    const int32_t* gidx;
    GridDev g;
    const int32_t* lex;
    const int32_t* rlex;
    double loff, ldu;
};

__device__ __forceinline__ int64_t to_internal(const TpArgs& a, int64_t c) { return c < a.N ? a.rlex[c] : c; }
__device__ __forceinline__ int64_t to_reference(const TpArgs& a, int64_t r) { return r < a.N ? a.lex[r] : r; }

// Error bits of the builder (any bit: the operator keeps the ELL format).
enum : int {
    TP_STENCIL = 2, TP_OUTSIDE = 4, TP_GAP = 8, TP_VALUE = 16, TP_TAIL = 32, TP_LVAL = 64, TP_LDIAG = 128
};

// Pass 1, one thread per internal row: line starts, weights, L columns, the
// row's moved-slot mask and its exception counts.
__global__ void k_tp_rows(TpArgs a, int* tpc, double* tpw64, float* tpw32, int* lcol_out, uint8_t* flag_out,
                          unsigned long long* mask_out, int* n_exc, int* n_mov, int* err) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.D) return;
    const int64_t D = a.D;
    const int64_t R = to_reference(a, r);
    const uint8_t f = a.flag[R];
    flag_out[r] = f;
    mask_out[r] = 0;
    n_exc[r] = 0;
    n_mov[r] = 0;
    if (f & ROW_IDENTITY) {
        for (int l = 0; l < kS * kS; ++l) tpc[l * D + r] = 0;
        for (int q = 0; q < 3 * kS; ++q) { tpw64[q * D + r] = 0.0; tpw32[q * D + r] = 0.0f; }
        for (int s = 0; s < 6; ++s) lcol_out[s * D + r] = -1;
        return;
    }
    if (f & ROW_NODIAG) { atomicOr(err, TP_LDIAG); return; }
    int64_t b[3];
    double w[3][5];
    if (!stencil_weights(a.g, kS - 1, a.qpts + 3 * R, b, w)) { atomicOr(err, TP_STENCIL); return; }
    const int64_t N1 = a.g.N1d;
    // sign of the row (ghost rows of Dirichlet curves are negated)
    double sign = 0.0;
    for (int s = 0; s < kS * kS * kS && sign == 0.0; ++s) {
        const int i = s / 16, j = (s / 4) % 4, k = s % 4;
        const double v = w[2][k] * w[1][j] * w[0][i];
        if (v == 0.0) continue;
        const double e = a.eval[s * D + R];
        sign = e == v ? 1.0 : (e == -v ? -1.0 : 2.0);
    }
    if (sign != 1.0 && sign != -1.0) { atomicOr(err, TP_VALUE); return; }
    unsigned long long mask = 0;
    int moved = 0;
    for (int i = 0; i < kS; ++i)
        for (int j = 0; j < kS; ++j) {
            int64_t c0 = -1;
            for (int k = 0; k < kS; ++k) {
                const int s = (i * kS + j) * kS + k;
                const int32_t gv = a.gidx[((b[0] + i) * N1 + (b[1] + j)) * N1 + (b[2] + k)];
                if (!(gv >= 0 && gv < a.N)) { atomicOr(err, TP_OUTSIDE); return; }
                const int64_t ci = a.rlex[gv];
                if (k == 0) c0 = ci;
                else if (ci != c0 + k) { atomicOr(err, TP_GAP); return; }
                const double v = w[2][k] * w[1][j] * (sign * w[0][i]);
                if (!(a.eval[s * D + R] == v)) { atomicOr(err, TP_VALUE); return; }
                if (a.ecol[s * D + R] != gv) { mask |= 1ull << s; ++moved; }
            }
            tpc[(i * kS + j) * D + r] = (int)c0;
        }
    for (int q = 0; q < kS; ++q) {
        const double wx = sign * w[0][q];
        tpw64[q * D + r] = wx;
        tpw64[(kS + q) * D + r] = w[1][q];
        tpw64[(2 * kS + q) * D + r] = w[2][q];
        tpw32[q * D + r] = (float)wx;
        tpw32[(kS + q) * D + r] = (float)w[1][q];
        tpw32[(2 * kS + q) * D + r] = (float)w[2][q];
    }
    bool extra = false;
    for (int s = kS * kS * kS; s < a.ew; ++s) {
        const double v = a.eval[s * D + R];
        const int c = a.ecol[s * D + R];
        if (v == 0.0 && c == R) continue;                 // padding
        if (s == kS * kS * kS) { extra = true; continue; }
        atomicOr(err, TP_TAIL);
        return;
    }
    for (int s = 0; s < 6; ++s) {
        int out = -1;
        if (s < a.lw) {
            const double v = a.lval[s * D + R];
            const int c = a.lcol[s * D + R];
            if (!(v == 0.0 && c == R)) {
                if (!(v == a.loff)) { atomicOr(err, TP_LVAL); return; }
                out = (int)to_internal(a, c);
            }
        }
        lcol_out[s * D + r] = out;
    }
    if (!(a.ldiag[R] == a.ldu)) { atomicOr(err, TP_LDIAG); return; }
    mask_out[r] = mask;
    n_exc[r] = (moved > 0 || extra) ? 1 : 0;
    n_mov[r] = moved;
}

// Pass 2: exception records (n_exc now holds exclusive offsets): the
// internal column of every slot, and the extra entry.
__global__ void k_tp_exc(TpArgs a, const int* exc_off, int64_t nexc_total, int* exc, int* ecx, int* xcol,
                         double* xval64, float* xval32) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.D) return;
    const int64_t D = a.D;
    const int64_t R = to_reference(a, r);
    const int e = exc_off[r];
    const int e_next = r + 1 < a.D ? exc_off[r + 1] : (int)nexc_total;
    if (e_next == e || (a.flag[R] & ROW_IDENTITY)) { exc[r] = -1; return; }
    exc[r] = e;
    for (int s = 0; s < kS * kS * kS; ++s) ecx[(int64_t)e * 64 + s] = (int)to_internal(a, a.ecol[s * D + R]);
    xcol[e] = -1;
    xval64[e] = 0.0;
    xval32[e] = 0.0f;
    if (a.ew > kS * kS * kS) {
        const int s = kS * kS * kS;
        const double v = a.eval[s * D + R];
        const int c = a.ecol[s * D + R];
        if (!(v == 0.0 && c == R)) {
            xcol[e] = (int)to_internal(a, c);
            xval64[e] = v;
            xval32[e] = (float)v;
        }
    }
}

__global__ void k_perm_full(const int32_t* lex, int64_t N, int64_t D, int32_t* perm) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < D) perm[r] = r < N ? lex[r] : (int32_t)r;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// The tube's lexicographic order, computed on first use (on the calling
// context's stream; the lock keeps concurrent first uses apart).
static void ensure_lex(Tube& t, Ctx* ctx) {
    static std::mutex lock;
    std::lock_guard<std::mutex> g(lock);
    if (t.lex.p) return;
    const hipStream_t st = ctx->stream;
    const int64_t G = t.N1d * t.N1d * t.N1d;
    const int64_t nb = (G + 255) / 256;
    Buffer bcount((size_t)(nb + 1) * 4), flags(64);
    Buffer lex((size_t)t.N * 4), rlex((size_t)t.N * 4);
    int* total = bcount.as<int>() + nb;
    HIPCHK(hipMemsetAsync(flags.p, 0, 64, st));
    hipLaunchKernelGGL(k_lex_count, dim3((unsigned)nb), dim3(256), 0, st, t.gidx.as<int32_t>(), G, bcount.as<int>());
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, bcount.as<int>(), nb, total);
    hipLaunchKernelGGL(k_lex_emit, dim3((unsigned)nb), dim3(256), 0, st, t.gidx.as<int32_t>(), G, bcount.as<int>(),
                       t.N, lex.as<int32_t>(), rlex.as<int32_t>(), flags.as<int>());
    HIPCHK(hipGetLastError());
    int h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&h[0], total, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&h[1], flags.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h[0] != t.N || h[1]) throw std::logic_error("tube index grid is inconsistent with the node list");
    t.lex = std::move(lex);
    t.rlex = std::move(rlex);
}

// CPMGPU_LDS_TILES selects the LDS tiles built with tensor-product
// operators: unset or "1", row tiles and shared-stencil tiles (these unless
// they would need more workgroups than the operator has row blocks); "rows",
// row tiles only; "0", none (every gather reads x directly).
static int lds_tile_mode() {
    const char* e = std::getenv("CPMGPU_LDS_TILES");
    if (!e) return 2;
    const std::string v(e);
    return v == "0" ? 0 : v == "rows" ? 1 : 2;
}

// The slots' weights in slot order (grouped copy of tpw).
template <typename T>
__global__ void k_slot_weights(const T* tpw, int64_t D, const int* grow, int64_t S, T* gw) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int64_t r = grow[s];
#pragma unroll
    for (int q = 0; q < 12; ++q) gw[s * 12 + q] = tpw[q * D + r];
}

// Tensor-product rows on the host, for the tile builders.
struct HostTp {
    int64_t D = 0, ne = 0;
    std::vector<uint8_t> flag;
    std::vector<int> tpc, exc, ecx, xcol;   // [16][D], [D], [ne][64], [ne]
    HostTp(int64_t D_, int64_t ne_, const Buffer& tpc_, const Buffer& exc_, const Buffer& ecx_, const Buffer& xcol_,
           const Buffer& flag_)
        : D(D_), ne(ne_), flag((size_t)D_), tpc((size_t)16 * D_), exc((size_t)D_),
          ecx((size_t)std::max<int64_t>(ne_, 1) * 64), xcol((size_t)std::max<int64_t>(ne_, 1)) {
        HIPCHK(hipMemcpy(flag.data(), flag_.p, flag.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(tpc.data(), tpc_.p, tpc.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(exc.data(), exc_.p, exc.size() * 4, hipMemcpyDeviceToHost));
        if (ne > 0) {
            HIPCHK(hipMemcpy(ecx.data(), ecx_.p, (size_t)ne * 64 * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(xcol.data(), xcol_.p, (size_t)ne * 4, hipMemcpyDeviceToHost));
        }
    }
    // the x rows row r's E row reads (kind 0: 16 lines of 4; exception rows:
    // the 64 slot columns and the extra column)
    template <typename F>
    void columns(int64_t r, F&& f) const {
        const int e = exc[(size_t)r];
        if (e >= 0) {
            for (int s = 0; s < 64; ++s) f(ecx[(size_t)e * 64 + s]);
            if (xcol[(size_t)e] >= 0) f(xcol[(size_t)e]);
            return;
        }
        for (int l = 0; l < 16; ++l)
            for (int k = 0; k < 4; ++k) f(tpc[(size_t)l * D + r] + k);
    }
};

static void upload(Buffer& b, const void* p, size_t bytes) {
    b = Buffer(bytes);
    HIPCHK(hipMemcpy(b.p, p, bytes, hipMemcpyHostToDevice));
}

template <typename Body>
static void parallel_for(int64_t n, Body&& body) {
    const int nth = (int)std::max<int64_t>(
        1, std::min<int64_t>({16, n, (int64_t)std::max(1u, std::thread::hardware_concurrency())}));
    std::vector<std::thread> th;
    for (int w = 0; w < nth; ++w)
        th.emplace_back([&, w] {
            for (int64_t i = n * w / nth; i < n * (w + 1) / nth; ++i) body(i);
        });
    for (auto& x : th) x.join();
}

// Shared-stencil tiles.  Plain rows are sorted by their stencil block (the
// first column of line (0, 0)) and rows with the same block are cut into
// items of up to kGrpRows rows; exception rows are single items placed by
// their smallest column, identity rows are items of up to kGrpRows rows at
// the end.  Tiles take items in that order while the tile's list of distinct
// x rows stays within kGrpLdsRows and its items within kBlock; inside a tile,
// plain items come first, by row count (larger first), then exception rows,
// then identity rows, so that a wave's threads take one code path and their
// loops over the rows of an item end together (which item a thread takes
// changes no row's arithmetic).  Returns false (nothing built) if the tiles
// would outnumber the operator's row blocks, which size every launch, or
// number fewer than kGrpMinTiles.
static bool build_group_tiles(Op& op, const HostTp& h, const Buffer& tpw64, const Buffer& tpw32) {
    const int64_t D = h.D;
    // CPMGPU_GROUP_ITEMS: fewer items per tile (more, smaller tiles); read once per process
    static const int max_items = [] {
        const char* e = std::getenv("CPMGPU_GROUP_ITEMS");
        return e ? std::max(1, std::min(kBlock, std::atoi(e))) : kBlock;
    }();
    const int64_t nblocks = (D + kBlock - 1) / kBlock;
    if (D >= (int64_t)kGrpOfsMask || nblocks < kGrpMinTiles) return false;
    struct Item {
        int r0, n, kind;   // rows trow[r0 .. r0 + n)
    };
    // non-identity rows by key with a counting sort (rows in increasing order
    // within a key, plain rows before exception rows): a plain row's key is its
    // stencil block (the first column of line (0, 0)), an exception row's its
    // smallest column; every key is a column, below D
    std::vector<int> key((size_t)D, -1), cnt((size_t)D + 1, 0), pcnt((size_t)D, 0), ident;
    int64_t nplain = 0, nexcr = 0;
    for (int64_t r = 0; r < D; ++r) {
        int k;
        if (h.flag[(size_t)r] & ROW_IDENTITY) {
            ident.push_back((int)r);
            continue;
        } else if (h.exc[(size_t)r] >= 0) {
            k = INT32_MAX;
            h.columns(r, [&](int c) { k = std::min(k, c); });
            ++nexcr;
        } else {
            k = h.tpc[(size_t)r];
            if (k >= 0 && k < D) ++pcnt[(size_t)k];
            ++nplain;
        }
        if (k < 0 || k >= D) throw std::logic_error("shared-stencil tiles: a row's key is out of range");
        key[(size_t)r] = k;
        ++cnt[(size_t)k + 1];
    }
    for (int64_t k = 0; k < D; ++k) cnt[(size_t)k + 1] += cnt[(size_t)k];
    std::vector<int> trow((size_t)D), fill(cnt.begin(), cnt.end() - 1);
    for (int pass = 0; pass < 2; ++pass)   // plain rows, then exception rows
        for (int64_t r = 0; r < D; ++r) {
            const int k = key[(size_t)r];
            if (k >= 0 && (h.exc[(size_t)r] >= 0) == (pass == 1)) trow[(size_t)fill[(size_t)k]++] = (int)r;
        }
    std::vector<Item> items;
    items.reserve((size_t)(nplain / kGrpRows + nexcr + (int64_t)ident.size() / kGrpRows + 2));
    int64_t items0 = 0;
    for (int64_t k = 0; k < D; ++k) {
        const int a = cnt[(size_t)k], b = cnt[(size_t)k + 1], pe = a + pcnt[(size_t)k];
        for (int c = a; c < pe; c += kGrpRows) {
            items.push_back({c, std::min(kGrpRows, pe - c), 0});
            ++items0;
        }
        for (int c = pe; c < b; ++c) items.push_back({c, 1, 1});
    }
    trow.resize((size_t)cnt[(size_t)D]);
    for (size_t c = 0; c < ident.size(); c += kGrpRows) {
        const int n = (int)std::min<size_t>(kGrpRows, ident.size() - c);
        items.push_back({(int)trow.size(), n, 2});
        for (int g = 0; g < n; ++g) trow.push_back(ident[c + g]);
    }
    const int64_t nitem = (int64_t)items.size();
    // greedy tiles over contiguous item ranges, one host thread each (a range
    // ends its last tile early); stamp[c] = tile whose list holds column c
    const int nparts = (int)std::max<int64_t>(
        1, std::min<int64_t>({16, nitem / 4096 + 1, (int64_t)std::max(1u, std::thread::hardware_concurrency())}));
    std::vector<std::vector<int>> starts((size_t)nparts);
    parallel_for(nparts, [&](int64_t part) {
        const int64_t i0 = nitem * part / nparts, i1 = nitem * (part + 1) / nparts;
        std::vector<int> stamp((size_t)D, -1), fresh;
        std::vector<int>& first = starts[(size_t)part];
        if (i1 > i0) first.push_back((int)i0);
        int tile = 0, list = 0, nin = 0;
        for (int64_t i = i0; i < i1;) {
            const Item& it = items[(size_t)i];
            fresh.clear();
            if (it.kind != 2)
                h.columns(trow[(size_t)it.r0], [&](int c) {
                    if (stamp[(size_t)c] != tile) {
                        stamp[(size_t)c] = tile;
                        fresh.push_back(c);
                    }
                });
            if (nin > 0 && (list + (int)fresh.size() > kGrpLdsRows || nin + 1 > max_items)) {
                for (int c : fresh) stamp[(size_t)c] = -1;
                first.push_back((int)i);
                ++tile;
                list = 0;
                nin = 0;
                continue;   // the item again, in a new tile
            }
            list += (int)fresh.size();
            ++nin;
            ++i;
        }
    });
    std::vector<int> first;
    for (const auto& v : starts) first.insert(first.end(), v.begin(), v.end());
    if (first.empty()) first.push_back(0);
    first.push_back((int)nitem);
    const int64_t ngt = (int64_t)first.size() - 1;
    // with fewer tiles the row tiles' 2.3 times as many workgroups fill the
    // GPU better (measured on spheres of 41,870 and 65,294 rows)
    if (ngt > nblocks || ngt < kGrpMinTiles) return false;
    // per tile: items by row count, the sorted list, positions
    std::vector<std::vector<int>> lists((size_t)ngt);
    std::vector<int> order((size_t)nitem);
    const int64_t nip = ngt * kBlock;   // item places: tile t owns t kBlock .. (t + 1) kBlock - 1
    std::vector<int> gpl((size_t)8 * nip, 0);
    std::vector<int16_t> tlx((size_t)std::max<int64_t>(h.ne, 1) * kExcPos, -1);
    std::atomic<int> bad{0};
    parallel_for(ngt, [&](int64_t t) {
        const int i0 = first[(size_t)t], i1 = first[(size_t)t + 1];
        for (int i = i0; i < i1; ++i) order[(size_t)i] = i;
        auto rank = [&](int u) {   // plain items by row count, then exception rows, then identity rows
            const Item& it = items[(size_t)u];
            return it.kind == 0 ? kGrpRows - it.n : kGrpRows + it.kind;
        };
        std::stable_sort(order.begin() + i0, order.begin() + i1, [&](int u, int v) { return rank(u) < rank(v); });
        std::vector<int>& u = lists[(size_t)t];
        for (int i = i0; i < i1; ++i)
            if (items[(size_t)i].kind != 2) h.columns(trow[(size_t)items[(size_t)i].r0], [&](int c) { u.push_back(c); });
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if ((int64_t)u.size() > kGrpLdsRows || i1 - i0 > kBlock) bad = 1;
        auto pos = [&](int c) { return (int16_t)(std::lower_bound(u.begin(), u.end(), c) - u.begin()); };
        for (int q = i0; q < i1 && q - i0 < kBlock; ++q) {
            const Item& it = items[(size_t)order[(size_t)q]];
            const int r = trow[(size_t)it.r0];
            const size_t qp = (size_t)t * kBlock + (size_t)(q - i0);   // the item's place
            if (it.kind == 0) {
                for (int l = 0; l < 16; ++l) {
                    const int16_t p = pos(h.tpc[(size_t)l * D + r]);
                    if (p < 0 || p + 3 >= (int)u.size()) bad = 1;
                    gpl[(size_t)(l >> 1) * nip + qp] |= (int)((unsigned)(uint16_t)p << (16 * (l & 1)));
                }
            } else if (it.kind == 1) {
                const int e = h.exc[(size_t)r];
                for (int s = 0; s < 64; ++s) tlx[(size_t)e * kExcPos + s] = pos(h.ecx[(size_t)e * 64 + s]);
                if (h.xcol[(size_t)e] >= 0) tlx[(size_t)e * kExcPos + 64] = pos(h.xcol[(size_t)e]);
                gpl[qp] = e;   // read with the line positions, a step before the record
            } else {
                for (int g = 0; g < kGrpRows; ++g)   // the rows (kind 2 reads their x values after the barrier)
                    gpl[(size_t)g * nip + qp] = g < it.n ? trow[(size_t)it.r0 + g] : -1;
            }
        }
    });
    if (bad) throw std::logic_error("shared-stencil tiles: a list or position is out of range");
    // slots in final item order, item places padded to kBlock per tile with
    // empty items (kind 3, no slot)
    std::vector<int> gofs((size_t)nip + 1), grow;
    grow.reserve((size_t)D);
    for (int64_t t = 0; t < ngt; ++t) {
        const int i0 = first[(size_t)t], i1 = first[(size_t)t + 1];
        for (int j = 0; j < kBlock; ++j) {
            const size_t qp = (size_t)t * kBlock + (size_t)j;
            if (j >= i1 - i0) {
                gofs[qp] = (int)grow.size() | (3 << kGrpKindShift);
                continue;
            }
            const Item& it = items[(size_t)order[(size_t)(i0 + j)]];
            gofs[qp] = (int)grow.size() | (it.kind << kGrpKindShift);
            for (int g = 0; g < it.n; ++g) grow.push_back(trow[(size_t)it.r0 + g]);
        }
    }
    gofs[(size_t)nip] = (int)grow.size();
    if ((int64_t)grow.size() != D) throw std::logic_error("shared-stencil tiles: not every row has a slot");
    // lists at kGrpLdsRows places per tile, the unused places repeating the
    // last entry (the gather reads every place without a branch); an entry is
    // the x row with the row's class in bit 30 (the gather's kListPlainBit)
    std::vector<int> tln((size_t)ngt), tlr((size_t)ngt * kGrpLdsRows, 0);
    op.grp_list = 0;
    for (int64_t t = 0; t < ngt; ++t) {
        const std::vector<int>& u = lists[(size_t)t];
        tln[(size_t)t] = (int)u.size();
        int* dst = tlr.data() + (size_t)t * kGrpLdsRows;
        // bit 30: the row's P^-1 leaves it as it is (identity or no-diagonal row)
        for (size_t k = 0; k < u.size(); ++k)
            dst[k] = u[k] | ((h.flag[(size_t)u[k]] & (ROW_IDENTITY | ROW_NODIAG)) ? 1 << 30 : 0);
        std::fill(dst + u.size(), dst + kGrpLdsRows, u.empty() ? 0 : dst[u.size() - 1]);
        op.grp_list += tln[(size_t)t];
    }
    op.grp_items = nitem;
    op.grp_items0 = items0;
    op.grp_rows0 = nplain;
    op.grp_exc = nexcr;
    op.grp_ident = (int64_t)ident.size();
    op.grp_ident_items = ((int64_t)ident.size() + kGrpRows - 1) / kGrpRows;
    upload(op.gln, tln.data(), tln.size() * 4);
    upload(op.glr, tlr.data(), tlr.size() * 4);
    upload(op.glx, tlx.data(), tlx.size() * 2);
    upload(op.gofs, gofs.data(), gofs.size() * 4);
    upload(op.grow, grow.data(), grow.size() * 4);
    upload(op.gpl, gpl.data(), gpl.size() * 4);
    op.gw64 = Buffer((size_t)12 * D * 8);
    op.gw32 = Buffer((size_t)12 * D * 4);
    const hipStream_t st = op.ctx->stream;
    hipLaunchKernelGGL(k_slot_weights<double>, dim3(nblk(D)), dim3(256), 0, st, tpw64.as<double>(), D,
                       op.grow.as<int>(), D, op.gw64.as<double>());
    hipLaunchKernelGGL(k_slot_weights<float>, dim3(nblk(D)), dim3(256), 0, st, tpw32.as<float>(), D,
                       op.grow.as<int>(), D, op.gw32.as<float>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    op.ngt = (int)ngt;
    op.nitem = (int)nip;
    op.nslot = D;
    return true;
}

// LDS tiles: for every block of kBlock internal rows, the sorted distinct x
// rows its E rows read, and for each stencil line the position of its first
// row in that list (the line's four rows are consecutive integers, hence
// consecutive in the list); rows with an exception record get the positions
// of their 64 columns and of the extra entry instead (tlx, kExcPos per
// record, -1 = no extra entry).  Tiles whose list would exceed kLdsRows keep
// reading x directly (tln = -1).
static void build_lds_tiles(Op& op, const HostTp& h) {
    const int64_t D = h.D, ne = h.ne;
    const std::vector<uint8_t>& hf = h.flag;
    const std::vector<int>&hc = h.tpc, &he = h.exc, &hx = h.ecx, &hxc = h.xcol;
    const int64_t ntiles = (D + kBlock - 1) / kBlock;
    std::vector<int> tln((size_t)ntiles), tlo((size_t)ntiles), tlr;
    std::vector<int16_t> tpl((size_t)kS * kS * D, 0), tlx((size_t)std::max<int64_t>(ne, 1) * kExcPos, -1);
    std::vector<std::vector<int>> lists((size_t)ntiles);
    // tiles are independent: host threads over contiguous tile ranges
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>({16, ntiles, (int64_t)std::max(1u, std::thread::hardware_concurrency())}));
    auto parallel = [&](auto&& body) {
        std::vector<std::thread> th;
        for (int w = 0; w < nth; ++w)
            th.emplace_back([&, w] {
                for (int64_t tile = ntiles * w / nth; tile < ntiles * (w + 1) / nth; ++tile) body(tile);
            });
        for (auto& x : th) x.join();
    };
    // 1. each tile's sorted distinct x rows
    parallel([&](int64_t tile) {
        const int64_t r0 = tile * kBlock, r1 = std::min<int64_t>(D, r0 + kBlock);
        std::vector<int>& u = lists[(size_t)tile];
        for (int64_t r = r0; r < r1; ++r) {
            const int e = he[(size_t)r];
            if (e >= 0) {
                for (int s = 0; s < 64; ++s) u.push_back(hx[(size_t)e * 64 + s]);
                if (hxc[(size_t)e] >= 0) u.push_back(hxc[(size_t)e]);
                continue;
            }
            for (int l = 0; l < kS * kS; ++l) {
                const int c0 = hc[(size_t)l * D + r];
                for (int k = 0; k < kS; ++k) u.push_back(c0 + k);
            }
        }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
    });
    // 2. offsets, counts, the concatenated lists
    for (int64_t tile = 0; tile < ntiles; ++tile) {
        const int64_t r0 = tile * kBlock, r1 = std::min<int64_t>(D, r0 + kBlock);
        const std::vector<int>& u = lists[(size_t)tile];
        tlo[(size_t)tile] = (int)tlr.size();
        const bool fits = (int64_t)u.size() <= kLdsRows;
        for (int64_t r = r0; r < r1; ++r)
            if (!(hf[(size_t)r] & ROW_IDENTITY)) ++op.tile_rows[(fits ? 0 : 2) + (he[(size_t)r] >= 0 ? 1 : 0)];
        if (!fits) {
            tln[(size_t)tile] = -1;
            continue;
        }
        tln[(size_t)tile] = (int)u.size();
        op.tile_list += (int64_t)u.size();
        tlr.insert(tlr.end(), u.begin(), u.end());
    }
    // 3. positions of every line (or exception column) in its tile's list
    parallel([&](int64_t tile) {
        if (tln[(size_t)tile] < 0) return;
        const int64_t r0 = tile * kBlock, r1 = std::min<int64_t>(D, r0 + kBlock);
        const std::vector<int>& u = lists[(size_t)tile];
        auto pos = [&](int c) { return (int16_t)(std::lower_bound(u.begin(), u.end(), c) - u.begin()); };
        for (int64_t r = r0; r < r1; ++r) {
            const int e = he[(size_t)r];
            if (e >= 0) {
                for (int s = 0; s < 64; ++s) tlx[(size_t)e * kExcPos + s] = pos(hx[(size_t)e * 64 + s]);
                if (hxc[(size_t)e] >= 0) tlx[(size_t)e * kExcPos + 64] = pos(hxc[(size_t)e]);
                continue;
            }
            for (int l = 0; l < kS * kS; ++l) tpl[(size_t)l * D + r] = pos(hc[(size_t)l * D + r]);
        }
    });
    tlr.push_back(0);   // an entry at every tile's list offset: the gather clamps its list reads into the list
    auto up = [](Buffer& b, const void* p, size_t bytes) {
        b = Buffer(bytes);
        HIPCHK(hipMemcpy(b.p, p, bytes, hipMemcpyHostToDevice));
    };
    up(op.tln, tln.data(), tln.size() * 4);
    up(op.tlo, tlo.data(), tlo.size() * 4);
    up(op.tlr, tlr.data(), tlr.size() * 4);
    up(op.tpl, tpl.data(), tpl.size() * 2);
    up(op.tlx, tlx.data(), tlx.size() * 2);
}

bool op_build_tp(Op& op, Tube& t, const TpSource& src) {
    if (t.p != 3 || t.dim != 3 || src.ew < kS * kS * kS || src.lw > 6 || op.nD != 0) return false;
    if (src.D > (int64_t)INT32_MAX / 2) return false;
    ensure_lex(t, op.ctx);
    const hipStream_t st = op.ctx->stream;
    const int64_t D = src.D;
    TpArgs a{};
    a.D = D;
    a.N = src.N;
    a.ecol = src.ecol;
    a.eval = src.eval;
    a.ew = src.ew;
    a.lcol = src.lcol;
    a.lval = src.lval;
    a.lw = src.lw;
    a.ldiag = src.ldiag;
    a.flag = src.flag;
    a.qpts = src.qpts;
    a.gidx = t.gidx.as<int32_t>();
    a.g = GridDev{t.N1d, t.xstart, t.dx, t.r_t};
    a.lex = t.lex.as<int32_t>();
    a.rlex = t.rlex.as<int32_t>();
    a.loff = 1.0 / (t.dx * t.dx);          // FDMatrices.cpp:25-49: 1 / h^2 off the diagonal
    a.ldu = -6.0 / (t.dx * t.dx);          // -2 dim / h^2 on it

    Buffer tpc((size_t)16 * D * 4), tpw64((size_t)12 * D * 8), tpw32((size_t)12 * D * 4), lcol((size_t)6 * D * 4),
        flag((size_t)D), mask((size_t)D * 8), nexc((size_t)(D + 1) * 4), nmov((size_t)(D + 1) * 4), flags(64),
        perm((size_t)D * 4), exc((size_t)D * 4);
    HIPCHK(hipMemsetAsync(flags.p, 0, 64, st));
    int* tot = flags.as<int>() + 4;
    hipLaunchKernelGGL(k_tp_rows, dim3(nblk(D)), dim3(256), 0, st, a, tpc.as<int>(), tpw64.as<double>(),
                       tpw32.as<float>(), lcol.as<int>(), flag.as<uint8_t>(), mask.as<unsigned long long>(),
                       nexc.as<int>(), nmov.as<int>(), flags.as<int>());
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, nexc.as<int>(), D, tot);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, nmov.as<int>(), D, tot + 1);
    HIPCHK(hipGetLastError());
    int h[6] = {0};
    HIPCHK(hipMemcpyAsync(h, flags.p, 24, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h[0]) return false;
    const int64_t ne = h[4], nm = h[5];
    Buffer ecx((size_t)std::max<int64_t>(ne, 1) * 64 * 4), xcol((size_t)std::max<int64_t>(ne, 1) * 4),
        xval64((size_t)std::max<int64_t>(ne, 1) * 8), xval32((size_t)std::max<int64_t>(ne, 1) * 4);
    hipLaunchKernelGGL(k_tp_exc, dim3(nblk(D)), dim3(256), 0, st, a, nexc.as<int>(), ne, exc.as<int>(), ecx.as<int>(),
                       xcol.as<int>(), xval64.as<double>(), xval32.as<float>());
    hipLaunchKernelGGL(k_perm_full, dim3(nblk(D)), dim3(256), 0, st, t.lex.as<int32_t>(), src.N, D, perm.as<int32_t>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));

    if (const int mode = lds_tile_mode()) {
        const HostTp h(D, ne, tpc, exc, ecx, xcol, flag);
        build_lds_tiles(op, h);
        op.grp_mode = mode == 2;   // shared-stencil tiles: on first use (op_ensure_group_tiles)
    }
    op.fmt = OP_TP;
    op.ew = kS * kS * kS;
    op.lw = 6;
    op.loff = a.loff;
    op.ldu = a.ldu;
    op.nexc = ne;
    op.nmoved = nm;
    op.tpc = std::move(tpc);
    op.tpw64 = std::move(tpw64);
    op.tpw32 = std::move(tpw32);
    op.lcol = std::move(lcol);
    op.flag = std::move(flag);
    op.perm = std::move(perm);
    op.exc = std::move(exc);
    op.ecx = std::move(ecx);
    op.xcol = std::move(xcol);
    op.xval64 = std::move(xval64);
    op.xval32 = std::move(xval32);
    op.ecol = Buffer();
    op.eval64 = Buffer();
    op.eval32 = Buffer();
    op.lval64 = Buffer();
    op.lval32 = Buffer();
    op.ldiag64 = Buffer();
    op.ldiag32 = Buffer();
    return true;
}

void op_ensure_group_tiles(Op& op) {
    if (op.grp_tried || !op.grp_mode || op.fmt != OP_TP) return;
    op.grp_tried = true;
    const HostTp h(op.D, op.nexc, op.tpc, op.exc, op.ecx, op.xcol, op.flag);
    build_group_tiles(op, h, op.tpw64, op.tpw32);
}

void ensure_group_tiles_for(const std::vector<Op*>& ops, int tsize, int nr) {
    grp_refresh_override();
    if (!grp_build(tsize, nr)) return;
    for (Op* o : ops) op_ensure_group_tiles(*o);
}

}  // namespace cpmgpu
