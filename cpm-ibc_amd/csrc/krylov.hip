// Batched BiCGStab for the closest-point screened-Poisson operator.
//
// Restates the reference's right-preconditioned BiCGStab (src/BiCGStab.h:30-108)
// and its matrix-free operator (src/GeneralizedBiCGStabSystem.h:191-261) as
// bandwidth-bound kernels; one iteration is (B = BiCGStab.h,
// G = GeneralizedBiCGStabSystem.h)
//
//   pupdate   p = (p - w v) beta + r ; y = P^-1 p                 (B:73-77)
//   egather   tmp = E y          (identity rows copy)              (G:235-261)
//   lfinish   v = m y + n (Ld y + Loff tmp) ; partial (r0, v)       (G:93-115,166-189; B:78)
//   fin       alpha = rho / (r0, v)                                (B:79)
//   supdate   s = r - alpha v ; z = P^-1 s                         (B:81-83)
//   egather   tmp = E z
//   lfinish   t = m z + n (Ld z + Loff tmp) ; partials (t,t),(t,s)  (B:84-86)
//   fin       omega = (t, s) / (t, t)                              (B:87-94)
//   xrupdate  x += alpha y ; x += w z ; r = s - w t ; (r,r),(r0,r)  (B:96-98)
//   fin       loop test (B:59), rho (B:61), restart test (B:62), beta (B:73)
//
// Reductions: every workgroup writes its partial sums; a one-workgroup-per-
// instance "fin" kernel adds them in a fixed order and runs the scalar part
// of the recurrence on the device, so iterations need no host round trip and
// results are reproducible run to run.  (No device-scope fences: on the
// multi-XCD MI355X a release/acquire at agent scope writes back / invalidates
// L2 and costs more than the whole kernel.)
//
// One launch covers every instance of a batch: workgroups map to (instance,
// row block) through a segment table; finished instances are skipped.
// Right-hand sides of one operator are interleaved [row][rhs] so one read of
// an E slot serves all of them.  Each rhs lane has its own state; a kernel
// acts on the lanes whose status is in its mask.
//
// Precision CPM_FP32_REFINE: iterative refinement with FP32 BiCGStab inner
// solves, run as one continuous device-side recurrence.  The solution is
// accumulated in FP64; when the FP32 recurrence residual has dropped by
// delta (default 1e-2) since the last step, or meets the tolerance, or hits
// the rho-restart rule, the lane takes a refinement step inside the next
// iteration's kernels (no extra launches):
//
//   pupdate   x64 += x32 / scale ; x32 = 0
//   egather_upd  tmp64 = E64 x64   (a launch of its own after the FP32 gather)
//   lfinish   r64 = b - A64 x64 ; partial (r64, r64)
//   fin       converged on the true residual?  else scale = 1 / ||r64||
//   supdate   r32 = r64 scale ; r0 = r32 ; p = v = 0
//   lfinish   partial (r32, r32)
//   fin       rho = r0sq = (r32, r32) ; the lane rejoins the next iteration
//
// The FP64 vectors (x64, r64, b, tmp64) are lane-major within a segment,
// the FP32 ones interleave the lanes per row.
//
// (Keeping p across steps, as in reliable-update schemes for one operator,
// diverges here once the residual approaches the FP32 operator's own
// rounding: the FP32 and FP64 operators differ in the last bits.)
#include <hip/hip_runtime.h>

#include <atomic>

#include <cmath>
#include <cstdlib>
#include <string>
#include <type_traits>

#include "engine.h"
#include "krylov.h"

namespace cpmgpu {

// ---------------------------------------------------------------- helpers
template <typename T, int NR> struct VT { typedef T __attribute__((ext_vector_type(NR))) type; };
template <typename T> struct VT<T, 1> { typedef T type; };

// Every array lives in global memory, but pointers taken from the segment
// table reach the compiler as generic pointers and would be accessed with
// flat instructions; these helpers address them as global memory.
#define CPM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
    return *(const CPM_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ void stg(T* p, const T& v) {
    *(CPM_GLOBAL T*)p = v;
}
template <typename T, int NR>
__device__ __forceinline__ typename VT<T, NR>::type vload(const T* p) {
    return ldg(reinterpret_cast<const typename VT<T, NR>::type*>(p));
}
template <typename T, int NR>
__device__ __forceinline__ void vstore(T* p, const typename VT<T, NR>::type& v) {
    stg(reinterpret_cast<typename VT<T, NR>::type*>(p), v);
}
template <typename V> __device__ __forceinline__ auto lane(const V& v, int k) { return v[k]; }
__device__ __forceinline__ double lane(double v, int) { return v; }
__device__ __forceinline__ float lane(float v, int) { return v; }
template <typename V, typename S> __device__ __forceinline__ void setlane(V& v, int k, S s) { v[k] = s; }
__device__ __forceinline__ void setlane(double& v, int, double s) { v = s; }
__device__ __forceinline__ void setlane(float& v, int, float s) { v = s; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup sums of NV*NR per-thread values.  Partials are stored
// component-major, part[c * nblocks + lb], so that the finaliser reads each
// component with coalesced loads.
// Wave sums of C = 2^L values per lane by a butterfly reduce-scatter: at step
// s (lane offset 32 >> s) a lane keeps the half of its values that its bit of
// the offset selects, summed with its partner's, so after L steps lane l holds
// component (l >> (6 - L)) & (C - 1) summed over 2^L lanes; the last 6 - L
// steps sum those groups.  2 (C - 1) + 6 - L exchanges instead of 6 C.
// Returns the lane's component total.
template <int C>
__device__ __forceinline__ double wave_reduce_scatter(double (&v)[C], int lanei, int& comp) {
    constexpr int L = C == 1 ? 0 : C == 2 ? 1 : C == 4 ? 2 : C == 8 ? 3 : 4;
    static_assert((1 << L) == C, "component count must be 1, 2, 4, 8 or 16");
#pragma unroll
    for (int st = 0; st < L; ++st) {
        const int o = 32 >> st, m = C >> st, h = m >> 1;
        const bool up = (lanei & o) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            // both halves cross over and the lane selects its sum: a select
            // between two array elements would be folded into a dynamically
            // indexed load, and the array moved to LDS
            const double lo = v[j] + __shfl_xor(v[j], o, 64);
            const double hi = v[j + h] + __shfl_xor(v[j + h], o, 64);
            v[j] = up ? hi : lo;
        }
    }
    double w = v[0];
#pragma unroll
    for (int o = 32 >> L; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    comp = (lanei >> (6 - L)) & (C - 1);
    return w;
}

template <int NV, int NR>
__device__ void block_partials(const double (&v)[NV][NR], double* part, int lb, int nblocks) {
    constexpr int C = NV * NR;
    __shared__ double red[kBlock / 64][C];
    const int lanei = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double flat[C];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) flat[a * NR + k] = v[a][k];
    int comp;
    const double w = wave_reduce_scatter<C>(flat, lanei, comp);
    if ((lanei & ((64 / C) - 1)) == 0) red[wave][comp] = w;
    __syncthreads();
    if (threadIdx.x < NV * NR) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
        stg(&part[(size_t)threadIdx.x * nblocks + lb], s);
    }
}

constexpr int kXcdRun = CPMGPU_XCD_RUN;

template <typename T>
__device__ __forceinline__ const Seg<T>& seg_of(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, int& lb, int64_t& row) {
    // Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
    // b mod 8).  Renumber them so that each XCD takes runs of kXcdRun
    // consecutive row blocks: the x rows a run's stencils gather then come
    // from that XCD's own L2 instead of every XCD fetching all of them, and
    // costly rows (interior-boundary rows sit together at the end) still
    // spread over all XCDs.  The tail that does not fill whole runs keeps
    // the plain order.
    const int b = (int)blockIdx.x, n = (int)gridDim.x;
    const int nfull = n / (8 * kXcdRun) * (8 * kXcdRun);
    int vb = b;
    if (b < nfull) {
        const int xcd = b % 8, k = b / 8;
        vb = ((k / kXcdRun) * 8 + xcd) * kXcdRun + k % kXcdRun;
    }
    if (!bseg) {   // one segment: no map (a dependent load less)
        lb = vb;
        row = (int64_t)lb * kBlock + threadIdx.x;
        return segs[0];
    }
    const int e = bseg[vb];
    lb = e >> kSegBits;
    row = (int64_t)lb * kBlock + threadIdx.x;
    return segs[e & (kMaxSegs - 1)];
}

// Lane masks: bit s set = act on lanes whose status is s.
constexpr unsigned M_RUN = 1u << ST_RUN;
constexpr unsigned M_RESTART = 1u << ST_RESTART;
constexpr unsigned M_UPDATE = (1u << ST_UPDATE) | (1u << ST_UPDATE_MAXIT);
constexpr unsigned M_RESCALE = 1u << ST_RESCALE;
constexpr unsigned M_READY = 1u << ST_READY;
constexpr unsigned M_ALL = 0xffffffffu;
// The iteration's stages (the FP64 loop only ever has RUN lanes among these).
constexpr unsigned M_STAGE1 = M_RUN | M_UPDATE;    // pupdate .. fin alpha
constexpr unsigned M_STAGE2 = M_RUN | M_RESCALE;   // supdate, lfinish t .. fin omega
constexpr unsigned M_STAGE3 = M_RUN | M_READY;     // fin of the loop test

template <typename T>
__device__ __forceinline__ bool lane_in(const Seg<T>& sg, int k, unsigned mask) {
    return k < sg.nr && ((mask >> ldg(&sg.st[k].status)) & 1u);
}

// The FP64 vectors of mixed-precision segments are lane-major: lane k is a
// run of D values at off + k D, so a refinement step moves only the data of
// the lanes that refine and its gathers read with stride 1.
template <typename T>
__device__ __forceinline__ size_t lm_at(const Seg<T>& sg, int k, int64_t row) {
    return (size_t)(sg.off + (int64_t)k * sg.op.D + row);
}

template <typename T>
__device__ __forceinline__ bool seg_any(const Seg<T>& sg, unsigned mask) {
    bool a = false;
    for (int k = 0; k < sg.nr; ++k) a |= ((mask >> ldg(&sg.st[k].status)) & 1u) != 0;
    return a;
}

// The statuses of a segment's lanes, read once per kernel with all loads in
// flight together: a group always owns kMaxNR states, so every load is valid
// (testing lane by lane up to nr waits for each scalar load in turn, several
// microseconds before the first vector load of a kernel).  nr = 0: no lane.
template <int NR>
struct LaneStat {
    int s[NR];
    int nr;
    __device__ __forceinline__ unsigned of(unsigned mask) const {   // bit k: lane k's status is in mask
        unsigned b = 0;
#pragma unroll
        for (int k = 0; k < NR; ++k) b |= ((mask >> s[k]) & 1u) << k;
        return b & ((1u << nr) - 1u);
    }
};
template <int NR, typename T>
__device__ __forceinline__ LaneStat<NR> lane_stat(const Seg<T>& sg, bool valid = true) {
    LaneStat<NR> L;
#pragma unroll
    for (int k = 0; k < NR; ++k) L.s[k] = valid ? ldg(&sg.st[k].status) : 0;
    L.nr = valid ? sg.nr : 0;
    return L;
}
template <int NR>
__device__ __forceinline__ bool lane_in(const LaneStat<NR>& L, int k, unsigned mask) { return (L.of(mask) >> k) & 1u; }
template <int NR>
__device__ __forceinline__ bool seg_any(const LaneStat<NR>& L, unsigned mask) { return L.of(mask) != 0; }

// Diagonal_Inverse, GeneralizedBiCGStabSystem.h:303-338 (the branch the
// reference runs: Compute() is never called).  FP64 divides, as the
// reference does.  FP32 (a precision of this library's own, refined against
// FP64) multiplies by the segment's reciprocals of its trailing-row and (in
// the tensor-product format) its one plain diagonal value, formed once when
// the segment is set up: a division per value costs ~10 instructions where
// the gathers form P^-1 of every x row they load.
// TP: the operator is known to be in the tensor-product format.
// The diagonal of A = m I + n L at a row whose L_ii is ld.
template <typename T>
__device__ __forceinline__ T diag_denom(const Seg<T>& sg, T ld) { return (T)sg.m + ld * (T)sg.n; }

template <typename T, bool TP = false>
__device__ __forceinline__ T diag_inverse(const Seg<T>& sg, int64_t row, uint8_t f, T val) {
    constexpr bool kRecip = std::is_same<T, float>::value;
    const int64_t dof = sg.op.D - sg.op.nD;
    if (row >= dof) {   // :335-338 overwrite the trailing rows last
        if constexpr (kRecip) return val * sg.tinv;
        return val / (T)(sg.m - sg.n);
    }
    if (f & ROW_IDENTITY) return val;
    if (f & ROW_NODIAG) return val;
    if (TP || sg.op.fmt == OP_TP) {
        if constexpr (kRecip) return val * sg.dinv;
        return val / diag_denom(sg, sg.op.ldu);
    }
    return val / diag_denom(sg, ldg(&sg.op.ldiag[row]));
}

constexpr double kJacobiOmega = 2.0 / 3.0;   // GeneralizedBiCGStabSystem.h:364

// Precondition, GeneralizedBiCGStabSystem.h:350-385.  Jacobi: the first of the
// precond_iters damped sweeps from result = 0 (A 0 = 0), result = omega
// D^-1 x on the rows below dofs and 0 on the trailing nD rows; the further
// sweeps (k_jacobi) need an operator apply each.
template <typename T, bool TP = false>
__device__ __forceinline__ T precond_apply(const Seg<T>& sg, int64_t row, uint8_t f, T val) {
    if (sg.precond == 1 /*diagonal*/) return diag_inverse<T, TP>(sg, row, f, val);
    if (sg.precond == 2 /*jacobi*/) {
        if (sg.precond_iters <= 0 || row >= sg.op.D - sg.op.nD) return (T)0;
        return (T)kJacobiOmega * diag_inverse<T, TP>(sg, row, f, val);
    }
    return val;
}

// P^-1 of a row's NR values, lane by lane with precond_apply: the iteration
// forms y = P^-1 p and z = P^-1 s where it reads them (the gathers' LDS
// fills, lfinish, xrupdate) instead of storing them, with the same
// operations, so the results are those of stored vectors to the bit.
template <typename T, int NR, bool TP = false>
__device__ __forceinline__ typename VT<T, NR>::type precond_row(const Seg<T>& sg, int64_t row, uint8_t f,
                                                                const typename VT<T, NR>::type& v) {
    typename VT<T, NR>::type y;
#pragma unroll
    for (int k = 0; k < NR; ++k) setlane(y, k, precond_apply<T, TP>(sg, row, f, lane(v, k)));
    return y;
}

// P^-1 of one row as factors its lanes share, found once per row: y = c ((v
// / dd) mul), or 0 where `zero`.  A division by 1 and products with 1 change
// no value, so for every preconditioner, row class and format this performs
// precond_apply's operations in its order (the vector kernels' form of it:
// precond_apply's branches and scalar loads, taken for each of a row's
// lanes, made up two thirds of their instructions).  div: whether the
// segment's rows divide (FP64, or FP32 in the ELL format), uniform per
// segment.
template <typename T>
struct PRow {
    T dd, mul, c;
    bool zero;
    __device__ __forceinline__ T apply(T v, bool div) const {
        T q = div ? v / dd : v;
        q = q * mul;
        q = c * q;
        return zero ? (T)0 : q;
    }
};
template <typename T>
__device__ __forceinline__ bool prow_div(const Seg<T>& sg) {
    return !std::is_same<T, float>::value || sg.op.fmt != OP_TP;
}
template <typename T>
__device__ __forceinline__ PRow<T> prow_of(const Seg<T>& sg, int64_t row, uint8_t f) {
    constexpr bool kRecip = std::is_same<T, float>::value;
    PRow<T> p{(T)1, (T)1, (T)1, false};
    if (sg.precond != 1 && sg.precond != 2) return p;   // none: y = v
    const bool trail = row >= sg.op.D - sg.op.nD;
    if (sg.precond == 2) {   // Jacobi's first sweep: omega D^-1 v, 0 on the trailing rows
        p.c = (T)kJacobiOmega;
        p.zero = sg.precond_iters <= 0 || trail;
    }
    if (trail) {
        if constexpr (kRecip) p.mul = sg.tinv;
        else p.dd = (T)(sg.m - sg.n);
    } else if (!(f & (ROW_IDENTITY | ROW_NODIAG))) {
        if (sg.op.fmt == OP_TP) {
            if constexpr (kRecip) p.mul = sg.dinv;
            else p.dd = diag_denom(sg, sg.op.ldu);
        } else {
            p.dd = diag_denom(sg, ldg(&sg.op.ldiag[row]));
        }
    }
    return p;
}
template <typename T, int NR>
__device__ __forceinline__ typename VT<T, NR>::type prow_apply(const PRow<T>& p, bool div, const typename VT<T, NR>::type& v) {
    typename VT<T, NR>::type y;
#pragma unroll
    for (int k = 0; k < NR; ++k) setlane(y, k, p.apply(lane(v, k), div));
    return y;
}

// A shared-stencil list entry: the x row, and in bit 30 whether the row is an
// identity or no-diagonal row (its P^-1 leaves the value as it is).
constexpr int kListRowMask = (1 << 30) - 1;
constexpr int kListPlainBit = 30;

// E row of the tensor-product format: slot (i, j, k) in the reference's
// order, value (wz[k] wy[j]) wx[i] formed exactly as when E was built.  One
// stencil line (4 consecutive x rows) per step, two steps in flight (fully
// unrolled, the 64 NR-wide gathers would take every register).  A row whose
// columns an interior boundary moved runs the same loop with the line's four
// columns read from its exception record.  Both loops are kept free of
// branches: a conditional load inside the loop costs twice its time even
// when never taken.  `stride` = elements between consecutive rows of xb.
template <typename T, typename V, int W, bool EXC>
__device__ __forceinline__ void egather_tp_lines(const OpView<T>& op, int64_t row, const T* xb, int stride, int e,
                                                 V& acc) {
    typedef int I4 __attribute__((ext_vector_type(4)));
    const int64_t D = op.D;
    T wz[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) wz[q] = ldg(&op.tpw[(8 + q) * D + row]);
    const I4* ex = reinterpret_cast<const I4*>(op.ecx) + (size_t)e * 16;
#pragma unroll 2
    for (int l = 0; l < 16; ++l) {
        const T wx = ldg(&op.tpw[(l >> 2) * D + row]);
        const T wy = ldg(&op.tpw[(4 + (l & 3)) * D + row]);
        I4 c;
        if (EXC) {
            c = ldg(ex + l);
        } else {
            const int c0 = ldg(&op.tpc[l * D + row]);
            c = I4{c0, c0 + 1, c0 + 2, c0 + 3};
        }
        const V x0 = vload<T, W>(xb + (size_t)c.x * stride), x1 = vload<T, W>(xb + (size_t)c.y * stride),
                x2 = vload<T, W>(xb + (size_t)c.z * stride), x3 = vload<T, W>(xb + (size_t)c.w * stride);
        acc += ((wz[0] * wy) * wx) * x0;
        acc += ((wz[1] * wy) * wx) * x1;
        acc += ((wz[2] * wy) * wx) * x2;
        acc += ((wz[3] * wy) * wx) * x3;
    }
}

template <typename T, typename V, int W>
__device__ __forceinline__ void egather_tp(const OpView<T>& op, int64_t row, const T* xb, int stride, V& acc) {
    const int e = ldg(&op.exc[row]);
    if (e < 0) {
        egather_tp_lines<T, V, W, false>(op, row, xb, stride, 0, acc);
    } else {
        egather_tp_lines<T, V, W, true>(op, row, xb, stride, e, acc);
        const int xc = ldg(&op.xcol[e]);   // the extra entry after the stencil
        if (xc >= 0) acc += ldg(&op.xval[e]) * vload<T, W>(xb + (size_t)xc * stride);
    }
}

// LDS rows per tile for T x NR values per row (0: the list would not fit in
// kLdsBytes of LDS and every tile reads x directly).
template <typename T, int NR>
struct LdsTile {
    static constexpr int kRows = (kLdsRows * NR * (int)sizeof(T) <= kLdsBytes) ? kLdsRows : 0;
};

// LDS copy of a tile's x rows.  Rows of 32 bytes are kept as two planes of
// 16-byte halves (half h of row p at plane h, slot p): a ds_read_b128 of the
// lanes' rows then meets half the bank conflicts of 32-byte rows (measured on
// the torus-knot system; 8-byte planes were slower again).
template <typename V>
struct LdsRows {
    typedef int I4 __attribute__((ext_vector_type(4)));
    static constexpr bool kSplit = sizeof(V) == 32;
    V* base;
    int cap;
    __device__ __forceinline__ V load(int p) const {
        if constexpr (kSplit) {
            const I4* h = reinterpret_cast<const I4*>(base);
            const I4 a = h[p], b = h[cap + p];
            V v;
            __builtin_memcpy(reinterpret_cast<char*>(&v), &a, 16);
            __builtin_memcpy(reinterpret_cast<char*>(&v) + 16, &b, 16);
            return v;
        } else {
            return base[p];
        }
    }
    __device__ __forceinline__ void store(int p, const V& v) const {
        if constexpr (kSplit) {
            I4* h = reinterpret_cast<I4*>(base);
            I4 a, b;
            __builtin_memcpy(&a, reinterpret_cast<const char*>(&v), 16);
            __builtin_memcpy(&b, reinterpret_cast<const char*>(&v) + 16, 16);
            h[p] = a;
            h[cap + p] = b;
        } else {
            base[p] = v;
        }
    }
};

// A row's tensor-product metadata, loaded before the tile's x is in LDS so
// that its latency overlaps the fill.
template <typename T>
struct TpRow {
    T w[12];      // wx[4], wy[4], wz[4]
    int p[16];    // position of each line's first x row in the tile's list
};

template <typename T>
__device__ __forceinline__ void tp_row_load(const OpView<T>& op, int64_t row, TpRow<T>& m) {
    const int64_t D = op.D;
#pragma unroll
    for (int q = 0; q < 12; ++q) m.w[q] = ldg(&op.tpw[q * D + row]);
#pragma unroll
    for (int l = 0; l < 16; ++l) m.p[l] = ldg(&op.tpl[l * D + row]);
}

// E row of the tensor-product format from the tile's LDS copy of x: the same
// products in the same order as egather_tp_lines, so both paths agree to the
// bit.  lx.load(p) is the x row at position p of the tile's list.
template <typename T, typename V, int W>
__device__ __forceinline__ void egather_tp_lds(const TpRow<T>& m, const LdsRows<V>& lx, V& acc) {
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        T wx = m.w[l >> 2], wy = m.w[4 + (l & 3)];
        // the empty asm (no instruction) holds each line's products to their
        // line: formed all together ahead of the LDS reads, the 64 products
        // take 64 (FP64: 128) registers, which halved the FP64 kernel's
        // occupancy once the row's weights were loaded with the tile's list
        asm volatile("" : "+v"(wx), "+v"(wy));
        const int q = m.p[l];
        const V x0 = lx.load(q), x1 = lx.load(q + 1), x2 = lx.load(q + 2), x3 = lx.load(q + 3);
        acc += ((m.w[8] * wy) * wx) * x0;
        acc += ((m.w[9] * wy) * wx) * x1;
        acc += ((m.w[10] * wy) * wx) * x2;
        acc += ((m.w[11] * wy) * wx) * x3;
    }
}

// An exception row's LDS positions (kExcPos int16 packed in pairs).
struct ExcRow {
    int pk[kExcPos / 2];
    __device__ __forceinline__ int pos(int s) const { return (int)(int16_t)(pk[s >> 1] >> (16 * (s & 1))); }
};

__device__ __forceinline__ void exc_row_load(const int16_t* tlx, int e, ExcRow& m) {
    typedef int I4 __attribute__((ext_vector_type(4)));
    const I4* src = reinterpret_cast<const I4*>(tlx + (size_t)e * kExcPos);
#pragma unroll
    for (int q = 0; q < kExcPos / 8; ++q) {
        const I4 v = ldg(src + q);
        m.pk[4 * q] = v.x;
        m.pk[4 * q + 1] = v.y;
        m.pk[4 * q + 2] = v.z;
        m.pk[4 * q + 3] = v.w;
    }
}

// An exception row (a row an interior boundary moved) from a tile's LDS copy
// of x, with its weights (wx[4], wy[4], wz[4]) and extra value loaded before
// the tile's fill: the slots and the extra entry of egather_tp's exception
// path, in its order.
template <typename T, typename V>
__device__ __forceinline__ void egather_grp_exc(const T (&w)[12], T xv, const ExcRow& m, const LdsRows<V>& lx, V& acc) {
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        const T wx = w[l >> 2], wy = w[4 + (l & 3)];
        const V x0 = lx.load(m.pos(4 * l)), x1 = lx.load(m.pos(4 * l + 1)), x2 = lx.load(m.pos(4 * l + 2)),
                x3 = lx.load(m.pos(4 * l + 3));
        acc += ((w[8] * wy) * wx) * x0;
        acc += ((w[9] * wy) * wx) * x1;
        acc += ((w[10] * wy) * wx) * x2;
        acc += ((w[11] * wy) * wx) * x3;
    }
    // The extra entry's position, taken from its word only here: the empty
    // asm (no instruction) keeps the compiler from decoding it before the
    // tile's barrier, where it would wait for the record's load.
    int w64 = m.pk[32];
    asm volatile("" : "+v"(w64));
    const int xp = (int)(int16_t)w64;
    if (xp >= 0) acc += xv * lx.load(xp);
}

// A kind-0 item of a shared-stencil tile: up to kGrpRows rows with the same
// stencil block (absent rows have weight 0 and are not stored).
template <typename T>
struct GrpItem {
    T w[kGrpRows][12];   // wx[4], wy[4], wz[4] of each row
    int pp[8];           // positions of lines 2h and 2h + 1's first x rows in the tile's list (low, high 16 bits)
    int row[kGrpRows];
    __device__ __forceinline__ int p(int l) const { return (int)(((unsigned)pp[l >> 1] >> (16 * (l & 1))) & 0xffffu); }
};

// The rows and weights of the kGrpRows slots from s0 (slots past the last
// are read as the last: rows g >= n of an item are never used).  Loads with
// no condition, so that none waits on another; a slot's 12 weights are
// contiguous, read 16 bytes at a time.
template <typename T>
__device__ __forceinline__ void grp_item_load(const OpView<T>& op, int s0, GrpItem<T>& m) {
    constexpr int E = 16 / (int)sizeof(T);   // weights per 16-byte load
    typedef T V16 __attribute__((ext_vector_type(E)));
    const int64_t S = op.nslot;
#pragma unroll
    for (int g = 0; g < kGrpRows; ++g) {
        const int64_t sl = s0 + g < S ? s0 + g : S - 1;
        m.row[g] = ldg(&op.grow[sl]);
        const V16* src = reinterpret_cast<const V16*>(op.gw + sl * 12);
#pragma unroll
        for (int q = 0; q < 12 / E; ++q) {
            const V16 v = ldg(src + q);
#pragma unroll
            for (int e = 0; e < E; ++e) m.w[g][q * E + e] = v[e];
        }
    }
}

// The rows of a kind-0 item from the tile's LDS copy of x: every x row of the
// block is read once and used by all rows of the item; each row's products
// and sums are those of egather_tp_lds, in its order, so all paths agree to
// the bit.
template <typename T, typename V, int W>
__device__ __forceinline__ void egather_grp_lds(const GrpItem<T>& m, int n, const LdsRows<V>& lx, V (&acc)[kGrpRows]) {
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        const int q = m.p(l);
        const V x0 = lx.load(q), x1 = lx.load(q + 1), x2 = lx.load(q + 2), x3 = lx.load(q + 3);
#pragma unroll
        for (int g = 0; g < kGrpRows; ++g) {
            if (g < n) {
                const T wx = m.w[g][l >> 2], wy = m.w[g][4 + (l & 3)];
                acc[g] += ((m.w[g][8] * wy) * wx) * x0;
                acc[g] += ((m.w[g][9] * wy) * wx) * x1;
                acc[g] += ((m.w[g][10] * wy) * wx) * x2;
                acc[g] += ((m.w[g][11] * wy) * wx) * x3;
            }
        }
    }
}

// A thread's entries of a tile's list of x rows (entry threadIdx.x + k kBlock).
template <int LR>
struct FillIdx {
    static constexpr int KF = LR > 0 ? (LR + kBlock - 1) / kBlock : 1;
    int src[KF];
    int nu;
};

// The list entries (lst of nu rows; lst[0] readable even when nu = 0), all
// loads in flight together: entries past the list read its last entry
// instead of branching (a load in a branch is waited for inside it).
template <int LR>
__device__ __forceinline__ void lds_fill_index(const int* lst, int nu, FillIdx<LR>& f) {
    f.nu = nu;
    const int last = nu > 0 ? nu - 1 : 0;
#pragma unroll
    for (int k = 0; k < FillIdx<LR>::KF; ++k) {
        const int u = (int)threadIdx.x + k * kBlock;
        f.src[k] = ldg(&lst[u < last ? u : last]);
    }
}

// The listed x rows into LDS, all loads in flight together, each row put
// through xform(row value, list entry) on its way, then a workgroup barrier.
template <typename T, int NR, int LR, typename V, typename F>
__device__ __forceinline__ void lds_fill_rows(const FillIdx<LR>& f, const T* xb, const LdsRows<V>& lrows, F&& xform) {
    constexpr int KF = FillIdx<LR>::KF;
    V buf[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k)
        if ((int)threadIdx.x + k * kBlock < f.nu) buf[k] = vload<T, NR>(xb + (size_t)(f.src[k] & kListRowMask) * NR);
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        const int u = (int)threadIdx.x + k * kBlock;
        if (u < f.nu) lrows.store(u, xform(buf[k], f.src[k]));
    }
    __syncthreads();
}

// One shared-stencil tile of E x: thread t takes the item at place
// tile kBlock + t.
// Tile `tile` of op: ob[row] = (E x)_row from xb (both with NR values per
// row).  FLY: xb holds p (or s) and the tile gathers y = P^-1 p (z = P^-1 s)
// with the preconditioner of segment *pre, formed as the rows go into LDS.
template <typename T, int NR, int LR, bool FLY, typename V>
__device__ __forceinline__ void egather_grp_tile(const OpView<T>& op, const Seg<T>* pre, int tile, const T* xb, T* ob,
                                                 const LdsRows<V>& lrows) {
    // Loads in two rounds, each issued as soon as the previous one's values
    // are in: the list entries (first: the x rows wait on them), the item's
    // offsets and line words, all at fixed places of the tile; then the x
    // rows and the item's data.
    static_assert(FillIdx<LR>::KF * kBlock == LR && LR == kGrpLdsRows, "a tile's list places are read whole");
    FillIdx<LR> f;
    f.nu = ldg(&op.gln[tile]);
#pragma unroll
    for (int k = 0; k < FillIdx<LR>::KF; ++k) f.src[k] = ldg(&op.glr[(size_t)tile * LR + threadIdx.x + k * kBlock]);
    const int64_t NI = op.nitem;
    const int item = tile * kBlock + (int)threadIdx.x;
    const int a = ldg(&op.gofs[item]), b = ldg(&op.gofs[item + 1]);
    GrpItem<T> m;
#pragma unroll
    for (int h = 0; h < 8; ++h) m.pp[h] = ldg(&op.gpl[h * NI + item]);
    const int s0 = a & kGrpOfsMask;
    const int kind = a >> kGrpKindShift;   // 3: an empty place
    const int n = (b & kGrpOfsMask) - s0;
    // The items' global loads are issued before the barrier: a load after it
    // holds its wave (and the workgroup's next tile) for a whole memory
    // latency.  kind 0: the rows' weights, line positions and rows; kind 1
    // (a row an interior boundary moved): its weights in m.w[0], its row,
    // positions and extra value.  kind 2 (identity rows, their rows in the
    // line words) reads its x values after the barrier: such items sit in
    // the last tiles only, and loads of them here reused registers that the
    // kind 1 loads then had to wait for.
    ExcRow mx;
    T xval = (T)0;
    if (kind == 1) {   // issued before the weights: they need only the line words
        const int e = m.pp[0];
        exc_row_load(op.glx, e, mx);
        xval = ldg(&op.xval[e]);
    }
    grp_item_load(op, s0, m);   // kinds 0 and 1 (kind 1: slot s0's row and weights)
    lds_fill_rows<T, NR, LR>(f, xb, lrows, [&](const V& v, int e) {
        if constexpr (FLY) return precond_row<T, NR, true>(*pre, e & kListRowMask, (e >> kListPlainBit) & 1 ? ROW_IDENTITY : 0, v);
        else return v;
    });
    if (kind == 0) {
        V acc[kGrpRows];
#pragma unroll
        for (int g = 0; g < kGrpRows; ++g) acc[g] = (V)(T)0;
        egather_grp_lds<T, V, NR>(m, n, lrows, acc);
#pragma unroll
        for (int g = 0; g < kGrpRows; ++g)
            if (g < n) vstore<T, NR>(ob + (size_t)m.row[g] * NR, acc[g]);
    } else if (kind == 1) {
        V acc = (V)(T)0;
        egather_grp_exc<T, V>(m.w[0], xval, mx, lrows, acc);
        vstore<T, NR>(ob + (size_t)m.row[0] * NR, acc);
    } else if (kind == 2) {
#pragma unroll
        for (int g = 0; g < kGrpRows; ++g)
            if (g < n) {
                const int64_t r = m.pp[g];
                V v = vload<T, NR>(xb + (size_t)r * NR);
                if constexpr (FLY) v = precond_row<T, NR, true>(*pre, r, ldg(&op.flag[r]), v);
                vstore<T, NR>(ob + (size_t)r * NR, v);
            }
    }
}

// E row of the ELL format (eight slots in flight).
template <typename T, typename V, int W>
__device__ __forceinline__ void egather_ell(const OpView<T>& op, int64_t row, const T* xb, int stride, V& acc) {
    const int64_t D = op.D;
    const int* col = op.ecol + row;
    const T* val = op.eval + row;
    const int ew = op.ew;
    int s = 0;
    for (; s + 8 <= ew; s += 8) {
        int c[8];
        T w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) { c[u] = ldg(&col[(int64_t)(s + u) * D]); w[u] = ldg(&val[(int64_t)(s + u) * D]); }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += w[u] * vload<T, W>(xb + (size_t)c[u] * stride);
    }
    for (; s < ew; ++s) acc += ldg(&val[(int64_t)s * D]) * vload<T, W>(xb + (size_t)ldg(&col[(int64_t)s * D]) * stride);
}

// (E x)_row for W consecutive values per row of xb.
template <typename T, int W>
__device__ __forceinline__ typename VT<T, W>::type erow(const OpView<T>& op, int64_t row, const T* xb, int stride) {
    typedef typename VT<T, W>::type V;
    V acc = (V)(T)0;
    if (ldg(&op.flag[row]) & ROW_IDENTITY) return vload<T, W>(xb + (size_t)row * stride);
    if (op.fmt == OP_TP) egather_tp<T, V, W>(op, row, xb, stride, acc);
    else egather_ell<T, V, W>(op, row, xb, stride, acc);
    return acc;
}

// (A x)_row = m x + n (Ld x + Loff tmp), identity rows x, trailing nD rows
// (m - n) x; xv = x_row, tb = tmp of the segment, W values per row.
template <typename T, int W>
__device__ __forceinline__ typename VT<T, W>::type arow(const OpView<T>& op, double m, double n, int64_t row,
                                                        const typename VT<T, W>::type& xv, const T* tb, int stride) {
    typedef typename VT<T, W>::type V;
    const int64_t D = op.D;
    if (row >= D - op.nD) return (T)(m - n) * xv;
    if (ldg(&op.flag[row]) & ROW_IDENTITY) return xv;
    V lo = (V)(T)0;
    T ld;
    if (op.fmt == OP_TP) {
        // a missing neighbour (-1) reads the row's own entry with weight 0:
        // no conditional load in the loop
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int c = ldg(&op.lcol[(int64_t)s * D + row]);
            const T wl = c >= 0 ? op.loff : (T)0;
            lo += wl * vload<T, W>(tb + (size_t)(c >= 0 ? c : row) * stride);
        }
        ld = op.ldu;
    } else {
        for (int s = 0; s < op.lw; ++s) {
            const int c = ldg(&op.lcol[(int64_t)s * D + row]);
            lo += ldg(&op.lval[(int64_t)s * D + row]) * vload<T, W>(tb + (size_t)c * stride);
        }
        ld = ldg(&op.ldiag[row]);
    }
    return (T)m * xv + (T)n * (ld * xv + lo);
}

// arow of the tensor-product format with every load issued up front: the
// row's flag and L columns with x (and the caller's other loads), then the
// six neighbours' E x rows, then the products and sums of arow in its order,
// the trailing and identity rows taking their values by selection (a
// missing neighbour reads the row's own entry with weight 0).  pre: x goes
// through P^-1 first (x is p or s of an iteration that forms y and z where
// it reads them).
template <typename T, int W>
struct TpLRow {
    typename VT<T, W>::type xv;
    int c[6];
    uint8_t f;
};
template <typename T, int W>
__device__ __forceinline__ void arow_tp_load(const OpView<T>& op, int64_t row, const T* xr, TpLRow<T, W>& a) {
    const int64_t D = op.D;
    a.f = ldg(&op.flag[row]);
#pragma unroll
    for (int s = 0; s < 6; ++s) a.c[s] = ldg(&op.lcol[(int64_t)s * D + row]);
    a.xv = vload<T, W>(xr);
}
template <typename T, int W>
__device__ __forceinline__ typename VT<T, W>::type arow_tp(const Seg<T>& sg, int64_t row, TpLRow<T, W>& a, const T* tb,
                                                           int stride, bool pre) {
    typedef typename VT<T, W>::type V;
    const OpView<T>& op = sg.op;
    V nb[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) nb[s] = vload<T, W>(tb + (size_t)(a.c[s] >= 0 ? a.c[s] : row) * stride);
    V xv = a.xv;
    if (pre) xv = prow_apply<T, W>(prow_of(sg, row, a.f), prow_div(sg), xv);
    V lo = (V)(T)0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const T wl = a.c[s] >= 0 ? op.loff : (T)0;
        lo += wl * nb[s];
    }
    const T ld = op.ldu;
    V res = (T)sg.m * xv + (T)sg.n * (ld * xv + lo);
    if (a.f & ROW_IDENTITY) res = xv;
    if (row >= op.D - op.nD) res = (T)(sg.m - sg.n) * xv;
    return res;
}

// ---------------------------------------------------------------- vector kernels
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_init(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, const T* b, T* x, T* r,
                                                 T* r0, T* p, T* v) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    double pv[1][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) pv[0][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto bv = vload<T, NR>(b + i);
        typename VT<T, NR>::type zero = (typename VT<T, NR>::type)(T)0;
        vstore<T, NR>(r + i, bv);
        vstore<T, NR>(r0 + i, bv);
        vstore<T, NR>(x + i, zero);
        vstore<T, NR>(p + i, zero);
        vstore<T, NR>(v + i, zero);
#pragma unroll
        for (int k = 0; k < NR; ++k) pv[0][k] = (double)lane(bv, k) * (double)lane(bv, k);
    }
    block_partials<1, NR>(pv, sg.part, lb, sg.nblocks);
}

// RUN lanes: p = (p - w v) beta + r ; y = P^-1 p (fly: y is formed where it is read).
// UPDATE lanes (mixed): x64 += x32 / scale ; x32 = 0.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_pupdate(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, T* p, const T* v,
                                                    const T* r, T* y, T* x, MixVec mv, int fly) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (row >= sg.op.D || !seg_any(L, M_STAGE1)) return;
    const size_t i = (size_t)(sg.off + row * NR);
    if (seg_any(L, M_RUN)) {
        const uint8_t f = ldg(&sg.op.flag[row]);
        auto pv = vload<T, NR>(p + i);
        auto vv = vload<T, NR>(v + i);
        auto rv = vload<T, NR>(r + i);
        const unsigned run = L.of(M_RUN);
#pragma unroll
        for (int k = 0; k < NR; ++k) {   // a group owns kMaxNR states: every lane's load is valid
            const T beta = (T)ldg(&sg.st[k].beta), w = (T)ldg(&sg.st[k].omega);
            const T pk = (-w * lane(vv, k) + lane(pv, k)) * beta + lane(rv, k);   // p.Copy(-w,v,p); p*=beta; p+=r
            if ((run >> k) & 1u) setlane(pv, k, pk);
        }
        vstore<T, NR>(p + i, pv);
        if (!fly) vstore<T, NR>(y + i, prow_apply<T, NR>(prow_of(sg, row, f), prow_div(sg), pv));
    }
    if (std::is_same<T, float>::value && seg_any(L, M_UPDATE)) {
        auto xv = vload<T, NR>(x + i);
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (lane_in(L, k, M_UPDATE)) {
                double* xd = mv.x64 + lm_at(sg, k, row);
                stg(xd, ldg(xd) + (double)lane(xv, k) / ldg(&sg.st[k].scale));
                setlane(xv, k, (T)0);
            }
        vstore<T, NR>(x + i, xv);
    }
}

// tmp = E x.  Rows are independent; a thread owns one row and all NR rhs.
// UPDATE lanes (mixed, first gather of the iteration): tmp64 = E64 x64.
// E x for the instance segments.  Tiles with an LDS list copy their x rows
// into LDS once and gather every stencil line from there; the other tiles
// (lists longer than the LDS, or operators without tiles) read x directly.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_egather(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, const T* x, T* out,
                                                    unsigned lanes, MixVec mv) {
    typedef typename VT<T, NR>::type V;
    constexpr int LR = LdsTile<T, NR>::kRows;
    __shared__ V lx[LR > 0 ? LR : 1];
    const LdsRows<V> lrows{lx, LR > 0 ? LR : 1};
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg, lanes != 0);
    if (lanes && !seg_any(L, lanes)) return;               // uniform per workgroup
    const bool main = !lanes || seg_any(L, lanes & ~M_UPDATE);
    const size_t i = (size_t)(sg.off + row * NR);
    const T* xb = x + sg.off;
    const int nu = (LR > 0 && main && sg.op.fmt == OP_TP && sg.op.tln) ? ldg(&sg.op.tln[lb]) : -1;
    if (nu >= 0 && nu <= LR) {
        // Loads in two rounds, each issued as soon as the previous one's
        // values are in: the tile's list entries with the row's flag,
        // exception record and weights (rows past the end read the last
        // row's, unused), all unconditional; then the listed x rows with the
        // row's LDS positions (a plain row's 16 line positions, or a moved
        // row's exception record and extra value, in one register array) or
        // an identity row's own x.  Then the LDS copy, the barrier and every
        // stencil line from LDS.
        FillIdx<LR> f;
        lds_fill_index<LR>(sg.op.tlr + ldg(&sg.op.tlo[lb]), nu, f);
        const bool act = row < sg.op.D;
        const int64_t rr = act ? row : sg.op.D - 1;
        const int64_t D = sg.op.D;
        const int e = ldg(&sg.op.exc[rr]);
        const uint8_t fl = ldg(&sg.op.flag[rr]);
        TpRow<T> m;
#pragma unroll
        for (int q = 0; q < 12; ++q) m.w[q] = ldg(&sg.op.tpw[q * D + rr]);
        const int kind = act ? ((fl & ROW_IDENTITY) ? 2 : (e < 0 ? 0 : 1)) : 3;
        ExcRow mx;   // kind 0: words 0..15 are the line positions; kind 1: the exception record
        T xval = (T)0;
        V xi = (V)(T)0;
        if (kind == 0) {
#pragma unroll
            for (int l = 0; l < 16; ++l) mx.pk[l] = ldg(&sg.op.tpl[l * D + row]);
        } else if (kind == 1) {   // rows an interior boundary moved
            exc_row_load(sg.op.tlx, e, mx);
            xval = ldg(&sg.op.xval[e]);
        } else if (kind == 2) {
            xi = vload<T, NR>(xb + (size_t)row * NR);
        }
        lds_fill_rows<T, NR, LR>(f, xb, lrows, [](const V& v, int) { return v; });
        if (act) {
            V acc = (V)(T)0;
#pragma unroll
            for (int l = 0; l < 16; ++l) m.p[l] = mx.pk[l];
            if (kind == 2) acc = xi;
            else if (kind == 0) egather_tp_lds<T, V, NR>(m, lrows, acc);
            else egather_grp_exc<T, V>(m.w, xval, mx, lrows, acc);
            vstore<T, NR>(out + i, acc);
        }
    } else if (main && row < sg.op.D) {
        vstore<T, NR>(out + i, erow<T, NR>(sg.op, row, xb, NR));
    }
}

// FP64 gathers of the refinement lanes of FP32 segments: tmp64 = E64 x64 for
// every lane in an UPDATE status (a launch of its own beside the FP32
// gather, which then carries only FP32 work).  A lane's FP64 vectors are a
// run of D values, the one-right-hand-side layout, so block lb gathers tile
// lb of the operator's shared-stencil tiles for each such lane in turn;
// operators without tiles read x64 directly, a row per thread.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_egather_upd(const Seg<float>* __restrict__ segs, const int* __restrict__ bseg,
                                                        MixVec mv) {
    constexpr int LR = GrpTile<double, 1>::kRows;
    static_assert(LR > 0, "FP64 shared-stencil tiles fit the LDS");
    __shared__ double lx[LR];
    const LdsRows<double> lrows{lx, LR};
    int lb;
    int64_t row;
    const Seg<float>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (!seg_any(L, M_UPDATE)) return;   // uniform per workgroup
    if (sg.op64.ngt > 0) {
        if (lb >= sg.op64.ngt) return;
        bool first = true;
        for (int k = 0; k < sg.nr; ++k)
            if (lane_in(L, k, M_UPDATE)) {
                if (!first) __syncthreads();   // the previous lane's LDS reads are done
                first = false;
                egather_grp_tile<double, 1, LR, false>(sg.op64, (const Seg<double>*)nullptr, lb, mv.x64 + lm_at(sg, k, 0),
                                                       mv.tmp64 + lm_at(sg, k, 0), lrows);
            }
        return;
    }
    if (row >= sg.op.D) return;
    for (int k = 0; k < sg.nr; ++k)
        if (lane_in(L, k, M_UPDATE))
            stg(&mv.tmp64[lm_at(sg, k, row)], erow<double, 1>(sg.op64, row, mv.x64 + lm_at(sg, k, 0), 1));
}

// E x for launches whose every segment has shared-stencil tiles: block lb of
// a segment takes its tile lb (blocks past the last tile have nothing to
// do).  40 KB of LDS and at most 128 VGPRs, so four
// workgroups share a CU (a direct-read path for other segments, even out of
// line, made the FP32 x 8 kernel spill).
// FLY: x is p (or s) and the launch gathers y = P^-1 p (z = P^-1 s).
template <typename T, int NR, int WPE, bool FLY>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_egather_grp(
    const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, const T* x, T* out, unsigned lanes, MixVec mv) {
    typedef typename VT<T, NR>::type V;
    constexpr int LR = GrpTile<T, NR>::kRows;
    static_assert(LR > 0, "shared-stencil tiles do not fit this value size");
    __shared__ V lx[LR];
    const LdsRows<V> lrows{lx, LR};
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg, lanes != 0);
    if (lanes && !seg_any(L, lanes)) return;               // uniform per workgroup
    const bool main = !lanes || seg_any(L, lanes & ~M_UPDATE);
    if (main && lb < sg.op.ngt) egather_grp_tile<T, NR, LR, FLY>(sg.op, &sg, lb, x + sg.off, out + sg.off, lrows);
}

// out = A x from tmp = E x.  RED = 0: no reduction; 1: partial (other, out);
// 2: partials (out, out), (out, other).
// Mixed, RED = 1: UPDATE lanes r64 = b - A64 x64, partial (r64, r64).
// Mixed, RED = 2: RESCALE lanes partial (r32, r32) (rv32 = r32).
template <typename T, int NR, int RED>
__global__ __launch_bounds__(kBlock) void k_lfinish(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, const T* x,
                                                    const T* tmp, T* out, const T* other, const T* rv32,
                                                    unsigned lanes, MixVec mv, int fly) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg, lanes != 0);
    if (lanes && !seg_any(L, lanes)) return;   // uniform per segment
    constexpr int NV = RED == 2 ? 2 : 1;
    double pv[NV][NR];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) pv[a][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        if (!lanes || seg_any(L, lanes & ~(M_UPDATE | M_RESCALE))) {
            typename VT<T, NR>::type res, ov;
            if (sg.op.fmt == OP_TP) {   // every load that waits on no other in one round
                TpLRow<T, NR> a;
                arow_tp_load<T, NR>(sg.op, row, x + i, a);
                if (RED > 0) ov = vload<T, NR>(other + i);
                res = arow_tp<T, NR>(sg, row, a, tmp + sg.off, NR, fly);   // fly: x is p or s
            } else {
                auto xv = vload<T, NR>(x + i);
                if (RED > 0) ov = vload<T, NR>(other + i);
                if (fly) xv = precond_row<T, NR>(sg, row, ldg(&sg.op.flag[row]), xv);
                res = arow<T, NR>(sg.op, sg.m, sg.n, row, xv, tmp + sg.off, NR);
            }
            vstore<T, NR>(out + i, res);
            if (RED > 0) {
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const double rk = (double)lane(res, k);
                    if (RED == 1) {
                        pv[0][k] = (double)lane(ov, k) * rk;
                    } else {
                        pv[0][k] = rk * rk;
                        pv[NV - 1][k] = rk * (double)lane(ov, k);
                    }
                }
            }
        }
        if (std::is_same<T, float>::value && RED == 1 && seg_any(L, M_UPDATE)) {
#pragma unroll
            for (int k = 0; k < NR; ++k)   // constant indices keep pv in registers
                if (k < sg.nr && lane_in(L, k, M_UPDATE)) {
                    const size_t j = lm_at(sg, k, row);
                    const double ax =
                        arow<double, 1>(sg.op64, sg.m, sg.n, row, ldg(&mv.x64[j]), mv.tmp64 + lm_at(sg, k, 0), 1);
                    const double rk = ldg(&mv.b64[j]) - ax;
                    stg(&mv.r64[j], rk);
                    pv[0][k] = rk * rk;
                }
        }
        if (std::is_same<T, float>::value && RED == 2 && seg_any(L, M_RESCALE)) {
            auto rr = vload<T, NR>(rv32 + i);
#pragma unroll
            for (int k = 0; k < NR; ++k)
                if (lane_in(L, k, M_RESCALE)) {
                    const double rk = (double)lane(rr, k);
                    pv[0][k] = rk * rk;
                    pv[NV - 1][k] = 0.0;
                }
        }
    }
    if (RED > 0) block_partials<NV, NR>(pv, sg.part, lb, sg.nblocks);
}

// RUN lanes: s = r - alpha v ; z = P^-1 s (fly: z is formed where it is read).
// RESCALE lanes (mixed): r32 = r64 scale ; r0 = r32 ; p = v = 0.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_supdate(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, T* r, T* v, T* s, T* z,
                                                    T* r0, T* p, MixVec mv, int fly) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (row >= sg.op.D || !seg_any(L, M_STAGE2)) return;
    const size_t i = (size_t)(sg.off + row * NR);
    const uint8_t f = ldg(&sg.op.flag[row]);
    auto rv = vload<T, NR>(r + i);
    auto vv = vload<T, NR>(v + i);
    if (seg_any(L, M_RUN)) {
        typename VT<T, NR>::type sv;
#pragma unroll
        for (int k = 0; k < NR; ++k) {   // a group owns kMaxNR states: every lane's load is valid
            const T a = (T)ldg(&sg.st[k].alpha);
            setlane(sv, k, -a * lane(vv, k) + lane(rv, k));   // s.Copy(-alpha, v, r)
        }
        vstore<T, NR>(s + i, sv);
        if (!fly) vstore<T, NR>(z + i, prow_apply<T, NR>(prow_of(sg, row, f), prow_div(sg), sv));
    }
    if (std::is_same<T, float>::value && seg_any(L, M_RESCALE)) {
        auto r0v = vload<T, NR>(r0 + i);
        auto pv = vload<T, NR>(p + i);
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (lane_in(L, k, M_RESCALE)) {
                const T rk = (T)(ldg(&mv.r64[lm_at(sg, k, row)]) * ldg(&sg.st[k].scale));
                setlane(rv, k, rk);
                setlane(r0v, k, rk);
                setlane(pv, k, (T)0);
                setlane(vv, k, (T)0);
            }
        vstore<T, NR>(r + i, rv);
        vstore<T, NR>(r0 + i, r0v);
        vstore<T, NR>(p + i, pv);
        vstore<T, NR>(v + i, vv);
    }
}

// Jacobi sweep `sweep` (>= 1) of RUN lanes: y += omega D^-1 (x - A y) on the
// rows below dofs, with A y already in ay (GeneralizedBiCGStabSystem.h:366-378).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_jacobi(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg,
                                                   const T* x, const T* ay, T* y, int sweep) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (sg.precond != 2 || sweep >= sg.precond_iters) return;
    if (row >= sg.op.D - sg.op.nD || !seg_any(L, M_RUN)) return;
    const size_t i = (size_t)(sg.off + row * NR);
    const uint8_t f = ldg(&sg.op.flag[row]);
    auto xv = vload<T, NR>(x + i);
    auto av = vload<T, NR>(ay + i);
    auto yv = vload<T, NR>(y + i);
#pragma unroll
    for (int k = 0; k < NR; ++k)
        if (lane_in(L, k, M_RUN))
            setlane(yv, k, lane(yv, k) + (T)kJacobiOmega * diag_inverse(sg, row, f, (T)(lane(xv, k) - lane(av, k))));
    vstore<T, NR>(y + i, yv);
}

template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_xrupdate(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, T* x, const T* y,
                                                     const T* z, T* r, const T* s, const T* t, const T* r0, int fly) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (!seg_any(L, M_RUN)) return;
    double pv[2][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) { pv[0][k] = 0.0; pv[1][k] = 0.0; }
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto xv = vload<T, NR>(x + i);
        auto yv = vload<T, NR>(y + i);   // fly: p
        auto sv = vload<T, NR>(s + i);
        auto tv = vload<T, NR>(t + i);
        auto r0v = vload<T, NR>(r0 + i);
        const uint8_t f = ldg(&sg.op.flag[row]);
        // r is rewritten on every RUN lane: read only to keep the others
        const unsigned run = L.of(M_RUN);
        const bool all_run = run == (1u << NR) - 1u;
        typename VT<T, NR>::type rv = all_run ? (typename VT<T, NR>::type)(T)0 : vload<T, NR>(r + i);
        typename VT<T, NR>::type zv;
        if (fly) {   // y = P^-1 p, z = P^-1 s
            const PRow<T> pr = prow_of(sg, row, f);
            const bool div = prow_div(sg);
            yv = prow_apply<T, NR>(pr, div, yv);
            zv = prow_apply<T, NR>(pr, div, sv);
        } else {
            zv = vload<T, NR>(z + i);
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {   // a group owns kMaxNR states: every lane's load is valid
            if (!((run >> k) & 1u)) continue;
            const T a = (T)ldg(&sg.st[k].alpha), w = (T)ldg(&sg.st[k].omega);
            T xk = a * lane(yv, k) + lane(xv, k);   // x.Copy(alpha, y, x)
            xk = w * lane(zv, k) + xk;               // x.Copy(w, z, x)
            T rk = -w * lane(tv, k) + lane(sv, k);   // r.Copy(-w, t, s)
            setlane(xv, k, xk);
            setlane(rv, k, rk);
            pv[0][k] = (double)rk * (double)rk;
            pv[1][k] = (double)lane(r0v, k) * (double)rk;
        }
        vstore<T, NR>(x + i, xv);
        vstore<T, NR>(r + i, rv);
    }
    block_partials<2, NR>(pv, sg.part, lb, sg.nblocks);
}

// FP64 restart, BiCGStab.h:62-71: r = b - A x (A x already in ax), r0 = r; partial (r, r).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_restart(const Seg<T>* __restrict__ segs, const int* __restrict__ bseg, const T* b,
                                                    const T* ax, T* r, T* r0) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (!seg_any(L, M_RESTART)) return;
    double pv[1][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) pv[0][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto bv = vload<T, NR>(b + i);
        auto av = vload<T, NR>(ax + i);
        auto rv = vload<T, NR>(r + i);
        auto r0v = vload<T, NR>(r0 + i);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!lane_in(L, k, M_RESTART)) continue;
            T rk = lane(bv, k) - lane(av, k);
            setlane(rv, k, rk);
            setlane(r0v, k, rk);
            pv[0][k] = (double)rk * (double)rk;
        }
        vstore<T, NR>(r + i, rv);
        vstore<T, NR>(r0 + i, r0v);
    }
    block_partials<1, NR>(pv, sg.part, lb, sg.nblocks);
}

// ---------------------------------------------------------------- mixed precision start
// x64 = 0, r64 = b ; partial (b, b)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_mx_start(const Seg<double>* __restrict__ segs, const int* __restrict__ bseg, const double* b,
                                                     double* x, double* r) {
    int lb;
    int64_t row;
    const Seg<double>& sg = seg_of(segs, bseg, lb, row);
    double pv[1][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) pv[0][k] = 0.0;
    if (row < sg.op.D) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const size_t j = lm_at(sg, k, row);
            const double bk = b[j];
            r[j] = bk;
            x[j] = 0.0;
            pv[0][k] = bk * bk;
        }
    }
    block_partials<1, NR>(pv, sg.part, lb, sg.nblocks);
}

// FP32 start of the lanes in ST_RESCALE: r32 = r64 * scale, r0 = r32,
// p = v = x32 = 0 ; partial (r32, r32).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_mx_down(const Seg<double>* __restrict__ segs, const int* __restrict__ bseg, const double* r64,
                                                    float* r, float* r0, float* p, float* v, float* x) {
    int lb;
    int64_t row;
    const Seg<double>& sg = seg_of(segs, bseg, lb, row);
    const auto L = lane_stat<NR>(sg);
    if (!seg_any(L, M_RESCALE)) return;
    double pv[1][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) pv[0][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto ro = vload<float, NR>(r + i);
        auto r0o = vload<float, NR>(r0 + i);
        auto po = vload<float, NR>(p + i);
        auto vo = vload<float, NR>(v + i);
        auto xo = vload<float, NR>(x + i);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (!lane_in(L, k, M_RESCALE)) continue;
            const float rk = (float)(r64[lm_at(sg, k, row)] * ldg(&sg.st[k].scale));
            setlane(ro, k, rk);
            setlane(r0o, k, rk);
            setlane(po, k, 0.0f);
            setlane(vo, k, 0.0f);
            setlane(xo, k, 0.0f);
            pv[0][k] = (double)rk * (double)rk;
        }
        vstore<float, NR>(r + i, ro);
        vstore<float, NR>(r0 + i, r0o);
        vstore<float, NR>(p + i, po);
        vstore<float, NR>(v + i, vo);
        vstore<float, NR>(x + i, xo);
    }
    block_partials<1, NR>(pv, sg.part, lb, sg.nblocks);
}

// ---------------------------------------------------------------- finalisers
enum FinKind : int {
    FIN_INIT = 0, FIN_ALPHA, FIN_OMEGA, FIN_ITER64, FIN_RESTART64, FIN_MX_START, FIN_MX_INIT, FIN_ITER_MX
};

// Lanes each finaliser acts on (the producing kernel used the same mask) and
// the number of partial vectors it reads.
template <int KIND> struct FinTraits;
template <> struct FinTraits<FIN_INIT> { static constexpr int NV = 1; static constexpr unsigned M = M_ALL; };
template <> struct FinTraits<FIN_ALPHA> { static constexpr int NV = 1; static constexpr unsigned M = M_STAGE1; };
template <> struct FinTraits<FIN_OMEGA> { static constexpr int NV = 2; static constexpr unsigned M = M_STAGE2; };
template <> struct FinTraits<FIN_ITER64> { static constexpr int NV = 2; static constexpr unsigned M = M_STAGE3; };
template <> struct FinTraits<FIN_RESTART64> { static constexpr int NV = 1; static constexpr unsigned M = M_RESTART; };
template <> struct FinTraits<FIN_MX_START> { static constexpr int NV = 1; static constexpr unsigned M = M_ALL; };
template <> struct FinTraits<FIN_MX_INIT> { static constexpr int NV = 1; static constexpr unsigned M = M_RESCALE; };
template <> struct FinTraits<FIN_ITER_MX> { static constexpr int NV = 2; static constexpr unsigned M = M_STAGE3; };

constexpr double kEps = 2.220446049250313e-16;   // NumTraits<double>::epsilon (BiCGStab.h:56)

// FP32 recurrence (re)started from the scaled true residual.
// The scalar part of the recurrence for one (instance, rhs).
template <int KIND>
__device__ void fin_scalar(KState& s, const double* tot) {
    if (KIND == FIN_INIT) {   // BiCGStab.h:41-58 with x0 = 0: r = r0 = b
        const double bsq = tot[0];
        s.bsq = s.r0sq = s.rr = bsq;
        s.rho_old = s.alpha = s.omega = 1.0;
        s.iters = s.restarts = 0;
        if (!std::isfinite(bsq)) { s.status = ST_BREAKDOWN; return; }
        if (bsq == 0.0) { s.status = ST_ZERO_RHS; return; }
        if (!(bsq > s.tol2 * bsq)) { s.status = ST_CONVERGED; return; }   // tol >= 1
        if (s.max_it <= 0) { s.status = ST_MAXIT; return; }
        s.tol2 *= bsq;                 // tol2 held tol^2
        s.rho = bsq;                   // (r0, r) with r0 = r = b
        s.beta = s.rho;                // (rho / 1) * (1 / 1)
        s.status = ST_RUN;
    } else if (KIND == FIN_ALPHA) {
        if (s.status != ST_RUN) { mx_update(s, tot[0]); return; }
        s.alpha = s.rho / tot[0];                                      // BiCGStab.h:79
        if (!std::isfinite(s.alpha)) s.status = ST_BREAKDOWN;
    } else if (KIND == FIN_OMEGA) {
        if (s.status != ST_RUN) { mx_init(s, tot[0], ST_READY); return; }
        s.omega = tot[0] > 0.0 ? tot[1] / tot[0] : 0.0;                // BiCGStab.h:86-94
        if (!std::isfinite(s.omega)) s.status = ST_BREAKDOWN;
    } else if (KIND == FIN_ITER64) {
        ++s.iters;
        s.rr = tot[0];
        if (!(s.rr > s.tol2)) { s.status = std::isnan(s.rr) ? ST_BREAKDOWN : ST_CONVERGED; return; }   // BiCGStab.h:59
        if (s.iters >= s.max_it) { s.status = ST_MAXIT; return; }
        s.rho_old = s.rho;
        s.rho = tot[1];                                                // BiCGStab.h:61
        if (fabs(s.rho) < kEps * kEps * s.r0sq) { s.status = ST_RESTART; return; }   // BiCGStab.h:62
        s.beta = (s.rho / s.rho_old) * (s.alpha / s.omega);           // BiCGStab.h:73
        if (!std::isfinite(s.beta)) s.status = ST_BREAKDOWN;
    } else if (KIND == FIN_RESTART64) {                                // BiCGStab.h:62-71
        s.rho = s.r0sq = tot[0];
        if (s.restarts++ == 0) s.iters = 0;
        s.beta = (s.rho / s.rho_old) * (s.alpha / s.omega);
        s.status = std::isfinite(s.beta) ? ST_RUN : ST_BREAKDOWN;
    } else if (KIND == FIN_MX_START) {
        const double bsq = tot[0];
        s.bsq = s.rr_true = bsq;
        s.iters = s.restarts = s.updates = 0;
        if (!std::isfinite(bsq)) { s.status = ST_BREAKDOWN; return; }
        if (bsq == 0.0) { s.status = ST_ZERO_RHS; return; }
        if (!(bsq > s.tol2 * bsq)) { s.status = ST_CONVERGED; return; }
        if (s.max_it <= 0) { s.status = ST_MAXIT; return; }
        s.tol2 *= bsq;
        s.scale = 1.0 / sqrt(bsq);
        s.status = ST_RESCALE;
    } else if (KIND == FIN_MX_INIT) {
        mx_init(s, tot[0], ST_RUN);
    } else if (KIND == FIN_ITER_MX) {
        if (s.status == ST_READY) { s.status = ST_RUN; return; }
        ++s.iters;
        const double rr = tot[0];
        s.rr = rr;
        if (std::isnan(rr)) { s.status = ST_BREAKDOWN; return; }
        // scaled tolerance: ||r|| <= tol ||b|| with r scaled by `scale`
        s.judged = 0;
        if (!(rr > s.tol2 * s.scale * s.scale)) { s.judged = 2; s.status = ST_UPDATE; return; }
        if (s.iters >= s.max_it) { s.status = ST_UPDATE_MAXIT; return; }
        if (rr < s.delta2 * s.maxrr) { s.judged = 1; s.status = ST_UPDATE; return; }
        s.rho_old = s.rho;
        s.rho = tot[1];
        if (fabs(s.rho) < kEps * kEps * s.r0sq) { ++s.restarts; s.status = ST_UPDATE; return; }   // BiCGStab.h:62
        s.beta = (s.rho / s.rho_old) * (s.alpha / s.omega);
        if (!std::isfinite(s.beta)) { s.status = ST_BREAKDOWN; return; }
        if (rr > s.maxrr) s.maxrr = rr;
    }
}

// One workgroup per segment: fixed-order sum of the producing kernel's
// block partials, then the scalar recurrence for each active lane.
template <typename T, int NR, int KIND>
__global__ __launch_bounds__(kBlock) void k_fin(const Seg<T>* __restrict__ segs, int nseg) {
    constexpr int NV = FinTraits<KIND>::NV;
    constexpr unsigned M = FinTraits<KIND>::M;
    // One workgroup per (segment, rhs lane): the lane's NV components over all
    // block partials, U blocks per thread in flight at once (the partials were
    // just written by another kernel, so every round costs a full memory
    // latency), then a workgroup sum and the lane's scalar update.
    const Seg<T>& sg = segs[blockIdx.x / NR];
    const int k = (int)(blockIdx.x % NR);
    if (k >= sg.nr || !lane_in(sg, k, M)) return;   // same test the producer used
    constexpr int U = 8;
    __shared__ double red[NV][kBlock / 64];
    const int lanei = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) acc[a] = 0.0;
    for (int b0 = threadIdx.x; b0 < sg.nblocks; b0 += kBlock * U) {
        double v[NV][U];
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + kBlock * u;
                v[a][u] = b < sg.nblocks ? ldg(&sg.part[(size_t)(a * NR + k) * sg.nblocks + b]) : 0.0;
            }
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int u = 0; u < U; ++u) acc[a] += v[a][u];
    }
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        const double t = wave_sum(acc[a]);
        if (lanei == 0) red[a][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot[NV];
    for (int a = 0; a < NV; ++a) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[a][w];
        tot[a] = t;
    }
    fin_scalar<KIND>(sg.st[k], tot);
}

__global__ void k_reset_states(KState* st, int n, double tol2, int max_it, double delta2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    KState z = {};
    z.tol2 = tol2;
    z.max_it = max_it;
    z.delta2 = delta2;
    z.status = ST_RUN;
    st[i] = z;
}

void launch_reset_states(KState* st, int n, double tol2, int max_it, double delta2, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_states, dim3((n + 255) / 256), dim3(256), 0, s, st, n, tol2, max_it, delta2);
}

// ---------------------------------------------------------------- launchers
#define LAUNCH(kern, ...) \
    hipLaunchKernelGGL(kern, dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.nseg == 1 ? nullptr : c.bseg, __VA_ARGS__)
// The E gathers run the round-robin block map.
#define LAUNCH_GATHER(kern, ...)                                                                         \
    hipLaunchKernelGGL(kern, dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs,                        \
                       c.nseg == 1 ? nullptr : c.bseg + c.nblocks, __VA_ARGS__)
// E gather of launch c; FL: form P^-1 of the gathered vector (only with
// shared-stencil tiles, grp_launch).
#define EGATHER(FL, ...)                                                     \
    do {                                                                     \
        if constexpr (GrpTile<T, NR>::kRows > 0) {                           \
            if (c.grp > 0 && grp_gather(sizeof(T), NR, c.grp)) {             \
                if (FL) LAUNCH_GATHER((k_egather_grp<T, NR, GrpWaves<T, NR>::kWaves, true>), __VA_ARGS__); \
                else LAUNCH_GATHER((k_egather_grp<T, NR, GrpWaves<T, NR>::kWaves, false>), __VA_ARGS__); \
                break;                                                       \
            }                                                                \
        }                                                                    \
        LAUNCH_GATHER((k_egather<T, NR>), __VA_ARGS__);                      \
    } while (0)
#define FIN(T, NR, KIND) hipLaunchKernelGGL((k_fin<T, NR, KIND>), dim3(c.nseg * NR), dim3(kBlock), 0, c.stream, c.segs, c.nseg)

// Waves per SIMD the shared-stencil gather is compiled for: four (128 VGPRs)
// where that does not spill, two (256 VGPRs) for FP32 x 4 and the FP64 shapes.
template <typename T, int NR>
struct GrpWaves {
    static constexpr int kWaves = (sizeof(T) == 4 && NR != 4) ? 4 : CPMGPU_GROUP_WAVES_WIDE;
};

// Value shapes whose E gathers use shared-stencil tiles when every operator
// of the launch has at least ntiles of them: by default FP32 x 8, x 2 and
// FP64 x 4 from kGrpMinTiles tiles, FP32 x 1 from 768, where they were
// measured faster than row tiles (tubes of 42 k to 406 k rows).
// CPMGPU_GROUP_GATHER=all uses them for every shape whose tile fits the LDS,
// =none for none.
// The override is read from the environment once per solve, tile build or
// query (grp_refresh_override, called where a launch sequence starts), not
// per launch: a solve makes thousands of launches.
static std::atomic<int> g_grp_override{-2};   // -2 unread, -1 none, 0 "none", 1 "all"
// y = P^-1 p and z = P^-1 s formed where they are read (launches that allow
// it, Launch::iteration); CPMGPU_PRECOND_FLY=0 stores them instead.
static std::atomic<int> g_fly_override{-1};   // -1 none, 0 off
static bool precond_fly() { return g_fly_override.load(std::memory_order_relaxed) != 0; }

void grp_refresh_override() {
    const char* e = std::getenv("CPMGPU_GROUP_GATHER");
    const std::string v = e ? e : "";
    g_grp_override.store(v == "all" ? 1 : v == "none" ? 0 : -1, std::memory_order_relaxed);
    const char* fe = std::getenv("CPMGPU_PRECOND_FLY");
    g_fly_override.store(fe && std::string(fe) == "0" ? 0 : -1, std::memory_order_relaxed);
}

bool grp_gather(int tsize, int nr, int ntiles) {
    (void)tsize;
    if (g_grp_override.load(std::memory_order_relaxed) == -2) grp_refresh_override();
    const int ov = g_grp_override.load(std::memory_order_relaxed);
    if (ov >= 0) return ov == 1;
    // every shape whose rows fit the tiles' LDS is faster through them where the
    // operator has them (torus-knot system, one / eight groups per launch: FP64 x
    // 1 20.0 / 87.0 -> 16.9 / 69.8 us, FP64 x 2 23.6 / 99.8 -> 19.6 / 90.9 us,
    // FP32 x 4 21.3 / 92.1 -> 17.7 / 72.2 us); one right-hand side: row tiles
    // were faster up to ~600 tiles
    return ntiles >= (nr == 1 ? 768 : kGrpMinTiles);
}

bool grp_build(int tsize, int nr) {
    if (g_grp_override.load(std::memory_order_relaxed) == -2) grp_refresh_override();
    const int ov = g_grp_override.load(std::memory_order_relaxed);
    if (ov >= 0) return ov == 1;
    // the shapes of the solves that repeat on one operator: FP32 x 1, 2, 8 (FP32+refine
    // batches and streams) and FP64 x 4 (FP64 batches); an operator served at FP32 x 4
    // is solved once and would not win back the build (0.14 s at 1.27 M rows)
    return (tsize == 4 && nr != 4) || (tsize == 8 && nr == 4);
}

// Whether the E gathers of launch c use shared-stencil tiles (EGATHER's choice).
template <typename T, int NR>
static bool grp_launch(const LaunchCfg<T>& c) {
    if constexpr (GrpTile<T, NR>::kRows > 0) return c.grp > 0 && grp_gather(sizeof(T), NR, c.grp);
    return false;
}

template <typename T, int NR>
void Launch<T, NR>::init(const LaunchCfg<T>& c, const T* b, T* x, T* r, T* r0, T* p, T* v) {
    LAUNCH((k_init<T, NR>), b, x, r, r0, p, v);
    FIN(T, NR, FIN_INIT);
}
template <typename T, int NR>
void Launch<T, NR>::iteration(const LaunchCfg<T>& c, Vecs<T>& w, bool mixed, const MixVec& mv) {
    // fly: y = P^-1 p and z = P^-1 s are formed where they are read (the
    // gathers' LDS fills, lfinish, xrupdate) rather than stored: launches
    // whose gathers fill LDS from shared-stencil tiles and whose
    // preconditioner is pointwise (no Jacobi sweeps after the first)
    const bool fly = grp_launch<T, NR>(c) && c.jacobi_sweeps <= 1 && precond_fly();
    const int fl = fly ? 1 : 0;
    const T* yp = fly ? w.p : w.y;   // the vector the first gather and lfinish read
    const T* zs = fly ? w.s : w.z;   // ... the second
    LAUNCH((k_pupdate<T, NR>), w.p, w.v, w.r, w.y, w.x, mv, fl);
    for (int sw = 1; sw < c.jacobi_sweeps; ++sw) {   // y = P^-1 p, Jacobi sweeps after the first
        apply(c, w.y, w.tmp, w.t, M_RUN);
        LAUNCH((k_jacobi<T, NR>), w.p, w.t, w.y, sw);
    }
    if constexpr (std::is_same<T, float>::value) {
        if (mixed) {   // FP32 lanes in the gather, refinement lanes' FP64 gathers in a launch of their own
            EGATHER(fly, yp, w.tmp, M_RUN, mv);
            LAUNCH((k_egather_upd<NR>), mv);   // few groups refine at once: their blocks together
        } else {
            EGATHER(fly, yp, w.tmp, M_STAGE1, mv);
        }
    } else {
        EGATHER(fly, yp, w.tmp, M_STAGE1, mv);
    }
    LAUNCH((k_lfinish<T, NR, 1>), yp, w.tmp, w.v, w.r0, (const T*)nullptr, M_STAGE1, mv, fl);
    FIN(T, NR, FIN_ALPHA);
    LAUNCH((k_supdate<T, NR>), w.r, w.v, w.s, w.z, w.r0, w.p, mv, fl);
    for (int sw = 1; sw < c.jacobi_sweeps; ++sw) {   // z = P^-1 s
        apply(c, w.z, w.tmp, w.t, M_RUN);
        LAUNCH((k_jacobi<T, NR>), w.s, w.t, w.z, sw);
    }
    EGATHER(fly, zs, w.tmp, M_RUN, mv);
    LAUNCH((k_lfinish<T, NR, 2>), zs, w.tmp, w.t, w.s, (const T*)w.r, M_STAGE2, mv, fl);
    FIN(T, NR, FIN_OMEGA);
    LAUNCH((k_xrupdate<T, NR>), w.x, yp, w.z, w.r, w.s, w.t, w.r0, fl);
    if (mixed) FIN(T, NR, FIN_ITER_MX);
    else FIN(T, NR, FIN_ITER64);
}
template <typename T, int NR>
void Launch<T, NR>::restart(const LaunchCfg<T>& c, Vecs<T>& w) {
    const MixVec none{};
    EGATHER(false, w.x, w.tmp, M_RESTART, none);
    LAUNCH((k_lfinish<T, NR, 0>), w.x, w.tmp, w.t, (const T*)nullptr, (const T*)nullptr, M_RESTART, none, 0);
    LAUNCH((k_restart<T, NR>), w.b, w.t, w.r, w.r0);
    FIN(T, NR, FIN_RESTART64);
}
template <typename T, int NR>
void Launch<T, NR>::apply(const LaunchCfg<T>& c, const T* x, T* tmp, T* y, unsigned lanes) {
    const MixVec none{};
    EGATHER(false, x, tmp, lanes, none);
    LAUNCH((k_lfinish<T, NR, 0>), x, tmp, y, (const T*)nullptr, (const T*)nullptr, lanes, none, 0);
}

template <typename T>
void launch_lfinish_plain(const LaunchCfg<T>& c, const T* x, const T* tmp, T* out) {
    const MixVec none{};
    LAUNCH((k_lfinish<T, 1, 0>), x, tmp, out, (const T*)nullptr, (const T*)nullptr, 0u, none, 0);
}
template void launch_lfinish_plain<double>(const LaunchCfg<double>&, const double*, const double*, double*);

template <typename T, int NR>
void launch_egather(const LaunchCfg<T>& c, const T* x, T* out, bool fly) {
    const MixVec none{};
    const bool fl = fly && grp_launch<T, NR>(c) && precond_fly();
    EGATHER(fl, x, out, 0u, none);
}

template <int NR>
void MixedLaunch<NR>::start(const LaunchCfg<double>& c, const double* b, double* x64, double* r64) {
    LAUNCH((k_mx_start<NR>), b, x64, r64);
    FIN(double, NR, FIN_MX_START);
}
template <int NR>
void MixedLaunch<NR>::down(const LaunchCfg<double>& c, const double* r64, Vecs<float>& w) {
    LAUNCH((k_mx_down<NR>), r64, w.r, w.r0, w.p, w.v, w.x);
    FIN(double, NR, FIN_MX_INIT);
}

#define CPM_INST(T, NR) template struct Launch<T, NR>; template void launch_egather<T, NR>(const LaunchCfg<T>&, const T*, T*, bool);
CPM_INST(double, 1) CPM_INST(double, 2) CPM_INST(double, 4) CPM_INST(double, 8)
CPM_INST(float, 1) CPM_INST(float, 2) CPM_INST(float, 4) CPM_INST(float, 8)
template struct MixedLaunch<1>;
template struct MixedLaunch<2>;
template struct MixedLaunch<4>;
template struct MixedLaunch<8>;

}  // namespace cpmgpu
