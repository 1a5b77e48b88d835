// Batched BiCGStab for the closest-point screened-Poisson operator.
//
// Restates the reference's right-preconditioned BiCGStab (src/BiCGStab.h:30-108)
// and its matrix-free operator (src/GeneralizedBiCGStabSystem.h:191-261) as
// seven bandwidth-bound kernels per iteration:
//
//   pupdate   p = (p - w v) beta + r ; y = P^-1 p                 (:268-271)
//   egather   tmp = E y          (identity rows copy)              (:235-261)
//   lfinish   v = m y + n (Ld y + Loff tmp) ; partial (r0, v)       (:93-115,:166-189)
//   supdate   s = r - alpha v ; z = P^-1 s                         (:275-277)
//   egather   tmp = E z
//   lfinish   t = m z + n (Ld z + Loff tmp) ; partials (t,t),(t,s)  (:278-288)
//   xrupdate  x += alpha y ; x += w z ; r = s - w t ; (r,r),(r0,r)  (:290-292, loop test :253, rho :255)
//
// Every scalar of the recurrence (rho, alpha, omega, beta, the loop test and
// the rho-restart test) is computed on the device by the last workgroup of the
// reducing kernel ("finaliser"), so iterations run without host round trips.
// One launch covers every instance of a batch: workgroups map to (instance,
// row-block) through a segment table; converged instances are skipped.
// Right-hand sides of one operator are interleaved [row][rhs] so one read of
// E's ELL slot serves all of them.
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.h"
#include "krylov.h"

namespace cpmgpu {

// ---------------------------------------------------------------- helpers
template <typename T, int NR> struct VT { typedef T __attribute__((ext_vector_type(NR))) type; };
template <typename T> struct VT<T, 1> { typedef T type; };

template <typename T, int NR>
__device__ __forceinline__ typename VT<T, NR>::type vload(const T* p) {
    return *reinterpret_cast<const typename VT<T, NR>::type*>(p);
}
template <typename T, int NR>
__device__ __forceinline__ void vstore(T* p, const typename VT<T, NR>::type& v) {
    *reinterpret_cast<typename VT<T, NR>::type*>(p) = v;
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

// Reduce NV*NR per-thread doubles over the workgroup, publish the block's
// partials and let the last-arriving workgroup of the segment sum all of them
// in a fixed order.  Returns true in every thread of the finaliser block, with
// the totals in tot[] (valid in all threads).
template <typename T, int NR, int NV>
__device__ bool reduce_arrive(double (&v)[NV][NR], const Seg<T>& sg, int lb, double (&tot)[NV][NR]) {
    __shared__ double red[kBlock / 64][NV * NR];
    __shared__ int is_last;
    const int lanei = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            double w = wave_sum(v[a][k]);
            if (lanei == 0) red[wave][a * NR + k] = w;
        }
    __syncthreads();
    if (threadIdx.x < NV * NR) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
        sg.part[(size_t)lb * NV * NR + threadIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned prev = atomicAdd(sg.cnt, 1u);
        is_last = (prev == (unsigned)sg.nblocks - 1u);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    // fixed-order sum of the partials: thread j takes blocks j, j+256, ...
    double acc[NV][NR];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) acc[a][k] = 0.0;
    for (int b = threadIdx.x; b < sg.nblocks; b += kBlock)
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int k = 0; k < NR; ++k)
                acc[a][k] += __hip_atomic_load(sg.part + (size_t)b * NV * NR + a * NR + k, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            double w = wave_sum(acc[a][k]);
            if (lanei == 0) red[wave][a * NR + k] = w;
        }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) s += red[w][a * NR + k];
            tot[a][k] = s;
        }
    if (threadIdx.x == 0) atomicExch(sg.cnt, 0u);
    return true;
}

template <typename T>
__device__ __forceinline__ const Seg<T>& seg_of(const Seg<T>* segs, const int* bseg, int& lb, int64_t& row) {
    const Seg<T>& sg = segs[bseg[blockIdx.x]];
    lb = (int)blockIdx.x - sg.block0;
    row = (int64_t)lb * kBlock + threadIdx.x;
    return sg;
}

template <typename T, int NR>
__device__ __forceinline__ bool any_running(const Seg<T>& sg) {
    bool run = false;
    for (int k = 0; k < sg.nr; ++k) run |= (sg.st[k].status == ST_RUN);
    return run;
}

// Diagonal preconditioner, GeneralizedBiCGStabSystem.h:303-338 (the branch the
// reference runs: Compute() is never called).
template <typename T>
__device__ __forceinline__ T precond_apply(const Seg<T>& sg, int64_t row, uint8_t f, T val) {
    if (sg.precond != 1 /*diagonal*/) return val;
    const int64_t dof = sg.op.D - sg.op.nD;
    if (row >= dof) return val / (T)(sg.m - sg.n);   // :335-338 overwrite the trailing rows last
    if (f & ROW_IDENTITY) return val;
    if (f & ROW_NODIAG) return val;
    return val / ((T)sg.m + sg.op.ldiag[row] * (T)sg.n);
}

// ---------------------------------------------------------------- kernels
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_init(const Seg<T>* segs, const int* bseg, const T* b, T* x, T* r,
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
    double tot[1][NR];
    if (reduce_arrive<T, NR, 1>(pv, sg, lb, tot) && threadIdx.x == 0) {
        for (int k = 0; k < sg.nr; ++k) {
            KState& s = sg.st[k];
            const double bsq = tot[0][k];
            s.bsq = bsq;
            s.r0sq = bsq;
            s.rr = bsq;
            s.rho_old = 1.0;
            s.alpha = 1.0;
            s.omega = 1.0;
            s.iters = 0;
            s.restarts = 0;
            if (std::isnan(bsq) || std::isinf(bsq)) { s.status = ST_BREAKDOWN; continue; }
            if (bsq == 0.0) { s.status = ST_ZERO_RHS; continue; }
            if (!(bsq > s.tol2 * bsq)) { s.status = ST_CONVERGED; continue; }  // tol >= 1
            if (s.max_it <= 0) { s.status = ST_MAXIT; continue; }
            s.tol2 = s.tol2 * bsq;   // tol2 held tol^2 until now
            s.rho = bsq;             // (r0, r) with r0 = r = b
            s.beta = s.rho;          // (rho / 1) * (1 / 1)
            s.status = ST_RUN;
        }
    }
}

template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_pupdate(const Seg<T>* segs, const int* bseg, T* p, const T* v,
                                                    const T* r, T* y) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    if (row >= sg.op.D || !any_running<T, NR>(sg)) return;
    const size_t i = (size_t)(sg.off + row * NR);
    const uint8_t f = sg.op.flag[row];
    auto pv = vload<T, NR>(p + i);
    auto vv = vload<T, NR>(v + i);
    auto rv = vload<T, NR>(r + i);
    typename VT<T, NR>::type yv;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const T beta = (T)sg.st[k < sg.nr ? k : 0].beta, w = (T)sg.st[k < sg.nr ? k : 0].omega;
        T pk = (-w * lane(vv, k) + lane(pv, k)) * beta + lane(rv, k);   // p.Copy(-w,v,p); p*=beta; p+=r
        setlane(pv, k, pk);
        setlane(yv, k, precond_apply(sg, row, f, pk));
    }
    vstore<T, NR>(p + i, pv);
    vstore<T, NR>(y + i, yv);
}

// tmp = E x.  Rows are independent; a thread owns one row and all NR rhs.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_egather(const Seg<T>* segs, const int* bseg, const T* x, T* out,
                                                    int skip_done) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    if (row >= sg.op.D) return;
    if (skip_done && !any_running<T, NR>(sg)) return;
    const size_t i = (size_t)(sg.off + row * NR);
    typename VT<T, NR>::type acc = (typename VT<T, NR>::type)(T)0;
    if (sg.op.flag[row] & ROW_IDENTITY) {
        acc = vload<T, NR>(x + i);
    } else {
        const int64_t D = sg.op.D;
        const int* col = sg.op.ecol + row;
        const T* val = sg.op.eval + row;
        const T* xb = x + sg.off;
        const int ew = sg.op.ew;
        int s = 0;
        for (; s + 8 <= ew; s += 8) {
            int c[8];
            T w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) { c[u] = col[(int64_t)(s + u) * D]; w[u] = val[(int64_t)(s + u) * D]; }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                auto xv = vload<T, NR>(xb + (size_t)c[u] * NR);
                acc += w[u] * xv;
            }
        }
        for (; s < ew; ++s) {
            const int c = col[(int64_t)s * D];
            acc += val[(int64_t)s * D] * vload<T, NR>(xb + (size_t)c * NR);
        }
    }
    vstore<T, NR>(out + i, acc);
}

// out = m x + n (Ld x + Loff tmp); identity rows out = x; trailing nD rows
// (m - n) x.  RED = 0: no reduction; 1: (r0, out) -> alpha; 2: (out,out),
// (out, s) -> omega.  Returns early for finished instances only when RED > 0.
template <typename T, int NR, int RED>
__global__ __launch_bounds__(kBlock) void k_lfinish(const Seg<T>* segs, const int* bseg, const T* x,
                                                    const T* tmp, T* out, const T* other) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    if (RED > 0 && !any_running<T, NR>(sg)) return;   // uniform per segment: every block of it leaves
    constexpr int NV = RED == 2 ? 2 : 1;
    double pv[NV][NR];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int k = 0; k < NR; ++k) pv[a][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        const int64_t D = sg.op.D;
        const uint8_t f = sg.op.flag[row];
        auto xv = vload<T, NR>(x + i);
        typename VT<T, NR>::type res;
        if (row >= D - sg.op.nD) {
            res = (T)(sg.m - sg.n) * xv;
        } else if (f & ROW_IDENTITY) {
            res = xv;
        } else {
            typename VT<T, NR>::type lo = (typename VT<T, NR>::type)(T)0;
            const T* tb = tmp + sg.off;
            for (int s = 0; s < sg.op.lw; ++s) {
                const int c = sg.op.lcol[(int64_t)s * D + row];
                lo += sg.op.lval[(int64_t)s * D + row] * vload<T, NR>(tb + (size_t)c * NR);
            }
            res = (T)sg.m * xv + (T)sg.n * (sg.op.ldiag[row] * xv + lo);
        }
        vstore<T, NR>(out + i, res);
        if (RED > 0) {
            auto ov = vload<T, NR>(other + i);
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (RED == 1) {
                    pv[0][k] = (double)(lane(ov, k) * lane(res, k));
                } else {
                    pv[0][k] = (double)(lane(res, k) * lane(res, k));
                    if (NV == 2) pv[NV - 1][k] = (double)(lane(res, k) * lane(ov, k));
                }
            }
        }
    }
    if (RED == 0) return;
    double tot[NV][NR];
    if (reduce_arrive<T, NR, NV>(pv, sg, lb, tot) && threadIdx.x == 0) {
        for (int k = 0; k < sg.nr; ++k) {
            KState& s = sg.st[k];
            if (s.status != ST_RUN) continue;
            if (RED == 1) {
                s.alpha = s.rho / tot[0][k];                       // BiCGStab.h:273
            } else {
                const double val = tot[0][k];
                s.omega = val > 0.0 ? tot[NV - 1][k] / val : 0.0;  // :280-288
            }
            if (!std::isfinite(RED == 1 ? s.alpha : s.omega)) s.status = ST_BREAKDOWN;
        }
    }
}

template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_supdate(const Seg<T>* segs, const int* bseg, const T* r,
                                                    const T* v, T* s, T* z) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    if (row >= sg.op.D || !any_running<T, NR>(sg)) return;
    const size_t i = (size_t)(sg.off + row * NR);
    const uint8_t f = sg.op.flag[row];
    auto rv = vload<T, NR>(r + i);
    auto vv = vload<T, NR>(v + i);
    typename VT<T, NR>::type sv, zv;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const T a = (T)sg.st[k < sg.nr ? k : 0].alpha;
        T sk = -a * lane(vv, k) + lane(rv, k);   // s.Copy(-alpha, v, r)
        setlane(sv, k, sk);
        setlane(zv, k, precond_apply(sg, row, f, sk));
    }
    vstore<T, NR>(s + i, sv);
    vstore<T, NR>(z + i, zv);
}

template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_xrupdate(const Seg<T>* segs, const int* bseg, T* x, const T* y,
                                                     const T* z, T* r, const T* s, const T* t, const T* r0) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    if (!any_running<T, NR>(sg)) return;
    double pv[2][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) { pv[0][k] = 0.0; pv[1][k] = 0.0; }
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto xv = vload<T, NR>(x + i);
        auto yv = vload<T, NR>(y + i);
        auto zv = vload<T, NR>(z + i);
        auto sv = vload<T, NR>(s + i);
        auto tv = vload<T, NR>(t + i);
        auto r0v = vload<T, NR>(r0 + i);
        typename VT<T, NR>::type rv;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const KState& st = sg.st[k < sg.nr ? k : 0];
            const T a = (T)st.alpha, w = (T)st.omega;
            T xk = a * lane(yv, k) + lane(xv, k);   // x.Copy(alpha, y, x)
            xk = w * lane(zv, k) + xk;               // x.Copy(w, z, x)
            T rk = -w * lane(tv, k) + lane(sv, k);   // r.Copy(-w, t, s)
            if (st.status != ST_RUN) { xk = lane(xv, k); rk = (T)0; }
            setlane(xv, k, xk);
            setlane(rv, k, rk);
            pv[0][k] = (double)(rk * rk);
            pv[1][k] = (double)(lane(r0v, k) * rk);
        }
        vstore<T, NR>(x + i, xv);
        vstore<T, NR>(r + i, rv);
    }
    double tot[2][NR];
    if (reduce_arrive<T, NR, 2>(pv, sg, lb, tot) && threadIdx.x == 0) {
        for (int k = 0; k < sg.nr; ++k) {
            KState& st = sg.st[k];
            if (st.status != ST_RUN) continue;
            ++st.iters;
            const double rr = tot[0][k];
            st.rr = rr;
            // loop test of the next iteration, BiCGStab.h:253
            if (!(rr > st.tol2)) { st.status = std::isnan(rr) ? ST_BREAKDOWN : ST_CONVERGED; continue; }
            if (st.iters >= st.max_it) { st.status = ST_MAXIT; continue; }
            st.rho_old = st.rho;
            st.rho = tot[1][k];                                  // :255
            const double eps = 2.220446049250313e-16;
            if (std::fabs(st.rho) < eps * eps * st.r0sq) {       // :256
                st.status = ST_RESTART;
                continue;
            }
            st.beta = (st.rho / st.rho_old) * (st.alpha / st.omega);   // :267
            if (!std::isfinite(st.beta)) st.status = ST_BREAKDOWN;
        }
    }
}

// Restart, BiCGStab.h:257-264: r = b - A x (A x already in ax), r0 = r,
// rho = r0sq = (r, r).
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_restart(const Seg<T>* segs, const int* bseg, const T* b,
                                                    const T* ax, T* r, T* r0) {
    int lb;
    int64_t row;
    const Seg<T>& sg = seg_of(segs, bseg, lb, row);
    bool any = false;
    for (int k = 0; k < sg.nr; ++k) any |= (sg.st[k].status == ST_RESTART);
    if (!any) return;
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
            if (k < sg.nr && sg.st[k].status == ST_RESTART) {
                T rk = lane(bv, k) - lane(av, k);
                setlane(rv, k, rk);
                setlane(r0v, k, rk);
                pv[0][k] = (double)(rk * rk);
            }
        }
        vstore<T, NR>(r + i, rv);
        vstore<T, NR>(r0 + i, r0v);
    }
    double tot[1][NR];
    if (reduce_arrive<T, NR, 1>(pv, sg, lb, tot) && threadIdx.x == 0) {
        for (int k = 0; k < sg.nr; ++k) {
            KState& st = sg.st[k];
            if (st.status != ST_RESTART) continue;
            st.rho = st.r0sq = tot[0][k];
            if (st.restarts++ == 0) st.iters = 0;
            st.beta = (st.rho / st.rho_old) * (st.alpha / st.omega);
            st.status = std::isfinite(st.beta) ? ST_RUN : ST_BREAKDOWN;
        }
    }
}

// ---------------------------------------------------------------- refinement
// FP64 iterative refinement around an FP32 BiCGStab (rows interleave NR rhs).
//   start:   x64 = 0, r64 = b64, (b,b)
//   step:    r32 = r64 / |r64|  ->  FP32 solve A d = r32  ->  x64 += |r64| d
//            r64 = b64 - A64 x64 ; (r64, r64)
template <int NR>
__global__ __launch_bounds__(kBlock) void k_ref_start(const Seg<double>* segs, const int* bseg, const double* b,
                                                      double* x, double* r) {
    int lb;
    int64_t row;
    const Seg<double>& sg = seg_of(segs, bseg, lb, row);
    double pv[1][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) pv[0][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto bv = vload<double, NR>(b + i);
        vstore<double, NR>(r + i, bv);
        vstore<double, NR>(x + i, (typename VT<double, NR>::type)0.0);
#pragma unroll
        for (int k = 0; k < NR; ++k) pv[0][k] = lane(bv, k) * lane(bv, k);
    }
    double tot[1][NR];
    if (reduce_arrive<double, NR, 1>(pv, sg, lb, tot) && threadIdx.x == 0) {
        for (int k = 0; k < sg.nr; ++k) {
            KState& st = sg.st[k];
            st.bsq = tot[0][k];
            st.rr = st.bsq;
            st.refinements = 0;
            st.total_iters = 0;
            st.scale = std::sqrt(st.bsq);
            if (!std::isfinite(st.bsq)) st.outer_status = ST_BREAKDOWN;
            else if (st.bsq == 0.0) st.outer_status = ST_ZERO_RHS;
            else if (!(st.bsq > st.tol2 * st.bsq)) st.outer_status = ST_CONVERGED;
            else { st.tol2 = st.tol2 * st.bsq; st.outer_status = ST_RUN; }
        }
    }
}

// b32 = r64 / scale for running instances (zero otherwise so the inner solve
// of a finished instance sees a zero rhs and stops at once).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_ref_down(const Seg<double>* segs, const int* bseg, const double* r,
                                                     float* b32) {
    int lb;
    int64_t row;
    const Seg<double>& sg = seg_of(segs, bseg, lb, row);
    if (row >= sg.op.D) return;
    const size_t i = (size_t)(sg.off + row * NR);
    auto rv = vload<double, NR>(r + i);
    typename VT<float, NR>::type o;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const KState& st = sg.st[k < sg.nr ? k : 0];
        const bool run = k < sg.nr && st.outer_status == ST_RUN;
        setlane(o, k, run ? (float)(lane(rv, k) / st.scale) : 0.0f);
    }
    vstore<float, NR>(b32 + i, o);
}

template <int NR>
__global__ __launch_bounds__(kBlock) void k_ref_up(const Seg<double>* segs, const int* bseg, const float* d32,
                                                   double* x) {
    int lb;
    int64_t row;
    const Seg<double>& sg = seg_of(segs, bseg, lb, row);
    if (row >= sg.op.D) return;
    const size_t i = (size_t)(sg.off + row * NR);
    auto dv = vload<float, NR>(d32 + i);
    auto xv = vload<double, NR>(x + i);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const KState& st = sg.st[k < sg.nr ? k : 0];
        if (k < sg.nr && st.outer_status == ST_RUN) setlane(xv, k, lane(xv, k) + st.scale * (double)lane(dv, k));
    }
    vstore<double, NR>(x + i, xv);
}

// r64 = b - ax ; (r64, r64) ; outer loop test.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_ref_resid(const Seg<double>* segs, const int* bseg, const double* b,
                                                      const double* ax, double* r, const KState* inner) {
    int lb;
    int64_t row;
    const Seg<double>& sg = seg_of(segs, bseg, lb, row);
    bool any = false;
    for (int k = 0; k < sg.nr; ++k) any |= (sg.st[k].outer_status == ST_RUN);
    if (!any) return;
    double pv[1][NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) pv[0][k] = 0.0;
    if (row < sg.op.D) {
        const size_t i = (size_t)(sg.off + row * NR);
        auto bv = vload<double, NR>(b + i);
        auto av = vload<double, NR>(ax + i);
        auto rv = vload<double, NR>(r + i);
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (k < sg.nr && sg.st[k].outer_status == ST_RUN) {
                const double rk = lane(bv, k) - lane(av, k);
                setlane(rv, k, rk);
                pv[0][k] = rk * rk;
            }
        vstore<double, NR>(r + i, rv);
    }
    double tot[1][NR];
    if (reduce_arrive<double, NR, 1>(pv, sg, lb, tot) && threadIdx.x == 0) {
        const int seg_index = bseg[blockIdx.x];
        for (int k = 0; k < sg.nr; ++k) {
            KState& st = sg.st[k];
            if (st.outer_status != ST_RUN) continue;
            const KState& in = inner[seg_index * kMaxNR + k];
            st.total_iters += in.iters;
            st.restarts += in.restarts;
            ++st.refinements;
            st.rr = tot[0][k];
            st.scale = std::sqrt(st.rr);
            if (!(st.rr > st.tol2)) st.outer_status = std::isnan(st.rr) ? ST_BREAKDOWN : ST_CONVERGED;
            else if (in.status == ST_BREAKDOWN) st.outer_status = ST_BREAKDOWN;
            else if (st.refinements >= st.max_it) st.outer_status = ST_MAXIT;
        }
    }
}

__global__ void k_reset_states(KState* st, int n, double tol2, int max_it) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    KState z = {};
    z.tol2 = tol2;
    z.max_it = max_it;
    z.status = ST_RUN;
    z.outer_status = ST_RUN;
    st[i] = z;
}

void launch_reset_states(KState* st, int n, double tol2, int max_it, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_states, dim3((n + 255) / 256), dim3(256), 0, s, st, n, tol2, max_it);
}

// ---------------------------------------------------------------- launchers
template <typename T, int NR>
void Launch<T, NR>::init(const LaunchCfg<T>& c, const T* b, T* x, T* r, T* r0, T* p, T* v) {
    hipLaunchKernelGGL((k_init<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, b, x, r, r0, p, v);
}
template <typename T, int NR>
void Launch<T, NR>::iteration(const LaunchCfg<T>& c, Vecs<T>& w) {
    hipLaunchKernelGGL((k_pupdate<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.p, w.v, w.r, w.y);
    hipLaunchKernelGGL((k_egather<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.y, w.tmp, 1);
    hipLaunchKernelGGL((k_lfinish<T, NR, 1>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.y, w.tmp, w.v, w.r0);
    hipLaunchKernelGGL((k_supdate<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.r, w.v, w.s, w.z);
    hipLaunchKernelGGL((k_egather<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.z, w.tmp, 1);
    hipLaunchKernelGGL((k_lfinish<T, NR, 2>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.z, w.tmp, w.t, w.s);
    hipLaunchKernelGGL((k_xrupdate<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.x, w.y, w.z, w.r, w.s, w.t, w.r0);
}
template <typename T, int NR>
void Launch<T, NR>::restart(const LaunchCfg<T>& c, Vecs<T>& w) {
    hipLaunchKernelGGL((k_egather<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.x, w.tmp, 0);
    hipLaunchKernelGGL((k_lfinish<T, NR, 0>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.x, w.tmp, w.t, (const T*)nullptr);
    hipLaunchKernelGGL((k_restart<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, w.b, w.t, w.r, w.r0);
}
template <typename T, int NR>
void Launch<T, NR>::apply(const LaunchCfg<T>& c, const T* x, T* tmp, T* y) {
    hipLaunchKernelGGL((k_egather<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, x, tmp, 0);
    hipLaunchKernelGGL((k_lfinish<T, NR, 0>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, x, tmp, y, (const T*)nullptr);
}

template <int NR>
void RefLaunch<NR>::start(const LaunchCfg<double>& c, const double* b, double* x, double* r) {
    hipLaunchKernelGGL((k_ref_start<NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, b, x, r);
}
template <int NR>
void RefLaunch<NR>::down(const LaunchCfg<double>& c, const double* r, float* b32) {
    hipLaunchKernelGGL((k_ref_down<NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, r, b32);
}
template <int NR>
void RefLaunch<NR>::up(const LaunchCfg<double>& c, const float* d32, double* x) {
    hipLaunchKernelGGL((k_ref_up<NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, d32, x);
}
template <int NR>
void RefLaunch<NR>::resid(const LaunchCfg<double>& c, const double* b, const double* ax, double* r,
                          const KState* inner) {
    hipLaunchKernelGGL((k_ref_resid<NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, b, ax, r, inner);
}

template <typename T, int NR>
void launch_egather(const LaunchCfg<T>& c, const T* x, T* out) {
    hipLaunchKernelGGL((k_egather<T, NR>), dim3(c.nblocks), dim3(kBlock), 0, c.stream, c.segs, c.bseg, x, out, 0);
}

#define CPM_INST(T, NR) template struct Launch<T, NR>; template void launch_egather<T, NR>(const LaunchCfg<T>&, const T*, T*);
CPM_INST(double, 1) CPM_INST(double, 2) CPM_INST(double, 4) CPM_INST(double, 8)
CPM_INST(float, 1) CPM_INST(float, 2) CPM_INST(float, 4) CPM_INST(float, 8)
template struct RefLaunch<1>;
template struct RefLaunch<2>;
template struct RefLaunch<4>;
template struct RefLaunch<8>;

}  // namespace cpmgpu
