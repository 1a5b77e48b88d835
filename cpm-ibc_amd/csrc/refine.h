// Krylov state of one (instance, right-hand side) and the refinement-step
// rules of the mixed-precision solve.  Plain C++ apart from the function
// qualifiers, so that the state machine also compiles for the host (a CPU
// test defines __host__ and __device__ away: tests/cpp/refine_test.cpp).
#pragma once
#include <cmath>

namespace cpmgpu {

// Per (instance, rhs) Krylov state.  Written only by the one-block finalise
// kernels (krylov.hip) and read uniformly by every block of the next kernels.
enum : int {
    ST_RUN = 0, ST_CONVERGED = 1, ST_MAXIT = 2, ST_RESTART = 3, ST_ZERO_RHS = 4, ST_BREAKDOWN = 5,
    // mixed precision: the FP32 recurrence asks for an FP64 refinement step
    // (UPDATE*), the true residual is rescaled into the FP32 vectors
    // (RESCALE), and the lane rejoins at the next iteration (READY)
    ST_UPDATE = 6, ST_UPDATE_MAXIT = 7, ST_RESCALE = 8, ST_READY = 9,
    // mixed precision: refinement stopped reducing the true residual; the
    // host re-solves the system in FP64 from zero
    ST_FALLBACK = 10
};

struct KState {
    double rho, rho_old, alpha, omega, beta;
    double bsq, r0sq, tol2, rr;
    double rr_true;      // mixed: ||b - A x64||^2 at the last refinement step
    double scale;        // mixed: 1 / ||b - A x64|| (FP32 vectors are scaled by it)
    double maxrr;        // mixed: largest scaled recurrence residual^2 since the last step
    double delta2;       // mixed: refine when rr < delta2 * maxrr
    int iters, restarts, status, max_it;
    int updates;
    int fails;           // mixed: refinement steps that reduced the true residual by less than 10%
    int judged;          // mixed: what asked for the pending step: 1 the delta test (its reduction is
                         // judged), 2 the scaled tolerance test, 0 the rho restart rule or the iteration limit
    int stalls;          // mixed: restart-rule steps that did not reduce the true residual
};

__host__ __device__ inline void mx_init(KState& s, double rr, int next) {
    s.rr = s.rho = s.r0sq = s.maxrr = rr;
    s.rho_old = s.alpha = s.omega = 1.0;
    s.beta = s.rho;
    s.status = (rr > 0.0 && std::isfinite(rr)) ? next : ST_BREAKDOWN;
}

// Refinement step: the true residual ||b - A64 x64||^2 is rt.  A lane that
// took a new right-hand side (bsq < 0, x64 = 0) learns ||b||^2 = rt here, as
// FIN_MX_START does for the first ones.
//
// The FP32 recurrence can only resolve the correction to about eps32 times
// the condition number; on fine grids (h = 0.01, condition ~1e5) asking it
// for delta = 1e-2 per step leaves corrections that do not reduce the true
// residual.  For steps asked for by the delta test (the recurrence residual
// fell by delta), a true reduction that falls short of the recurrence's by
// more than ~3x raises delta 10x (to at most 0.5); such a step that does
// not reduce the true residual, or three that reduce it by less than 10%,
// hand the lane to an FP64 re-solve from zero (ST_FALLBACK).  So does a step
// asked for by the scaled tolerance test (the FP32 recurrence claims the
// whole solve converged) that does not reduce the true residual: the lane
// would otherwise rescale and stall short of the tolerance until max_iters.
// A step of the rho restart rule may legitimately raise the true residual
// (a right-hand side that lives on identity rows is exact on them after one
// iteration, which leaves r orthogonal to r0 and a larger residual on the
// other rows), so only its third such step falls back.
__host__ __device__ inline void mx_update(KState& s, double rt) {
    ++s.updates;
    const double prev = s.rr_true;
    const bool fresh = s.bsq < 0.0;
    s.rr_true = rt;
    if (fresh) {
        s.bsq = rt;
        if (!std::isfinite(rt)) { s.status = ST_BREAKDOWN; return; }
        if (rt == 0.0) { s.status = ST_ZERO_RHS; return; }
        if (!(rt > s.tol2 * rt)) { s.status = ST_CONVERGED; return; }
        if (s.max_it <= 0) { s.status = ST_MAXIT; return; }
        s.tol2 *= rt;
    }
    if (!std::isfinite(rt)) { s.status = ST_FALLBACK; return; }
    if (!(rt > s.tol2)) { s.status = ST_CONVERGED; return; }
    if (s.status == ST_UPDATE_MAXIT) { s.status = ST_MAXIT; return; }
    if (!fresh) {
        const double a = rt / prev;          // squared true reduction of this step
        if (s.judged == 1) {
            const double c = s.rr / s.r0sq;  // squared reduction the recurrence reached
            if (!(a < 1.0)) { s.status = ST_FALLBACK; return; }
            if (a > 10.0 * c) s.delta2 = fmin(0.25, 100.0 * s.delta2);
            if (a > 0.81 && ++s.fails >= 3) { s.status = ST_FALLBACK; return; }
        } else if (!(a < 1.0)) {
            if (s.judged == 2 || ++s.stalls >= 3) { s.status = ST_FALLBACK; return; }
        }
    }
    s.scale = 1.0 / sqrt(rt);
    s.status = ST_RESCALE;
}

}  // namespace cpmgpu
