// CPU test of the mixed-precision refinement-step rules (cpm-ibc_amd/csrc/refine.h,
// mx_update), compiled for the host.  Prints "refine ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>

#define __host__
#define __device__
#include "refine.h"

using namespace cpmgpu;

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                             \
        }                                                             \
    } while (0)

// A lane after its first refinement step: ||b||^2 = 1, tol 1e-8, true residual rt0.
static KState lane(double rt0) {
    KState s{};
    s.bsq = 1.0;
    s.tol2 = 1e-16;        // tol^2 * ||b||^2
    s.max_it = 100000;
    s.delta2 = 1e-4;
    s.rr_true = rt0;
    s.status = ST_UPDATE;
    return s;
}

int main() {
    // a step asked for by the scaled tolerance test (judged = 2) that does not
    // reduce the true residual: the lane stalls short of the tolerance and
    // goes to the FP64 re-solve instead of rescaling until max_iters
    {
        KState s = lane(1e-12);
        s.judged = 2;
        mx_update(s, 1e-12);
        CHECK(s.status == ST_FALLBACK);
        KState t = lane(1e-12);
        t.judged = 2;
        mx_update(t, 2e-12);
        CHECK(t.status == ST_FALLBACK);
    }
    // a step of the rho restart rule may raise the true residual (a right-hand
    // side on identity rows after its first iteration): tolerated twice, the
    // third such step falls back, so a lane cannot restart until max_iters
    {
        KState s = lane(1e-6);
        double rt = 1e-6;
        for (int k = 0; k < 2; ++k) {
            s.judged = 0;
            rt *= 1.5;
            mx_update(s, rt);
            CHECK(s.status == ST_RESCALE && s.stalls == k + 1);
            s.status = ST_UPDATE;
        }
        s.judged = 0;
        mx_update(s, rt * 1.5);
        CHECK(s.status == ST_FALLBACK);
    }
    // an unjudged step that reduces the true residual rescales; delta and the
    // failure count are left alone (they judge delta-test steps only)
    {
        KState s = lane(1e-6);
        s.judged = 0;
        s.rr = 1e-16;
        s.r0sq = 1.0;
        mx_update(s, 0.9e-6);
        CHECK(s.status == ST_RESCALE && s.fails == 0 && s.delta2 == 1e-4);
        CHECK(std::fabs(s.scale - 1.0 / std::sqrt(0.9e-6)) < 1e-9 * s.scale);
    }
    // judged steps: too little true reduction three times -> re-solve; a true
    // reduction far short of the recurrence's raises delta
    {
        KState s = lane(1.0);
        s.judged = 1;
        s.rr = 1e-4;
        s.r0sq = 1.0;
        double rt = 1.0;
        for (int k = 0; k < 2; ++k) {
            rt *= 0.9;
            mx_update(s, rt);
            CHECK(s.status == ST_RESCALE);
            s.status = ST_UPDATE;
        }
        CHECK(s.delta2 > 1e-4);
        rt *= 0.9;
        mx_update(s, rt);
        CHECK(s.status == ST_FALLBACK && s.fails == 3);
    }
    // convergence on the true residual wins over every other rule
    {
        KState s = lane(1e-15);
        s.judged = 0;
        mx_update(s, 1e-17);
        CHECK(s.status == ST_CONVERGED);
    }
    // a fresh lane (new right-hand side) learns ||b||^2 and is never judged
    {
        KState s{};
        s.bsq = -1.0;
        s.tol2 = 1e-16;
        s.max_it = 10;
        s.rr_true = 0.0;
        s.status = ST_UPDATE;
        mx_update(s, 4.0);
        CHECK(s.status == ST_RESCALE && s.bsq == 4.0 && s.tol2 == 4e-16);
    }
    // maxit reported at the step asked for by the iteration limit
    {
        KState s = lane(1e-6);
        s.status = ST_UPDATE_MAXIT;
        mx_update(s, 1e-7);
        CHECK(s.status == ST_MAXIT);
    }
    std::printf("refine ok\n");
    return 0;
}
