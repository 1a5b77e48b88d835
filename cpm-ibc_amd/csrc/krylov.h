// Launch interface of the batched BiCGStab kernels (krylov.hip).
#pragma once
#include "engine.h"

namespace cpmgpu {

template <typename T>
struct LaunchCfg {
    const Seg<T>* segs;    // device segment table
    const int* bseg;       // device block maps (segment | block << kSegBits), nblocks entries segment by
                           // segment, then nblocks round-robin over the segments (the E gathers')
    int nblocks;
    int nseg;
    hipStream_t stream;
    int jacobi_sweeps;     // largest precond_iters over Jacobi-preconditioned segments (0: none)
    int grp = 0;           // fewest shared-stencil tiles of the segments' operators (0: some have none)
};

// Whether E gathers of T x nr values use shared-stencil tiles when every
// operator of the launch has at least ntiles of them.
bool grp_gather(int tsize, int nr, int ntiles);
// Whether a launch of T x nr values builds its operators' shared-stencil tiles
// (other shapes use them only where an earlier launch built them).
bool grp_build(int tsize, int nr);
// Re-read the CPMGPU_GROUP_GATHER override (once per launch sequence).
void grp_refresh_override();

// Workspace vectors of one precision; every vector spans all segments
// (segment g occupies [off_g, off_g + D_g * NR)).
template <typename T>
struct Vecs {
    T *b, *x, *r, *r0, *p, *v, *s, *t, *y, *z, *tmp;
};

// FP64 vectors of a mixed-precision solve (same element layout as the FP32
// workspace), used by the FP32 iteration kernels for refinement steps.
struct MixVec {
    double* x64;          // solution
    double* r64;          // true residual b - A64 x64
    const double* b64;    // right-hand sides
    double* tmp64;        // E64 x64
};

template <typename T, int NR>
struct Launch {
    static void init(const LaunchCfg<T>& c, const T* b, T* x, T* r, T* r0, T* p, T* v);
    // mixed: FP32 recurrence whose lanes take FP64 refinement steps in place
    // (mv holds the FP64 vectors); otherwise the reference's FP64 loop
    static void iteration(const LaunchCfg<T>& c, Vecs<T>& w, bool mixed, const MixVec& mv);
    static void restart(const LaunchCfg<T>& c, Vecs<T>& w);
    // y = A x; lanes = status mask restricting the work to instances with such a lane (0: all)
    static void apply(const LaunchCfg<T>& c, const T* x, T* tmp, T* y, unsigned lanes = 0);
};

// Start of a mixed-precision solve: x64 = 0, r64 = b, then the FP32 vectors
// from the scaled residual (the segments of both precisions share one state
// array).
template <int NR>
struct MixedLaunch {
    static void start(const LaunchCfg<double>& c, const double* b, double* x64, double* r64);
    static void down(const LaunchCfg<double>& c, const double* r64, Vecs<float>& w);
};

// out = m x + n (L_ii x + L_off tmp) alone (the tail of an apply).
template <typename T>
void launch_lfinish_plain(const LaunchCfg<T>& c, const T* x, const T* tmp, T* out);

// E-gather alone (identity rows copy), for measurement; fly: of P^-1 x, as
// the iterations gather (where the launch uses shared-stencil tiles).
template <typename T, int NR>
void launch_egather(const LaunchCfg<T>& c, const T* x, T* out, bool fly = false);

// State reset before a (re)started solve: status, tolerance^2, iteration cap.
void launch_reset_states(KState* st, int n, double tol2, int max_it, double delta2, hipStream_t s);

}  // namespace cpmgpu
