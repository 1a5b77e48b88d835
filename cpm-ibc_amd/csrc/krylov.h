// Launch interface of the batched BiCGStab kernels (krylov.hip).
#pragma once
#include "engine.h"

namespace cpmgpu {

template <typename T>
struct LaunchCfg {
    const Seg<T>* segs;    // device segment table
    const int* bseg;       // device block -> segment map
    int nblocks;
    int nseg;
    hipStream_t stream;
};

// Workspace vectors of one precision; every vector spans all segments
// (segment g occupies [off_g, off_g + D_g * NR)).
template <typename T>
struct Vecs {
    T *b, *x, *r, *r0, *p, *v, *s, *t, *y, *z, *tmp;
};

template <typename T, int NR>
struct Launch {
    static void init(const LaunchCfg<T>& c, const T* b, T* x, T* r, T* r0, T* p, T* v);
    // mixed = true: the loop test asks for reliable updates instead of stopping
    static void iteration(const LaunchCfg<T>& c, Vecs<T>& w, bool mixed);
    static void restart(const LaunchCfg<T>& c, Vecs<T>& w);
    // y = A x; act >= 0 restricts the work to instances with a lane in that state
    static void apply(const LaunchCfg<T>& c, const T* x, T* tmp, T* y, int act = -1);
};

// FP32 BiCGStab with reliable FP64 updates (the segments of both precisions
// share one state array).
template <int NR>
struct MixedLaunch {
    static void start(const LaunchCfg<double>& c, const double* b, double* x64, double* r64);
    static void down(const LaunchCfg<double>& c, const double* r64, Vecs<float>& w);
    static void update(const LaunchCfg<double>& c, const double* b, double* x64, double* r64, double* ax, double* tmp,
                       Vecs<float>& w);
};

// E-gather alone (identity rows copy), for measurement.
template <typename T, int NR>
void launch_egather(const LaunchCfg<T>& c, const T* x, T* out);

// State reset before a (re)started solve: status, tolerance^2, iteration cap.
void launch_reset_states(KState* st, int n, double tol2, int max_it, double delta2, hipStream_t s);

}  // namespace cpmgpu
