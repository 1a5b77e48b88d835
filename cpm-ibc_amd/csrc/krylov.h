// Launch interface of the batched BiCGStab kernels (krylov.hip).
#pragma once
#include "engine.h"

namespace cpmgpu {

template <typename T>
struct LaunchCfg {
    const Seg<T>* segs;    // device segment table
    const int* bseg;       // device block -> segment map
    int nblocks;
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
    static void iteration(const LaunchCfg<T>& c, Vecs<T>& w);
    static void restart(const LaunchCfg<T>& c, Vecs<T>& w);
    static void apply(const LaunchCfg<T>& c, const T* x, T* tmp, T* y);
};

template <int NR>
struct RefLaunch {
    static void start(const LaunchCfg<double>& c, const double* b, double* x, double* r);
    static void down(const LaunchCfg<double>& c, const double* r, float* b32);
    static void up(const LaunchCfg<double>& c, const float* d32, double* x);
    static void resid(const LaunchCfg<double>& c, const double* b, const double* ax, double* r, const KState* inner);
};

// E-gather alone (identity rows copy), for measurement.
template <typename T, int NR>
void launch_egather(const LaunchCfg<T>& c, const T* x, T* out);

// State reset before a (re)started solve: status, tolerance^2, iteration cap.
void launch_reset_states(KState* st, int n, double tol2, int max_it, hipStream_t s);

}  // namespace cpmgpu
