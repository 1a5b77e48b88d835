// Internal types of the MI355X closest-point engine (libcpmgpu.so).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace cpmgpu {

constexpr int kBlock = 256;          // 4 wave64 per workgroup
constexpr int kMaxNR = 8;            // right-hand sides per multi-RHS sweep

// Row flags (one byte per row).
enum : uint8_t { ROW_IDENTITY = 1, ROW_NODIAG = 2 };

// Per (instance, rhs) Krylov state, written only by the finaliser block of a
// reducing kernel and read (uniformly) by every block of the next kernels.
enum : int { ST_RUN = 0, ST_CONVERGED = 1, ST_MAXIT = 2, ST_RESTART = 3, ST_ZERO_RHS = 4, ST_BREAKDOWN = 5 };

struct KState {
    double rho, rho_old, alpha, omega, beta;
    double bsq, r0sq, tol2, rr;
    double scale;        // refinement: |r64| of the current correction solve
    int iters, restarts, status, max_it;
    int outer_status, refinements, total_iters, pad;
};

// Device view of an operator in one precision.  E and Loff are ELL matrices
// stored slot-major (entry s of row r at [s * D + r]) so that a wavefront's 64
// rows read 64 consecutive words per slot.
template <typename T>
struct OpView {
    int64_t D;
    int nD;
    int ew, lw;          // ELL widths of E and of L without its diagonal
    const int* ecol;
    const T* eval;
    const int* lcol;
    const T* lval;
    const T* ldiag;      // L_ii (0 where L has no diagonal)
    const uint8_t* flag;
};

// One instance group of a batched launch: an operator, NR right-hand sides
// and the contiguous block range that covers its D rows.
template <typename T>
struct Seg {
    OpView<T> op;
    double m, n;
    int64_t off;         // element offset of row 0 in every workspace vector (rows interleave NR rhs)
    int block0, nblocks;
    int nr;              // rhs in this group (<= template NR; extra lanes idle)
    int precond;         // CPM_PRECOND_*
    KState* st;          // nr states
    double* part;        // nblocks * NR * 2 partial sums
    unsigned* cnt;       // arrival counter for the finaliser
};

void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::cpmgpu::hip_check((x), #x)

struct Buffer {
    void* p = nullptr;
    size_t bytes = 0;
    Buffer() = default;
    explicit Buffer(size_t n);
    ~Buffer();
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    Buffer(Buffer&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    Buffer& operator=(Buffer&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    template <typename U> U* as() const { return static_cast<U*>(p); }
};

}  // namespace cpmgpu
