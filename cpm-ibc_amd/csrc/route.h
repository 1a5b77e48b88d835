// FP64 routing of FP32+refine solves: an operator's record of how its
// systems fared in FP32+refine, and the decision to start its next ones in
// FP64.  Plain C++ so that a CPU test compiles it (tests/cpp/route_test.cpp).
#pragma once
#include <cstdint>

namespace cpmgpu {

// Cost of an FP32 x 8 iteration per right-hand side relative to an FP64 x 4
// one (torus-knot system: 445 FP32 and 388 FP64 iterations per solve at 105
// and 62 solves/s give 0.52; rounded up, towards FP64).
constexpr double kFp32IterCost = 0.6;

// With a fraction f of an operator's systems falling back to FP64 (their FP32
// refinement stalled), FP32+refine costs per system about
//   (1 - f) c i32 + f (c i32_lost + i64)   FP64-iteration equivalents
// and plain FP64 i64, where i32 is the mean FP32 iterations of the systems
// that converged in FP32, i32_lost the mean FP32 iterations spent before a
// fallback, i64 the mean FP64 iterations of the re-solves and c =
// kFp32IterCost.  An operator without a fallback keeps FP32 (there is no FP64
// iteration count to compare with).
struct Fp32Record {
    int64_t solved = 0;      // FP32+refine systems
    int64_t fell = 0;        // of which fell back to FP64
    int64_t it32 = 0;        // FP32 iterations of the systems that converged in FP32
    int64_t it32_lost = 0;   // FP32 iterations of the systems that fell back
    int64_t it64 = 0;        // FP64 iterations of their re-solves

    void add(int iters32, int fallback_iters64) {
        ++solved;
        if (fallback_iters64 > 0) {
            ++fell;
            it32_lost += iters32;
            it64 += fallback_iters64;
        } else {
            it32 += iters32;
        }
    }
    bool prefer_fp64() const {
        if (fell == 0 || solved == 0) return false;
        const double f = (double)fell / (double)solved;
        const double i32 = solved > fell ? (double)it32 / (double)(solved - fell) : 0.0;
        const double lost = (double)it32_lost / (double)fell, i64 = (double)it64 / (double)fell;
        return (1.0 - f) * kFp32IterCost * i32 + f * (kFp32IterCost * lost + i64) > i64;
    }
};

}  // namespace cpmgpu
