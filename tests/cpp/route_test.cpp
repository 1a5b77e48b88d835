// CPU test of the FP64 routing decision (cpm-ibc_amd/csrc/route.h).  Prints
// "route ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>

#include "route.h"

using namespace cpmgpu;

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                                        \
        }                                                                        \
    } while (0)

int main() {
    {   // no history, or no fallback: FP32 (nothing to compare with)
        Fp32Record r;
        CHECK(!r.prefer_fp64());
        for (int k = 0; k < 50; ++k) r.add(2000, 0);   // however many FP32 iterations
        CHECK(!r.prefer_fp64());
    }
    {   // every system fell back (fine screened grids): FP64
        Fp32Record r;
        r.add(577, 1005);
        CHECK(r.prefer_fp64());
        CHECK(r.solved == 1 && r.fell == 1 && r.it32_lost == 577 && r.it64 == 1005 && r.it32 == 0);
    }
    {   // torus-knot Laplace system: 1 of 32 fell back, FP32 iterations ~1.15x FP64: FP32 stays
        Fp32Record r;
        for (int k = 0; k < 31; ++k) r.add(445, 0);
        r.add(300, 388);
        CHECK(!r.prefer_fp64());
    }
    {   // screened h = 0.02: 3 of 8 fell back, FP32 needs ~1.85x the FP64 iterations: FP64
        Fp32Record r;
        for (int k = 0; k < 5; ++k) r.add(891, 0);
        for (int k = 0; k < 3; ++k) r.add(500, 482);
        CHECK(r.prefer_fp64());
    }
    {   // the boundary: (1 - f) c i32 + f (c lost + i64) against i64, f = 1/2, lost = 0
        // FP64 exactly when c i32 / 2 + i64 / 2 > i64, i.e. i32 > i64 / c
        Fp32Record a, b;
        a.add(100 * 10 / 6 + 1, 0);   // i32 just above i64 / c (c = 0.6, i64 = 100)
        a.add(0, 100);
        b.add(100 * 10 / 6 - 1, 0);   // just below
        b.add(0, 100);
        CHECK(a.prefer_fp64());
        CHECK(!b.prefer_fp64());
    }
    std::printf("route ok\n");
    return 0;
}
