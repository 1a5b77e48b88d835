// The reference's calling sequence, unchanged, on the MI355X engine:
//   cpm::Solver solver; solver.Init(m, n, L, E, diagonal, identity_rows);
//   VectorX u = solver.Solve(rhs);
// (e.g. examples/LaplaceTorusTorusKnot/LaplaceTorusTorusKnot.cpp:38-41), with
// L, E, identity rows and rhs read from files written by the test, plus the
// GPU-built tube path through the same facade.  Writes the solutions for the
// test to compare with the oracle.
//
// usage: solver_dropin <dir>   (dir holds D.txt, L_*.bin, E_*.bin, ident.bin, b.bin, mn.txt)
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cpm_solver.h"

template <typename T>
static std::vector<T> read_bin(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v((size_t)n / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), n);
    return v;
}

template <typename T>
static void write_bin(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string d = argv[1];
    try {
        int64_t D = 0;
        double m = 0, n = 0;
        {
            std::ifstream f(d + "/mn.txt");
            f >> D >> m >> n;
        }
        cpm::SpMat L, E;
        L.assign(D, D, read_bin<int32_t>(d + "/L_ptr.bin"), read_bin<int32_t>(d + "/L_col.bin"),
                 read_bin<double>(d + "/L_val.bin"));
        E.assign(D, D, read_bin<int32_t>(d + "/E_ptr.bin"), read_bin<int32_t>(d + "/E_col.bin"),
                 read_bin<double>(d + "/E_val.bin"));
        const auto id8 = read_bin<uint8_t>(d + "/ident.bin");
        std::vector<bool> identity_rows(id8.begin(), id8.end());
        cpm::VectorX rhs = read_bin<double>(d + "/b.bin");

        // exactly the reference's calls
        cpm::Solver solver;
        solver.Init(m, n, L, E, cpm::diagonal, identity_rows);
        cpm::VectorX u = solver.Solve(rhs);
        write_bin(d + "/x_fp64.bin", u);
        const cpm_solve_info i64 = solver.LastInfo();

        cpm::VectorX Au;
        solver.Multiply(u, Au, m, n, L, E);
        write_bin(d + "/Ax.bin", Au);

        solver.SetPrecision(CPM_FP32_REFINE);
        cpm::VectorX u32 = solver.Solve(rhs);
        write_bin(d + "/x_fp32.bin", u32);
        const cpm_solve_info i32 = solver.LastInfo();

        // batch of three right-hand sides (rhs, 2 rhs, -rhs)
        std::vector<cpm::VectorX> bs{rhs, rhs, rhs};
        for (auto& v : bs[1]) v *= 2.0;
        for (auto& v : bs[2]) v = -v;
        solver.SetPrecision(CPM_FP64);
        auto xs = solver.Solve(bs);
        std::vector<double> cat;
        for (auto& x : xs) cat.insert(cat.end(), x.begin(), x.end());
        write_bin(d + "/x_batch.bin", cat);

        // error path: a wrong-size right-hand side throws cpm::Error
        bool threw = false;
        try {
            cpm::VectorX bad(3, 1.0);
            solver.Solve(bad);
        } catch (const cpm::Error& e) {
            threw = e.status == CPM_ERR_ARG;
        }
        std::printf("{\"D\": %lld, \"fp64_status\": %d, \"fp64_iters\": %d, \"fp64_rel\": %.3e, "
                    "\"fp32_status\": %d, \"fp32_iters\": %d, \"fp32_updates\": %d, \"threw\": %d}\n",
                    (long long)D, i64.status, i64.iters, i64.rel_residual, i32.status, i32.iters, i32.updates,
                    threw ? 1 : 0);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
