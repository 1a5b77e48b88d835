// CPU test of the serving loop's host bookkeeping (cpm-ibc_amd/csrc/pool.h):
// the first-fit range pool, and retirement of finished requests when writing
// out one request's results throws.  Prints "pool ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "pool.h"

using cpmgpu::RangePool;

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                             \
        }                                                             \
    } while (0)

struct Req {
    int64_t off, len;
    int slot;
    int tag;
};

int main() {
    // first fit, split, merge on both sides
    RangePool p(1000);
    int64_t a, b, c, d;
    CHECK(p.take(100, a) && a == 0);
    CHECK(p.take(200, b) && b == 100);
    CHECK(p.take(300, c) && c == 300);
    CHECK(!p.take(401, d));
    CHECK(p.take(400, d) && d == 600);
    CHECK(p.free_ranges().empty());
    CHECK(p.give(b, 200));                 // hole [100, 300)
    CHECK(!p.give(b, 200));                // twice: refused, nothing changes
    CHECK(!p.give(150, 10));               // inside a free range: refused
    CHECK(p.free_ranges().size() == 1);
    CHECK(p.take(50, b) && b == 100);      // first fit inside the hole
    CHECK(p.give(b, 50));
    CHECK(p.give(a, 100));                 // merges with [100, 300)
    CHECK(p.free_ranges().size() == 1 && p.free_ranges().begin()->second == 300);
    CHECK(p.give(d, 400));
    CHECK(p.give(c, 300));                 // merges both neighbours
    CHECK(p.free_ranges().size() == 1 && p.free_ranges().begin()->first == 0 &&
          p.free_ranges().begin()->second == 1000);
    CHECK(!p.give(990, 20));               // past the pool

    // retirement: four requests leave; writing out the second and fourth throws
    RangePool pool(4096);
    std::vector<int> free_slots;
    std::vector<Req> leaving;
    for (int i = 0; i < 4; ++i) {
        int64_t off;
        CHECK(pool.take(1024, off));
        leaving.push_back(Req{off, 1024, i, 100 + i});
    }
    std::vector<int> finished, released;
    bool threw = false;
    try {
        cpmgpu::retire_requests(
            leaving,
            [&](const Req& q) {
                if (q.tag == 101) throw std::runtime_error("fallback allocation failed");
                if (q.tag == 103) throw std::logic_error("second failure");
                finished.push_back(q.tag);
            },
            [&](const Req& q) {
                CHECK(pool.give(q.off, q.len));   // each exactly once
                free_slots.push_back(q.slot);
                released.push_back(q.tag);
            });
    } catch (const std::runtime_error& e) {
        threw = true;   // the first failure is the one rethrown
    }
    CHECK(threw);
    CHECK(leaving.empty());
    CHECK((finished == std::vector<int>{100, 102}));
    CHECK((released == std::vector<int>{100, 101, 102, 103}));
    CHECK(free_slots.size() == 4);
    CHECK(pool.free_ranges().size() == 1 && pool.free_ranges().begin()->second == 4096);
    // the pool hands the memory out again without overlap
    int64_t o1, o2;
    CHECK(pool.take(2048, o1) && pool.take(2048, o2) && o1 != o2);
    std::printf("pool ok\n");
    return 0;
}
