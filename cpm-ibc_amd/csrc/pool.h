// Host-only bookkeeping of the serving loop (no HIP): the first-fit range
// pool that hands out workspace vector ranges, and the retirement of
// finished requests, which returns every request's range and state slot
// exactly once even when writing out one request's results throws.
// Header-only so that a CPU test can exercise it (tests/cpp/pool_test.cpp).
#pragma once
#include <cstdint>
#include <exception>
#include <iterator>
#include <map>
#include <vector>

namespace cpmgpu {

// Ranges [off, off + len) of a pool of `cap` elements, first fit, adjacent
// free ranges merged.
class RangePool {
public:
    explicit RangePool(int64_t cap = 0) { reset(cap); }
    void reset(int64_t cap) {
        free_.clear();
        cap_ = cap;
        if (cap > 0) free_[0] = cap;
    }
    bool take(int64_t len, int64_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= len) {
                off = it->first;
                const int64_t rest = it->second - len;
                free_.erase(it);
                if (rest > 0) free_[off + len] = rest;
                return true;
            }
        return false;
    }
    // Returns false (and changes nothing) if the range overlaps a free one:
    // a range given back twice.
    bool give(int64_t off, int64_t len) {
        if (len <= 0 || off < 0 || off + len > cap_) return false;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && nx->first < off + len) return false;
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second > off) return false;
        }
        auto it = free_.emplace(off, len).first;
        nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
        return true;
    }
    int64_t cap() const { return cap_; }
    const std::map<int64_t, int64_t>& free_ranges() const { return free_; }

private:
    int64_t cap_ = 0;
    std::map<int64_t, int64_t> free_;   // start -> length
};

// Retire the requests that left the running set (the caller has already
// removed them from it): finish(q) writes out q's results; release(q) gives
// q's range and slot back.  Every request is released exactly once, after
// its finish ran or threw; if any finish threw, the other requests are still
// finished and released, and the first exception is rethrown at the end.
template <typename Req, typename Finish, typename Release>
void retire_requests(std::vector<Req>& leaving, Finish finish, Release release) {
    std::exception_ptr first;
    for (Req& q : leaving) {
        try {
            finish(q);
        } catch (...) {
            if (!first) first = std::current_exception();
        }
        release(q);
    }
    leaving.clear();
    if (first) std::rethrow_exception(first);
}

}  // namespace cpmgpu
