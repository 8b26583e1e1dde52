// Host-side pieces shared by the run drivers (capi.cpp) and the serving
// loop (serve.cpp): the output::run_with chunk schedule and a small
// parallel-for over the host's cores.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <thread>
#include <vector>

#include "host.hpp"

namespace cpm {

// One step of a replica's driver schedule: advance `mcs`, then either push
// the energy onto the burn-in trace or take a readout sample.
struct Step {
    uint64_t mcs;
    bool trace, sample;
};

// output.rs:441-468: burn-in in chunks of min(interval, remaining), the
// energy traced once burned >= burn - floor(burn * 0.25); then
// max(readout / interval, 1) readout chunks of `interval`, one sample each.
inline std::vector<Step> schedule(const Config& c) {
    std::vector<Step> s;
    const uint64_t burn = c.burn_in, interval = c.interval;
    const uint64_t window = (uint64_t)((double)burn * 0.25);
    const uint64_t start = burn > window ? burn - window : 0;
    uint64_t burned = 0;
    while (burned < burn) {
        uint64_t chunk = std::min(interval, burn - burned);
        burned += chunk;
        s.push_back({chunk, burned >= start, false});
    }
    const uint64_t n = std::max<uint64_t>(c.readout / std::max<uint64_t>(interval, 1), 1);
    for (uint64_t i = 0; i < n; ++i) s.push_back({interval, false, true});
    return s;
}

// Runs f(i) for i in [0, n) over up to 16 host threads.
template <class F>
void parallel_for(size_t n, F&& f) {
    size_t nt = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    nt = std::min(nt, n);
    if (nt <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::exception_ptr> errs(nt);
    std::vector<std::thread> ts;
    for (size_t t = 0; t < nt; ++t)
        ts.emplace_back([&, t] {
            try {
                for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
            } catch (...) {
                errs[t] = std::current_exception();
            }
        });
    for (auto& t : ts) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

}  // namespace cpm
