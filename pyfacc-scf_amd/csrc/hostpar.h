// Host-side parallel loop for set-up work that is independent per molecule
// (pair selection, quartet plans, guess densities of a batch).
#pragma once

#include <algorithm>
#include <exception>
#include <thread>
#include <vector>

namespace pf {

// Threads for n independent items, at least `grain` items per thread, at
// most 16 (the CPU share of one GPU's process on a shared node).
inline int host_threads(int n, int grain) {
    const int hw = std::max(1, (int)std::thread::hardware_concurrency());
    return std::max(1, std::min({16, hw, n / std::max(grain, 1)}));
}

// f(t, begin, end) over [0, n) cut into T contiguous ranges of about equal
// size, range t on its own thread (range 0 on the caller's).  The first
// exception any range throws is rethrown after all have ended.
template <class F>
void parallel_ranges(int n, int T, F &&f) {
    T = std::max(1, std::min(T, std::max(n, 1)));
    auto lo = [&](int t) { return (int)((long long)n * t / T); };
    if (T == 1) {
        f(0, 0, n);
        return;
    }
    std::vector<std::exception_ptr> err(T);
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (int t = 1; t < T; ++t)
        th.emplace_back([&, t] {
            try {
                f(t, lo(t), lo(t + 1));
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    try {
        f(0, 0, lo(1));
    } catch (...) {
        err[0] = std::current_exception();
    }
    for (auto &x : th) x.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace pf
