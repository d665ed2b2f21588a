// Host-side parallel loop over index ranges (std::thread; no OpenMP runtime needed).
#pragma once

#include <algorithm>
#include <cstddef>
#include <thread>
#include <vector>

namespace crvpt {

// Runs fn(begin, end, worker) over [0, n) split into contiguous chunks on up to 16 host threads (1 for small n).
template <class Fn>
void parallelFor(size_t n, Fn fn, size_t minPerThread = 4096) {
    size_t threads = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    threads = std::max<size_t>(1, std::min(threads, n / std::max<size_t>(minPerThread, 1)));
    if (threads <= 1) {
        fn(size_t(0), n, size_t(0));
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + threads - 1) / threads;
    for (size_t t = 0; t < threads; t++) {
        const size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([=, &fn] { fn(b, e, t); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace crvpt
