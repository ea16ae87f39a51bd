// Minimal host thread pool for the per-cell set-up loops (independent
// iterations; results do not depend on the thread count).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace nsem {

inline unsigned host_threads() {
    if (const char* e = std::getenv("NSEM_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) return (unsigned)v;
    }
    unsigned n = std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) n = std::min<unsigned>(n ? n : v, (unsigned)v);
    }
    return std::max(1u, std::min(n ? n : 1u, 32u));
}

// fn(begin, end) over [0, n) split into contiguous chunks.
template <typename F>
void parallel_for(uint64_t n, F fn) {
    const unsigned nt = (unsigned)std::min<uint64_t>(host_threads(), std::max<uint64_t>(1, n / 64));
    if (nt <= 1) { fn(uint64_t(0), n); return; }
    std::vector<std::thread> th;
    const uint64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const uint64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &fn] { fn(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace nsem
