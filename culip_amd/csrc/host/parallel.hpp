// Host thread pool for the symbolic phase: a loop over [0, n) split into
// chunks that threads claim in order. Callers only write to positions fixed
// by the loop index (or computed from counts by a prefix sum between two
// such loops), so every result is the same for any number of threads and
// any schedule.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace culip {

// threads for host loops: CULIP_HOST_THREADS if set (>= 1), else the
// hardware threads, at most 16
inline int host_threads()
{
    static const int t = [] {
        if(const char* e = std::getenv("CULIP_HOST_THREADS")) {
            const int v = std::atoi(e);
            if(v >= 1) return std::min(v, 256);
        }
        const unsigned hw = std::thread::hardware_concurrency();
        return (int) std::max(1u, std::min(hw, 16u));
    }();
    return t;
}

// fn(worker, lo, hi) for consecutive chunks [lo, hi) of [0, n); worker is
// in [0, workers) and identifies per-thread scratch. The first exception
// thrown by any call is rethrown once every thread has stopped.
template<typename Fn>
void parallel_chunks(int64_t n, int64_t chunk, Fn&& fn)
{
    if(n <= 0) return;
    chunk = std::max<int64_t>(1, chunk);
    const int64_t nch = (n + chunk - 1) / chunk;
    const int workers = (int) std::min<int64_t>(host_threads(), nch);
    if(workers <= 1) {
        fn(0, (int64_t) 0, n);
        return;
    }
    std::atomic<int64_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    auto run = [&](int w) {
        for(;;) {
            const int64_t c = next.fetch_add(1);
            if(c >= nch) return;
            try {
                fn(w, c * chunk, std::min(n, (c + 1) * chunk));
            } catch(...) {
                std::lock_guard<std::mutex> lk(mu);
                if(!err) err = std::current_exception();
                next.store(nch);
                return;
            }
        }
    };
    std::vector<std::thread> th;
    th.reserve(workers - 1);
    for(int w = 1; w < workers; ++w) th.emplace_back(run, w);
    run(0);
    for(auto& t : th) t.join();
    if(err) std::rethrow_exception(err);
}

// number of workers parallel_chunks uses for n items in chunks of `chunk`
inline int parallel_workers(int64_t n, int64_t chunk)
{
    if(n <= 0) return 1;
    const int64_t nch = (n + std::max<int64_t>(1, chunk) - 1) / std::max<int64_t>(1, chunk);
    return (int) std::max<int64_t>(1, std::min<int64_t>(host_threads(), nch));
}

}  // namespace culip
