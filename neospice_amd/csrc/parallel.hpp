// A small host-side parallel loop for per-instance CPU work (pivot
// orderings of the instances of a batch).
#pragma once
#include <algorithm>
#include <atomic>
#include <exception>
#include <thread>
#include <vector>

namespace nsa {

// Runs f(i) for i < count on up to 16 host threads; rethrows the first
// exception a call raised.
template <class F>
void parallel_for(int count, F&& f) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::min({hw, 16, count});
    if (nt <= 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> err(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            try {
                for (int i; (i = next.fetch_add(1)) < count;) f(i);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace nsa
