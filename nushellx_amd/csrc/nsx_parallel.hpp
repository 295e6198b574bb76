// Host-side parallel loop over independent items (instances of a batch).
#pragma once
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace nsx {

// f(i) for i in [0, n) on up to 16 host threads (dynamic assignment); serial
// when there are fewer than `min_per_thread` items per thread.
template <typename F>
void parallel_for(int n, F&& f, int min_per_thread = 8) {
  int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  nt = std::min(nt, std::max(1, n / std::max(1, min_per_thread)));
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
    });
  for (auto& th : pool) th.join();
}

}  // namespace nsx
