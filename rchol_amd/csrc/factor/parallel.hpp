// Minimal fork-join loop over [0, n) in contiguous chunks (host threads).
#pragma once
#include <algorithm>
#include <cstddef>
#include <thread>
#include <vector>

namespace rchol_host {

inline unsigned worker_count() {
  unsigned t = std::thread::hardware_concurrency();
  return std::max(1u, std::min(t ? t : 1u, 16u));
}

// Below serial_below items the loop runs on the caller's thread.
template <typename F>
void parallel_for(size_t n, F fn, size_t serial_below = 4096) {
  const unsigned T = worker_count();
  if (T == 1 || n < serial_below) {
    fn((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + T - 1) / T;
  for (unsigned t = 0; t < T; t++) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=]() { fn(a, b); });
  }
  for (auto &x : th) x.join();
}

}  // namespace rchol_host
