// Deterministic host parallelism for setup: each index is processed by
// exactly one thread and writes only its own outputs, so results do not
// depend on the thread count.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <thread>
#include <vector>

namespace whip {

template <typename F> void parallel_for(size_t n, F fn) {
  unsigned threads = std::thread::hardware_concurrency();
  if (const char *env = std::getenv("WASL_HIP_SETUP_THREADS"))
    threads = static_cast<unsigned>(std::max(1, std::atoi(env)));
  threads = std::max(1u, std::min(threads, 16u));
  if (n < 4096 || threads == 1) {
    fn(size_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  const size_t chunk = (n + threads - 1) / threads;
  for (unsigned t = 0; t < threads; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e)
      break;
    pool.emplace_back([=] { fn(b, e); });
  }
  for (auto &th : pool)
    th.join();
}

}  // namespace whip
