// host_parallel.h — minimal std::thread fan-out used by the host-side builders.
#pragma once

#include <cstdint>
#include <thread>
#include <vector>

namespace ctc {

// Runs f(t) for t in [0, T) on T threads (inline when T == 1).
template <class F>
void run_threads(unsigned T, F &&f) {
  if (T <= 1) {
    f(0u);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(T);
  for (unsigned t = 0; t < T; ++t) pool.emplace_back([&f, t] { f(t); });
  for (auto &th : pool) th.join();
}

// Splits [0, n) into T contiguous ranges and runs f(begin, end, t) on each.
template <class F>
void parallel_for(uint64_t n, unsigned T, F &&f) {
  if (n == 0) return;
  if (T > n) T = (unsigned)n;
  if (T < 1) T = 1;
  run_threads(T, [&](unsigned t) { f(n * t / T, n * (t + 1) / T, t); });
}

}  // namespace ctc
