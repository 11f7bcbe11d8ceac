// Shared host-side helpers: error reporting across the C ABI and a small
// std::thread parallel-for (no OpenMP runtime, so the library never competes
// with the one torch loads).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "jfm.h"

namespace jfm {

void set_error(const std::string& msg);
int32_t fail(int32_t code, const std::string& msg);

inline int resolve_threads(int requested) {
  if (requested > 0) return requested;
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? static_cast<int>(std::min(hc, 64u)) : 1;
}

// Runs f(begin, end) over [0, n) split into contiguous chunks, one per thread.
template <typename F>
void parallel_chunks(int64_t n, int threads, F&& f) {
  if (n <= 0) return;
  int t = std::max(1, std::min<int>(threads, static_cast<int>((n + 1023) / 1024)));
  if (t == 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(t);
  for (int k = 0; k < t; ++k) {
    int64_t b = n * k / t, e = n * (k + 1) / t;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

// Dynamic scheduling: f(i) for i in [0, n), chunks of `grain` handed out atomically.
template <typename F>
void parallel_dynamic(int64_t n, int threads, int64_t grain, F&& f) {
  if (n <= 0) return;
  int t = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(threads, (n + grain - 1) / grain)));
  std::atomic<int64_t> next{0};
  auto worker = [&] {
    for (;;) {
      int64_t b = next.fetch_add(grain);
      if (b >= n) break;
      int64_t e = std::min(n, b + grain);
      for (int64_t i = b; i < e; ++i) f(i);
    }
  };
  if (t == 1) {
    worker();
    return;
  }
  std::vector<std::thread> pool;
  for (int k = 0; k < t; ++k) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
}

}  // namespace jfm
