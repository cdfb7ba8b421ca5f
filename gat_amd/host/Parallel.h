#pragma once
// Static slicing of an index range over host threads.
#include <algorithm>
#include <thread>
#include <vector>

// Calls fn(i) for every i in [0, n) on up to `threads` threads; thread t
// takes the t-th contiguous slice, so each index is visited exactly once.
template <class Fn>
void parallel_for(int n, int threads, Fn&& fn) {
  const int t = std::max(1, std::min(threads, n));
  if (t <= 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(size_t(t));
  for (int w = 0; w < t; ++w) {
    const int b = int(int64_t(n) * w / t), e = int(int64_t(n) * (w + 1) / t);
    pool.emplace_back([&fn, b, e] {
      for (int i = b; i < e; ++i) fn(i);
    });
  }
  for (auto& th : pool) th.join();
}
