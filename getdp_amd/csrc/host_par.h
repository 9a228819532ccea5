// Host-side helpers for splitting independent work over threads.
#pragma once

#include <algorithm>
#include <thread>
#include <vector>

namespace gdp {

// Threads for `items` independent items: 1 below 20000, else up to 16.
inline int worker_count(size_t items)
{
  unsigned hw = std::thread::hardware_concurrency();
  int n = (int)std::min<size_t>(hw ? hw : 4, 16);
  if(items < 20000) n = 1;
  return std::max(1, n);
}

// fn(a, b) over disjoint ranges covering [0, n), one per thread.
template <typename Fn> void parallel_ranges(size_t n, Fn fn)
{
  const int nt = worker_count(n);
  if(nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  for(int t = 0; t < nt; t++) {
    size_t a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([=] { fn(a, b); });
  }
  for(auto &x : th) x.join();
}

}  // namespace gdp
