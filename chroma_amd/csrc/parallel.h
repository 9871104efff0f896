// Host-side parallel loop over [0, n) on std::thread workers (no OpenMP runtime in the library).
#pragma once
#include <algorithm>
#include <thread>
#include <vector>

namespace cmx {

template <typename F> void parallel_for(long n, F&& f) {
  const long hw = std::max(1u, std::thread::hardware_concurrency());
  const long nt = std::min<long>(std::min<long>(hw, 16), std::max<long>(1, n / 4096));
  if (nt <= 1) {
    for (long i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (long t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const long lo = n * t / nt, hi = n * (t + 1) / nt;
      for (long i = lo; i < hi; ++i) f(i);
    });
  for (auto& x : th) x.join();
}

// max-reduction variant
template <typename F> double parallel_max(long n, F&& f) {
  const long hw = std::max(1u, std::thread::hardware_concurrency());
  const long nt = std::min<long>(std::min<long>(hw, 16), std::max<long>(1, n / 4096));
  std::vector<double> part(nt, 0.0);
  std::vector<std::thread> th;
  for (long t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const long lo = n * t / nt, hi = n * (t + 1) / nt;
      double m = 0;
      for (long i = lo; i < hi; ++i) m = std::max(m, f(i));
      part[t] = m;
    });
  for (auto& x : th) x.join();
  return *std::max_element(part.begin(), part.end());
}

// sum-reduction variant (fixed partition, so the result is deterministic for a given thread count)
template <typename F> double parallel_sum(long n, F&& f) {
  const long nt = 16;
  std::vector<double> part(nt, 0.0);
  std::vector<std::thread> th;
  for (long t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const long lo = n * t / nt, hi = n * (t + 1) / nt;
      double m = 0;
      for (long i = lo; i < hi; ++i) m += f(i);
      part[t] = m;
    });
  for (auto& x : th) x.join();
  double s = 0;
  for (double v : part) s += v;
  return s;
}

}  // namespace cmx
