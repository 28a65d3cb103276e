// Host threads for the mesh builders: a blocking parallel for over
// contiguous chunks.  Work is split the same way for a given size and thread
// count, and callers combine per-chunk results in chunk order, so results
// never depend on timing.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace rmp {

// Threads used: RICH_MESH_THREADS if set, else the hardware's, at most 16.
inline int threads() {
  static const int t = [] {
    if (const char* e = std::getenv("RICH_MESH_THREADS")) {
      const int v = std::atoi(e);
      if (v >= 1) return std::min(v, 64);
    }
    const unsigned h = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(h, 16u)));
  }();
  return t;
}

// Number of chunks for n items (at least `grain` items per chunk).
inline int chunks(int64_t n, int64_t grain = 16384) {
  const int64_t c = (n + grain - 1) / grain;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(c, threads())));
}

// fn(chunk, begin, end) for `nc` contiguous chunks of [0, n); chunk 0 runs
// on the calling thread.
template <class Fn>
void for_chunks(int64_t n, int nc, Fn&& fn) {
  auto range = [&](int k, int64_t& b, int64_t& e) {
    b = n * k / nc;
    e = n * (k + 1) / nc;
  };
  if (nc <= 1) {
    fn(0, int64_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(static_cast<size_t>(nc - 1));
  for (int k = 1; k < nc; ++k)
    pool.emplace_back([&, k] {
      int64_t b, e;
      range(k, b, e);
      fn(k, b, e);
    });
  int64_t b, e;
  range(0, b, e);
  fn(0, b, e);
  for (auto& t : pool) t.join();
}

template <class Fn>
void parallel_for(int64_t n, Fn&& fn, int64_t grain = 16384) {
  for_chunks(n, chunks(n, grain), [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) fn(i);
  });
}

}  // namespace rmp
