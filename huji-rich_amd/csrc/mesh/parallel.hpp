// Host threads for the mesh builders and the mesh upload: a blocking
// parallel for over contiguous chunks, run by a pool of worker threads that
// persists for the process (starting threads per call costs far more than
// the shorter loops themselves).  Work is split the same way for a given
// size and thread count, and callers combine per-chunk results in chunk
// order, so results never depend on timing.
#pragma once

#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <deque>
#include <functional>
#include <mutex>
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

namespace detail {

inline bool& in_worker() {
  static thread_local bool v = false;
  return v;
}

// threads() - 1 detached workers taking jobs from one queue.  The pool is
// never destroyed: its workers wait on it until the process ends.
struct Pool {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> jobs;

  explicit Pool(int n) {
    for (int k = 0; k < n; ++k) std::thread([this] { work(); }).detach();
  }
  void work() {
    in_worker() = true;
    for (;;) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> l(mu);
        cv.wait(l, [&] { return !jobs.empty(); });
        job = std::move(jobs.front());
        jobs.pop_front();
      }
      job();
    }
  }
  void submit(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> l(mu);
      jobs.push_back(std::move(f));
    }
    cv.notify_one();
  }
};

// The process's pool; a forked child, which has none of its parent's
// threads, starts its own.
inline Pool& pool() {
  static std::mutex m;
  static Pool* p = nullptr;
  static pid_t owner = 0;
  std::lock_guard<std::mutex> l(m);
  if (!p || owner != getpid()) {
    p = new Pool(threads() - 1);
    owner = getpid();
  }
  return *p;
}

}  // namespace detail

// fn(chunk, begin, end) for `nc` contiguous chunks of [0, n); chunk 0 runs
// on the calling thread, the others on the pool.  Called from inside a
// chunk (a pool worker), or with one thread, it runs its chunks in order on
// the calling thread.
template <class Fn>
void for_chunks(int64_t n, int nc, Fn&& fn) {
  auto range = [&](int k, int64_t& b, int64_t& e) {
    b = n * k / nc;
    e = n * (k + 1) / nc;
  };
  int64_t b, e;
  if (nc <= 1) {
    fn(0, int64_t{0}, n);
    return;
  }
  if (threads() <= 1 || detail::in_worker()) {  // no pool workers to wait for
    for (int k = 0; k < nc; ++k) {
      range(k, b, e);
      fn(k, b, e);
    }
    return;
  }
  std::mutex dm;
  std::condition_variable dcv;
  int left = nc - 1;
  detail::Pool& p = detail::pool();
  for (int k = 1; k < nc; ++k)
    p.submit([&, k] {
      int64_t kb, ke;
      range(k, kb, ke);
      fn(k, kb, ke);
      std::lock_guard<std::mutex> l(dm);  // notified under the lock: the waiter
      if (--left == 0) dcv.notify_one();  // cannot return before it is released
    });
  range(0, b, e);
  fn(0, b, e);
  std::unique_lock<std::mutex> l(dm);
  dcv.wait(l, [&] { return left == 0; });
}

template <class Fn>
void parallel_for(int64_t n, Fn&& fn, int64_t grain = 16384) {
  for_chunks(n, chunks(n, grain), [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) fn(i);
  });
}

}  // namespace rmp
