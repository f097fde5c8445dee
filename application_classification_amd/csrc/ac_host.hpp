// Host-side helpers shared by the C-ABI implementation, the generators and the file readers.
#pragma once

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

namespace ac {

// Thread-local message for calls that have no context (generators, file IO, ac_create).
void set_global_error(const std::string& msg);
const char* global_error();

// Worker threads for host loops: the CPUs this process may run on (its affinity mask, not the machine's count),
// further limited by OMP_NUM_THREADS when that is set, at most 64.
inline int host_threads() {
  long n = 0;
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = (long)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const long v = std::atol(e);
    if (v > 0 && v < n) n = v;
  }
  if (n < 1) n = 1;
  if (n > 64) n = 64;
  return (int)n;
}

// Static-chunked parallel loop over [0, n): fn(begin, end, worker_id).  The calling thread takes chunk 0; a chunk
// whose thread cannot be started runs on the calling thread instead, so a failed thread start costs time, not the
// process.
template <typename Fn>
void parallel_chunks(int64_t n, Fn fn, int threads = 0) {
  if (threads <= 0) threads = host_threads();
  if (n < 1) return;
  if (threads == 1 || n < 4096) {
    fn((int64_t)0, n, 0);
    return;
  }
  int64_t per = (n + threads - 1) / threads;
  std::vector<std::thread> pool;
  pool.reserve(threads);
  std::vector<int> inline_chunks;
  for (int t = 1; t < threads; ++t) {
    int64_t b = (int64_t)t * per, e = std::min(n, b + per);
    if (b >= e) break;
    try {
      pool.emplace_back([=] { fn(b, e, t); });
    } catch (...) {
      inline_chunks.push_back(t);
    }
  }
  fn((int64_t)0, std::min(n, per), 0);
  for (int t : inline_chunks) fn((int64_t)t * per, std::min(n, (int64_t)t * per + per), t);
  for (auto& th : pool) th.join();
}

// Dynamic work queue over [0, n) items: fn(item).  The calling thread takes items too, so the queue drains even
// when no thread can be started.
template <typename Fn>
void parallel_items(int64_t n, Fn fn, int threads = 0) {
  if (threads <= 0) threads = host_threads();
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (;;) {
      const int64_t i = next.fetch_add(1);
      if (i >= n) break;
      fn(i);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads && t < n; ++t) {
    try {
      pool.emplace_back(work);
    } catch (...) {
      break;
    }
  }
  work();
  for (auto& th : pool) th.join();
}

// Counter-based random numbers: a SplitMix64 finaliser applied to (seed, stream, index).
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull)) ^ i);
}
// Uniform in the open interval (0, 1) with 53 random bits.
inline double unit_open(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// Sequential generator for the few places that need a stream of draws (pattern growth).
struct SeqRng {
  uint64_t seed, stream, ctr = 0;
  SeqRng(uint64_t s, uint64_t st) : seed(s), stream(st) {}
  uint64_t next() { return hash3(seed, stream, ctr++); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  double uniform() { return unit_open(next()); }
};

// Stream ids, so that independent draws never share counters.
enum : uint64_t {
  kStreamEr = 1,
  kStreamRmat = 2,
  kStreamPattern = 3,
};

}  // namespace ac
