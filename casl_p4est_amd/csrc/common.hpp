// Shared host-side helpers: error reporting, thread pool loops, Morton keys.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <string>
#include <exception>
#include <thread>
#include <vector>
#include <algorithm>

namespace casl {

// Thread-local last error message returned by casl_last_error().
void set_error(const std::string &msg);
const char *get_error();

// Worker count for host loops: CASL_NUM_THREADS, else OMP_NUM_THREADS, else
// hardware concurrency, capped at 64.
int num_threads();

// Runs fn(begin, end) over [0, n) split into contiguous chunks, one per thread.
template <class Fn>
void parallel_for(int64_t n, Fn fn, int64_t min_chunk = 4096) {
  int t = num_threads();
  int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(t, n / std::max<int64_t>(1, min_chunk)));
  if (chunks <= 1) {
    if (n > 0) fn(int64_t(0), n);
    return;
  }
  // an exception in a worker is kept and rethrown here after every worker
  // has joined (one escaping a thread would end the process)
  std::vector<std::exception_ptr> err(size_t(chunks), nullptr);
  std::vector<std::thread> pool;
  pool.reserve(chunks);
  for (int64_t c = 0; c < chunks; ++c) {
    int64_t b = n * c / chunks, e = n * (c + 1) / chunks;
    pool.emplace_back([=, &fn, &err]() {
      try {
        fn(b, e);
      } catch (...) {
        err[size_t(c)] = std::current_exception();
      }
    });
  }
  for (auto &th : pool) th.join();
  for (auto &x : err)
    if (x) std::rethrow_exception(x);
}

// Morton (z-order) keys.  Bit b of axis a lands at bit D*b + a, so at equal
// bit position z outranks y outranks x: the order of p4est_quadrant_compare
// on non-negative coordinates (p4est 2.8, p4est_bits.c).
inline uint64_t spread2(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
inline uint32_t compact2(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return uint32_t(x);
}
inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1FFFFF;
  x = (x | (x << 32)) & 0x001F00000000FFFFull;
  x = (x | (x << 16)) & 0x001F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
inline uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x | (x >> 2)) & 0x10C30C30C30C30C3ull;
  x = (x | (x >> 4)) & 0x100F00F00F00F00Full;
  x = (x | (x >> 8)) & 0x001F0000FF0000FFull;
  x = (x | (x >> 16)) & 0x001F00000000FFFFull;
  x = (x | (x >> 32)) & 0x00000000001FFFFFull;
  return uint32_t(x);
}

template <int D>
inline uint64_t morton(const int32_t *c) {
  if constexpr (D == 2)
    return spread2(uint32_t(c[0])) | (spread2(uint32_t(c[1])) << 1);
  else
    return spread3(uint32_t(c[0])) | (spread3(uint32_t(c[1])) << 1) |
           (spread3(uint32_t(c[2])) << 2);
}

template <int D>
inline void demorton(uint64_t k, int32_t *c) {
  if constexpr (D == 2) {
    c[0] = int32_t(compact2(k));
    c[1] = int32_t(compact2(k >> 1));
  } else {
    c[0] = int32_t(compact3(k));
    c[1] = int32_t(compact3(k >> 1));
    c[2] = int32_t(compact3(k >> 2));
  }
}

}  // namespace casl
