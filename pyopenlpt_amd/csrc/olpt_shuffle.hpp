// olpt_shuffle.hpp - the random subsampling of limit2DObjectNumber
// (IPR.cpp:19-39): std::shuffle with std::default_random_engine(1234), as
// the reference's toolchain (GCC 11's libstdc++) defines them, restated so
// that one device thread can run it (host and device share this code).
//
//  * default_random_engine is minstd_rand0: x <- 16807 x mod (2^31 - 1),
//    x0 = seed, each call returns the new x (range [1, 2^31 - 2]).
//  * uniform_int_distribution<unsigned long>{0, b} on it: the engine's range
//    R = 2^31 - 3 exceeds b, so the value is drawn by downscaling: with
//    s = R / (b + 1) (integer division) and past = (b + 1) s, draw
//    r = x - 1 until r < past and return r / s (b == R returns x - 1).
//  * shuffle of n elements: when R / n >= n, positions come two at a time:
//    for even n first swap element 1 with a draw from {0, 1}; then for
//    i = first unswapped, one draw x in [0, (i + 1)(i + 2) - 1] swaps i with
//    x / (i + 2) and i + 1 with x % (i + 2).  Otherwise, for i = 1 .. n-1,
//    swap i with a draw from [0, i].
#pragma once

#include <stdint.h>

namespace olpt {

struct Minstd0 {
  uint64_t x;
  __host__ __device__ uint64_t next() {
    x = (x * 16807u) % 2147483647u;
    return x;
  }
};

constexpr uint64_t kMinstd0Range = 2147483645u;  // max() - min()

// uniform_int_distribution{0, b}; b > kMinstd0Range never occurs for index
// ranges below 2^31 and returns b.
__host__ __device__ inline uint64_t uniform_upto(Minstd0 &g, uint64_t b) {
  if (b < kMinstd0Range) {
    const uint64_t er = b + 1;
    const uint64_t scaling = kMinstd0Range / er;
    const uint64_t past = er * scaling;
    uint64_t r;
    do
      r = g.next() - 1;
    while (r >= past);
    return r / scaling;
  }
  if (b == kMinstd0Range) return g.next() - 1;
  return b;
}

// Shuffle a[0, n) in place, as std::shuffle(a, a + n, g).
template <typename I>
__host__ __device__ inline void shuffle_like_libstdcxx(I *a, int64_t n, Minstd0 &g) {
  if (n <= 0) return;
  const uint64_t un = (uint64_t)n;
  auto swp = [&](int64_t i, int64_t j) {
    const I t = a[i];
    a[i] = a[j];
    a[j] = t;
  };
  if (kMinstd0Range / un >= un) {
    int64_t i = 1;
    if (un % 2 == 0) {
      swp(i, (int64_t)uniform_upto(g, 1));
      ++i;
    }
    while (i != n) {
      const uint64_t sr = (uint64_t)i + 1;
      const uint64_t x = uniform_upto(g, sr * (sr + 1) - 1);
      swp(i, (int64_t)(x / (sr + 1)));
      ++i;
      swp(i, (int64_t)(x % (sr + 1)));
      ++i;
    }
    return;
  }
  for (int64_t i = 1; i < n; ++i) swp(i, (int64_t)uniform_upto(g, (uint64_t)i));
}

} // namespace olpt
