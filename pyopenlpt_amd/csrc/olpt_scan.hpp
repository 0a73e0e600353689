// olpt_scan.hpp - exclusive prefix sums over int32 arrays (device-resident
// sizes), used to turn per-chunk / per-cell / per-candidate counts into
// output offsets without a host round trip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace olpt {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr int kScanTile = kScanThreads * kScanPerThread; // 4096

// Scratch needed by scan_exclusive for n elements and `segments` segments.
inline size_t scan_scratch_ints(long n, int segments) {
  long tiles = (n + kScanTile - 1) / kScanTile;
  return (size_t)(tiles + 1) * segments;
}

// Exclusive scan of `segments` arrays of length n (stride `seg_stride`):
//   out[s][i] = sum_{k<i} in[s][k];  total[s] = sum of all.
// in and out may alias.  total may be null.  Over a batch of nf frames every
// pointer is frame 0's and frame f's lies f * fstride bytes further.  With
// n_dev (a device count per frame, frame 0's pointer like the others) only
// the first min(n, *n_dev) entries are read and written: the rest of out is
// left as it was.
void scan_exclusive(const int32_t *in, int32_t *out, long n, int segments,
                    long seg_stride, int32_t *total, int32_t *scratch,
                    hipStream_t st, int nf = 1, long fstride = 0,
                    const int32_t *n_dev = nullptr);

// One segment, and in[] is left all zero (in and out must not alias): the
// counts of a count -> scan -> fill pass are ready for the next pass.
// total may be null; pointing it at out + n makes out[i + 1] - out[i] the
// count of entry i for every i < n.
void scan_exclusive_clear(int32_t *in, int32_t *out, long n, int32_t *total, int32_t *scratch,
                          hipStream_t st, int nf = 1, long fstride = 0);

} // namespace olpt
