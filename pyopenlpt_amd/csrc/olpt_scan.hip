// olpt_scan.hip - two-launch exclusive scan: tile sums, then the tile-local
// scan, whose workgroup adds up the sums of the tiles before its own for the
// carry.  Each tile is 4096 ints handled by one 256-thread workgroup (16
// consecutive ints per thread, wave-shuffle scan of the per-thread partials),
// so a 4M-entry array is 1024 tiles.
#include "olpt_scan.hpp"
#include "olpt_kernels.hpp"

namespace olpt {

// Workgroup exclusive scan: a shuffle scan inside each wave64, then every
// thread adds the totals of the waves before its own (kScanThreads / 64
// entries in LDS).  lds needs kScanThreads / 64 ints.
constexpr int kScanWaves = kScanThreads / 64;
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *lds,
                                                   int32_t &block_total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  int32_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanWaves; ++i) {
    const int32_t s = lds[i];
    pre += i < w ? s : 0;
    tot += s;
  }
  block_total = tot;
  __syncthreads();  // lds may be reused by the next call
  return pre + x - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_tile_sums(const int32_t *__restrict__ in, long n, long stride,
            int32_t *__restrict__ tsum, long tiles, long fstride, const int32_t *n_dev) {
  __shared__ int32_t lds[kScanWaves];
  const long fo = (long)blockIdx.z * fstride;
  in = fshift(in, fo);
  tsum = fshift(tsum, fo);
  const int seg = blockIdx.y;
  const long tile = blockIdx.x;
  if (n_dev) n = min(n, (long)max(0, *fshift(n_dev, fo)));
  if (tile * kScanTile >= n) {
    // past the frame's count: an empty tile
    if (threadIdx.x == 0) tsum[seg * (tiles + 1) + tile] = 0;
    return;
  }
  const int32_t *p = in + seg * stride;
  // a sum: lanes read consecutive entries (coalesced)
  const long base = tile * kScanTile + threadIdx.x;
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    const long i = base + (long)k * kScanThreads;
    if (i < n) s += p[i];
  }
  int32_t tot;
  block_excl_scan(s, lds, tot);
  if (threadIdx.x == 0) tsum[seg * (tiles + 1) + tile] = tot;
}

// CLEAR: the input is zeroed once read (in and out distinct), so a count
// array is ready for its next round of atomic increments without a launch
// of its own.  The tile goes through LDS: loaded and stored with lanes on
// consecutive entries (coalesced), scanned with each thread on 16
// consecutive ones; entry e sits at e + e / 16, so a thread's run starts 17
// words after the previous thread's and the 64 lanes of a wave hit 64
// different banks.
constexpr int kScanPad = kScanTile + kScanTile / kScanPerThread;
__device__ __forceinline__ int scan_slot(int e) { return e + e / kScanPerThread; }

template <bool CLEAR>
__global__ void __launch_bounds__(kScanThreads)
k_tile_scan(int32_t *__restrict__ in, int32_t *__restrict__ out, long n,
            long stride, const int32_t *__restrict__ tsum, long tiles, long fstride,
            const int32_t *n_dev, int32_t *total) {
  __shared__ int32_t lds[kScanWaves];
  const long fo = (long)blockIdx.z * fstride;
  in = fshift(in, fo);
  out = fshift(out, fo);
  tsum = fshift(tsum, fo);
  const int seg = blockIdx.y;
  const long tile = blockIdx.x;
  if (n_dev) n = min(n, (long)max(0, *fshift(n_dev, fo)));
  const bool past = tile * kScanTile >= n;  // uniform over the workgroup
  if (past && tile != tiles - 1) return;
  // the sums of the tiles before this one, added up here (a few hundred per
  // segment at most) rather than scanned by a launch of their own; the last
  // tile also writes the segment's total
  const int32_t *ts = tsum + seg * (tiles + 1);
  int32_t carry = 0;
  for (long b0 = 0; b0 < tile; b0 += kScanThreads) {
    const long j = b0 + threadIdx.x;
    int32_t tot;
    block_excl_scan(j < tile ? ts[j] : 0, lds, tot);
    carry += tot;
  }
  if (tile == tiles - 1 && total && threadIdx.x == 0) fshift(total, fo)[seg] = carry + ts[tile];
  if (past) return;
  __shared__ int32_t buf[kScanPad];
  int32_t *p = in + seg * stride;
  int32_t *q = out + seg * stride;
  const long t0 = tile * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    const int e = k * kScanThreads + (int)threadIdx.x;
    const long i = t0 + e;
    buf[scan_slot(e)] = i < n ? p[i] : 0;
    if constexpr (CLEAR) {
      if (i < n) p[i] = 0;
    }
  }
  __syncthreads();
  const int e0 = (int)threadIdx.x * kScanPerThread;
  int32_t v[kScanPerThread];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    v[k] = buf[scan_slot(e0 + k)];
    s += v[k];
  }
  int32_t tot;
  int32_t run = block_excl_scan(s, lds, tot) + carry;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    buf[scan_slot(e0 + k)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    const int e = k * kScanThreads + (int)threadIdx.x;
    const long i = t0 + e;
    if (i < n) q[i] = buf[scan_slot(e)];
  }
}

void scan_exclusive(const int32_t *in, int32_t *out, long n, int segments,
                    long seg_stride, int32_t *total, int32_t *scratch,
                    hipStream_t st, int nf, long fstride, const int32_t *n_dev) {
  if (n <= 0 || segments <= 0 || nf <= 0) return;
  const long tiles = (n + kScanTile - 1) / kScanTile;
  dim3 g((unsigned)tiles, (unsigned)segments, (unsigned)nf);
  k_tile_sums<<<g, kScanThreads, 0, st>>>(in, n, seg_stride, scratch, tiles, fstride, n_dev);
  k_tile_scan<false><<<g, kScanThreads, 0, st>>>(const_cast<int32_t *>(in), out, n, seg_stride,
                                                 scratch, tiles, fstride, n_dev, total);
}

void scan_exclusive_clear(int32_t *in, int32_t *out, long n, int32_t *total, int32_t *scratch,
                          hipStream_t st, int nf, long fstride) {
  if (n <= 0 || nf <= 0) return;
  const long tiles = (n + kScanTile - 1) / kScanTile;
  dim3 g((unsigned)tiles, 1, (unsigned)nf);
  k_tile_sums<<<g, kScanThreads, 0, st>>>(in, n, n, scratch, tiles, fstride, nullptr);
  k_tile_scan<true><<<g, kScanThreads, 0, st>>>(in, out, n, n, scratch, tiles, fstride, nullptr,
                                                total);
}

} // namespace olpt
