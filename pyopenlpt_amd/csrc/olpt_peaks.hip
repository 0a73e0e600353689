// olpt_peaks.hip - 2D tracer detection (ObjectFinder2D::findTracer2D,
// ObjectFinder.cpp:37-111).
//
// Layout: a "chunk" is 64 consecutive interior columns of one interior row;
// one wave64 takes up to 16 chunks of a row.  Pass 1 evaluates threshold +
// 4-neighbour local maximum
// (myMATH::isLocalMax tie-break: up/left may equal, down/right must be
// strictly lower, myMATH.h:398-439) with coalesced loads and writes the
// 64-bit ballot mask and its popcount; the predicate includes the fit's own
// rejections so counts and order match the reference exactly.  An exclusive
// scan of the per-chunk counts gives each chunk its output offset, so pass 2
// writes the peaks in the reference's row-major scan order with no atomics,
// fitting the separable three-point log-parabola in FP64 (the reference's
// absolute-coordinate formula, :67-97).
#include <algorithm>
#include <mutex>
#include <string>

#include "olpt_kernels.hpp"
#include "olpt_scan.hpp"
#include "olpt_shuffle.hpp"

namespace olpt {

constexpr int kPeakWaves = 4; // waves per workgroup

template <typename TIn>
__device__ __forceinline__ double pix(const TIn *p, long i) {
  return (double)p[i];
}

__device__ __forceinline__ double safe_ln(double v) {
  return log(v < 1e-4 ? 1e-4 : v);
}

// false where the reference rejects the peak (zero denominator or
// non-finite centre, ObjectFinder.cpp:81-83, :95-97)
template <typename TIn>
__device__ __forceinline__ bool fit_peak(const TIn *p, long i, int W, int row,
                                         int col, double &xc, double &yc) {
  const int x1 = col - 1, x2 = col, x3 = col + 1;
  const int y1 = row - 1, y2 = row, y3 = row + 1;
  const double l1x = safe_ln(pix(p, i - 1));
  const double l2 = safe_ln(pix(p, i));
  const double l3x = safe_ln(pix(p, i + 1));
  const double nx = l1x * (double)((x2 * x2) - (x3 * x3)) -
                    l2 * (double)((x1 * x1) - (x3 * x3)) +
                    l3x * (double)((x1 * x1) - (x2 * x2));
  const double dx = l1x * (double)(x3 - x2) - l3x * (double)(x1 - x2) +
                    l2 * (double)(x1 - x3);
  if (dx == 0.0) return false;
  xc = -0.5 * (nx / dx);
  if (!isfinite(xc)) return false;
  const double l1y = safe_ln(pix(p, i - W));
  const double l3y = safe_ln(pix(p, i + W));
  const double ny = l1y * (double)((y2 * y2) - (y3 * y3)) -
                    l2 * (double)((y1 * y1) - (y3 * y3)) +
                    l3y * (double)((y1 * y1) - (y2 * y2));
  const double dy = l1y * (double)(y3 - y2) - l3y * (double)(y1 - y2) +
                    l2 * (double)(y1 - y3);
  if (dy == 0.0) return false;
  yc = -0.5 * (ny / dy);
  return isfinite(yc);
}

// A wave takes a group of up to kPeakGroup consecutive chunks of one row.
// Candidates (threshold + local maximum) are found a chunk at a time with
// coalesced loads; the separable fit - five FP64 logarithms and two
// divisions, needed for its rejections and for the centre - then runs on
// the group's candidates packed densely into lanes, so a wave fits ~64
// candidates per pass instead of the two or three a single chunk holds.
constexpr int kPeakGroup = 16;

// index of the r-th (0-based) set bit of m (r < popcount(m))
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const uint64_t lo = m & ((1ull << w) - 1ull);
    const int c = __popcll(lo);
    if (r >= c) {
      r -= c;
      m >>= w;
      pos += w;
    } else {
      m = lo;
    }
  }
  return pos;
}

// the (chunk k, bit) of candidate t of a group with masks cm and exclusive
// prefix counts pre (wave-uniform)
__device__ __forceinline__ void group_locate(const uint64_t cm[kPeakGroup],
                                             const int pre[kPeakGroup + 1], int g, int t,
                                             int &k, int &bit) {
  k = 0;
#pragma unroll
  for (int q = 1; q < kPeakGroup; ++q)
    if (q < g && pre[q] <= t) k = q;
  uint64_t m = 0;
  int base = 0;
#pragma unroll
  for (int q = 0; q < kPeakGroup; ++q)
    if (q == k) {
      m = cm[q];
      base = pre[q];
    }
  bit = nth_set_bit(m, t - base);
}

template <typename TIn>
__global__ void __launch_bounds__(kPeakWaves * kWave)
k_peak_mask(const TIn *__restrict__ img, long cam_stride, int H, int W,
            double min_int, CamList cams, uint64_t *__restrict__ mask,
            int32_t *__restrict__ cnt, long nchunk, long fstride) {
  __shared__ unsigned long long s_vm[kPeakWaves][kPeakGroup];
  const long fo = (long)blockIdx.z * fstride;
  img = fshift(img, fo);
  mask = fshift(mask, fo);
  cnt = fshift(cnt, fo);
  const int s = blockIdx.y;
  const int wv = threadIdx.x / kWave;
  const int lane = lane_id();
  const int cpr = (W - 2 + 63) / 64;
  const int gpr = (cpr + kPeakGroup - 1) / kPeakGroup;
  const long grp = (long)blockIdx.x * kPeakWaves + wv;
  if (grp >= (long)(H - 2) * gpr) return;
  const int row = 1 + (int)(grp / gpr);
  const int c0 = (int)(grp % gpr) * kPeakGroup;
  const int g = min(kPeakGroup, cpr - c0);
  const TIn *p = img + cams.id[s] * cam_stride;
  uint64_t cm[kPeakGroup];
  int pre[kPeakGroup + 1];
  pre[0] = 0;
#pragma unroll
  for (int k = 0; k < kPeakGroup; ++k) {
    bool cand = false;
    const int col = 1 + (c0 + k) * 64 + lane;
    if (k < g && col < W - 1) {
      const long i = (long)row * W + col;
      const double v = pix(p, i);
      if (!(v < min_int)) {
        const double up = pix(p, i - W), lf = pix(p, i - 1);
        const double dn = pix(p, i + W), rt = pix(p, i + 1);
        cand = !(up > v) && !(lf > v) && !(dn >= v) && !(rt >= v);
      }
    }
    cm[k] = ballot(cand);
    pre[k + 1] = pre[k] + __popcll(cm[k]);
  }
  if (lane < kPeakGroup) s_vm[wv][lane] = 0ull;
  __builtin_amdgcn_wave_barrier();
  // the fit's rejections (ObjectFinder.cpp:81-83, :95-97) on the packed
  // candidates; the survivors' bits go back into their chunk masks
  const int total = pre[kPeakGroup];
  for (int t0 = 0; t0 < total; t0 += kWave) {
    const int t = t0 + lane;
    if (t < total) {
      int k, bit;
      group_locate(cm, pre, g, t, k, bit);
      const int col = 1 + (c0 + k) * 64 + bit;
      double xc, yc;
      if (fit_peak(p, (long)row * W + col, W, row, col, xc, yc))
        atomicOr(&s_vm[wv][k], 1ull << bit);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < g) {
    const uint64_t m = s_vm[wv][lane];
    const long chunk = (long)(row - 1) * cpr + c0 + lane;
    mask[s * nchunk + chunk] = m;
    cnt[s * nchunk + chunk] = __popcll(m);
  }
}

template <typename TIn>
__global__ void __launch_bounds__(kPeakWaves * kWave)
k_peak_emit(const TIn *__restrict__ img, long cam_stride, int H, int W,
            CamList cams, const uint64_t *__restrict__ mask,
            const int32_t *__restrict__ off, long nchunk,
            double *__restrict__ xy, int cap, long fstride) {
  const long fo = (long)blockIdx.z * fstride;
  img = fshift(img, fo);
  mask = fshift(mask, fo);
  off = fshift(off, fo);
  xy = fshift(xy, fo);
  const int s = blockIdx.y;
  const int wv = threadIdx.x / kWave;
  const int lane = lane_id();
  const int cpr = (W - 2 + 63) / 64;
  const int gpr = (cpr + kPeakGroup - 1) / kPeakGroup;
  const long grp = (long)blockIdx.x * kPeakWaves + wv;
  if (grp >= (long)(H - 2) * gpr) return;
  const int row = 1 + (int)(grp / gpr);
  const int c0 = (int)(grp % gpr) * kPeakGroup;
  const int g = min(kPeakGroup, cpr - c0);
  const long chunk0 = s * nchunk + (long)(row - 1) * cpr + c0;
  uint64_t cm[kPeakGroup];
  int pre[kPeakGroup + 1];
  pre[0] = 0;
#pragma unroll
  for (int k = 0; k < kPeakGroup; ++k) {
    cm[k] = k < g ? mask[chunk0 + k] : 0ull;
    pre[k + 1] = pre[k] + __popcll(cm[k]);
  }
  const int total = pre[kPeakGroup];
  if (total == 0) return;
  const int cam = cams.id[s];
  const TIn *p = img + cam * cam_stride;
  for (int t0 = 0; t0 < total; t0 += kWave) {
    const int t = t0 + lane;
    if (t >= total) continue;
    int k, bit;
    group_locate(cm, pre, g, t, k, bit);
    // rank inside the chunk = t - pre[k]: the reference's row-major order
    int base = 0;
#pragma unroll
    for (int q = 0; q < kPeakGroup; ++q)
      if (q == k) base = pre[q];
    const int pos = off[chunk0 + k] + (t - base);
    if (pos >= cap) continue;
    const int col = 1 + (c0 + k) * 64 + bit;
    double xc = 0.0, yc = 0.0;
    fit_peak(p, (long)row * W + col, W, row, col, xc, yc);
    double *o = xy + ((long)cam * cap + pos) * 2;
    o[0] = xc;
    o[1] = yc;
  }
}

// Per-camera counts and capacity check.
__global__ void k_peak_finish(CamList cams, const int32_t *__restrict__ totals,
                              int cap, int32_t *__restrict__ npts,
                              int32_t *__restrict__ overflow, long fstride) {
  const long fo = (long)blockIdx.z * fstride;
  totals = fshift(totals, fo);
  npts = fshift(npts, fo);
  overflow = fshift(overflow, fo);
  const int s = threadIdx.x;
  if (s >= cams.n) return;
  const int t = totals[s];
  npts[cams.id[s]] = t < cap ? t : cap;
  if (t > cap) atomicOr(&overflow[0], 1);
}

template <typename TIn>
void launch_find_peaks(const TIn *img, long cam_stride, int H, int W,
                       double min_int, CamList cams, uint64_t *mask,
                       int32_t *cnt, int32_t *off, int32_t *scan_scratch,
                       int32_t *totals, double *xy, int cap, int32_t *npts,
                       int32_t *overflow, hipStream_t st, int nf, long fstride) {
  const long nchunk = peak_chunks(H, W);
  if (nchunk <= 0 || cams.n <= 0 || nf <= 0) return;
  const int cpr = (W - 2 + 63) / 64;
  const long ngroup = (long)(H - 2) * ((cpr + kPeakGroup - 1) / kPeakGroup);
  dim3 g((unsigned)((ngroup + kPeakWaves - 1) / kPeakWaves), (unsigned)cams.n, (unsigned)nf);
  stage_begin(ST_PEAKS, st);
  k_peak_mask<TIn><<<g, kPeakWaves * kWave, 0, st>>>(img, cam_stride, H, W,
                                                     min_int, cams, mask, cnt,
                                                     nchunk, fstride);
  scan_exclusive(cnt, off, nchunk, cams.n, nchunk, totals, scan_scratch, st, nf, fstride);
  k_peak_emit<TIn><<<g, kPeakWaves * kWave, 0, st>>>(img, cam_stride, H, W, cams,
                                                     mask, off, nchunk, xy, cap, fstride);
  k_peak_finish<<<dim3(1, 1, nf), 64, 0, st>>>(cams, totals, cap, npts, overflow, fstride);
  stage_end(ST_PEAKS, st);
}

// limit2DObjectNumber (IPR.cpp:19-39): the listed cameras' peak lists
// longer than cap are shuffled with one engine seeded 1234, shared across
// the cameras in order, and cut to cap.  The draws are sequential, so one
// thread computes each camera's permutation (it only runs when a list is
// over the cap); a grid gathers the kept points.
__global__ void k_limit_perm(const int32_t *npts, CamList cams, int cap, int cap2d,
                             int32_t *perm, int32_t *lim, long fstride) {
  if (threadIdx.x || blockIdx.x) return;
  const long fo = (long)blockIdx.z * fstride;
  npts = fshift(npts, fo);
  perm = fshift(perm, fo);
  lim = fshift(lim, fo);
  Minstd0 g{1234};
  for (int q = 0; q < cams.n; ++q) {
    const int k = cams.id[q];
    const int n = npts[k];
    lim[k] = 0;
    if (n <= cap) continue;
    int32_t *a = perm + (long)k * cap2d;
    for (int i = 0; i < n; ++i) a[i] = i;
    shuffle_like_libstdcxx(a, (int64_t)n, g);
    lim[k] = 1;
  }
}

__global__ void k_limit_gather(const double *__restrict__ xy, const int32_t *__restrict__ perm,
                               const int32_t *__restrict__ lim, CamList cams, int cap, int cap2d,
                               double *__restrict__ tmp, long fstride) {
  const long fo = (long)blockIdx.z * fstride;
  xy = fshift(xy, fo);
  perm = fshift(perm, fo);
  lim = fshift(lim, fo);
  tmp = fshift(tmp, fo);
  const int k = cams.id[blockIdx.y];
  if (!lim[k]) return;
  const long o = (long)k * cap2d;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
    const long src = o + perm[o + i];
    tmp[2 * (o + i)] = xy[2 * src];
    tmp[2 * (o + i) + 1] = xy[2 * src + 1];
  }
}

__global__ void k_limit_store(double *__restrict__ xy, int32_t *__restrict__ npts,
                              const int32_t *__restrict__ lim, CamList cams, int cap, int cap2d,
                              const double *__restrict__ tmp, long fstride) {
  const long fo = (long)blockIdx.z * fstride;
  xy = fshift(xy, fo);
  npts = fshift(npts, fo);
  lim = fshift(lim, fo);
  tmp = fshift(tmp, fo);
  const int k = cams.id[blockIdx.y];
  if (!lim[k]) return;
  const long o = (long)k * cap2d;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
    xy[2 * (o + i)] = tmp[2 * (o + i)];
    xy[2 * (o + i) + 1] = tmp[2 * (o + i) + 1];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) npts[k] = cap;
}

void launch_limit2d(double *xy, int32_t *npts, CamList cams, int cap, int cap2d,
                    int32_t *perm, int32_t *lim, double *tmp, hipStream_t st, int nf,
                    long fstride) {
  if (cams.n <= 0 || cap <= 0 || cap >= cap2d || nf <= 0) return;
  k_limit_perm<<<dim3(1, 1, nf), 64, 0, st>>>(npts, cams, cap, cap2d, perm, lim, fstride);
  dim3 g((unsigned)std::min(1024, (cap + 255) / 256), (unsigned)cams.n, (unsigned)nf);
  k_limit_gather<<<g, 256, 0, st>>>(xy, perm, lim, cams, cap, cap2d, tmp, fstride);
  k_limit_store<<<g, 256, 0, st>>>(xy, npts, lim, cams, cap, cap2d, tmp, fstride);
}

#define INST(T)                                                                \
  template void launch_find_peaks<T>(const T *, long, int, int, double,        \
                                     CamList, uint64_t *, int32_t *,           \
                                     int32_t *, int32_t *, int32_t *,          \
                                     double *, int, int32_t *, int32_t *,      \
                                     hipStream_t, int, long);
INST(uint8_t)
INST(uint16_t)
INST(float)
INST(double)
#undef INST

} // namespace olpt

// ===========================================================================
// olpt_find_peaks_static: the peak kernels on a fixed per-device work array
// (no context, no allocation).
// ===========================================================================
namespace {

using namespace olpt;

constexpr int kStaticMaxSide = 4096;
constexpr long kStaticMaxChunks = (long)(kStaticMaxSide - 2) * ((kStaticMaxSide - 2 + 63) / 64);

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// layout: mask u64 | cnt i32 | off i32 | scan partials | totals[64] |
// npts[OLPT_MAX_CAM] | flags[8]
constexpr size_t kOffMask = 0;
constexpr size_t kOffCnt = kOffMask + align256(8 * (size_t)kStaticMaxChunks);
constexpr size_t kOffOff = kOffCnt + align256(4 * (size_t)kStaticMaxChunks);
constexpr size_t kOffScan = kOffOff + align256(4 * (size_t)kStaticMaxChunks);
constexpr size_t kScanInts = ((size_t)kStaticMaxChunks + kScanTile - 1) / kScanTile + 1;
constexpr size_t kOffTotals = kOffScan + align256(4 * kScanInts);
constexpr size_t kOffNpts = kOffTotals + align256(4 * 64);
constexpr size_t kOffFlags = kOffNpts + align256(4 * OLPT_MAX_CAM);
constexpr size_t kStaticBytes = kOffFlags + align256(4 * 8);   // ~4.2 MB

// The work memory of olpt_find_peaks_static: one copy per GPU, held until
// the process ends.  Host code uses only the device address returned by
// hipGetSymbolAddress (per device), never this name.
__device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

__global__ void k_static_clear(int32_t *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void k_static_count(const int32_t *__restrict__ totals, int32_t *__restrict__ n_out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *n_out = totals[0];
}

thread_local std::string g_static_err;

olpt_status static_fail(olpt_status s, const std::string &m) {
  g_static_err = m;
  return olpt::set_error(s, g_static_err.c_str());
}

} // namespace

extern "C" olpt_status olpt_find_peaks_static(void *stream, const void *d_img, int dtype, int n_row,
                                              int n_col, const olpt_ipr_params *params, double *d_xy,
                                              int cap, int32_t *d_n_out) {
  if (!d_img || !params || !d_n_out || (cap > 0 && !d_xy))
    return static_fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (n_row < 3 || n_col < 3 || cap < 0)
    return static_fail(OLPT_E_INVALID_ARGUMENT, "image must be at least 3 x 3 and cap >= 0");
  if (n_row > kStaticMaxSide || n_col > kStaticMaxSide)
    return static_fail(OLPT_E_CAPACITY, "olpt_find_peaks_static takes images up to 4096 x 4096");
  if (dtype < OLPT_U8 || dtype > OLPT_F64) return static_fail(OLPT_E_UNSUPPORTED, "unsupported image dtype");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return static_fail(OLPT_E_GPU, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return static_fail(OLPT_E_GPU, "device index out of range");
  hipStream_t st = (hipStream_t)stream;
  if (st) {
    int sdev = -1;
    e = hipStreamGetDevice(st, &sdev);
    if (e != hipSuccess) return static_fail(OLPT_E_GPU, std::string("hipStreamGetDevice: ") + hipGetErrorString(e));
    if (sdev != dev) return static_fail(OLPT_E_INVALID_ARGUMENT, "the stream belongs to another device than the current one");
  }
  unsigned char *base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return static_fail(OLPT_E_GPU, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      g_scratch_ptr[dev] = p;
    }
    base = (unsigned char *)g_scratch_ptr[dev];
  }
  uint64_t *mask = (uint64_t *)(base + kOffMask);
  int32_t *cnt = (int32_t *)(base + kOffCnt);
  int32_t *off = (int32_t *)(base + kOffOff);
  int32_t *scan = (int32_t *)(base + kOffScan);
  int32_t *totals = (int32_t *)(base + kOffTotals);
  int32_t *npts = (int32_t *)(base + kOffNpts);
  int32_t *flags = (int32_t *)(base + kOffFlags);
  // the overflow flag is or-ed into: clear it (a kernel, so the call can be
  // captured into a graph); every other word is written before it is read
  k_static_clear<<<1, 64, 0, st>>>(flags, 8);
  CamList cl{1, {0}};
  const double mi = (double)params->min_obj_int;
  double *xy = d_xy;
  const int c = cap;
  switch (dtype) {
  case OLPT_U8: launch_find_peaks<uint8_t>((const uint8_t *)d_img, 0, n_row, n_col, mi, cl, mask, cnt, off, scan, totals, xy, c, npts, flags, st, 1, 0); break;
  case OLPT_U16: launch_find_peaks<uint16_t>((const uint16_t *)d_img, 0, n_row, n_col, mi, cl, mask, cnt, off, scan, totals, xy, c, npts, flags, st, 1, 0); break;
  case OLPT_F32: launch_find_peaks<float>((const float *)d_img, 0, n_row, n_col, mi, cl, mask, cnt, off, scan, totals, xy, c, npts, flags, st, 1, 0); break;
  default: launch_find_peaks<double>((const double *)d_img, 0, n_row, n_col, mi, cl, mask, cnt, off, scan, totals, xy, c, npts, flags, st, 1, 0); break;
  }
  k_static_count<<<1, 64, 0, st>>>(totals, d_n_out);
  e = hipGetLastError();
  if (e != hipSuccess) return static_fail(OLPT_E_GPU, std::string("launch: ") + hipGetErrorString(e));
  return OLPT_OK;
}
