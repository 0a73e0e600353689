// olpt_stb.hip - shake-the-box tracking over the IPR library.
//
// The tracking loop of STB.cpp (initial phase with the predictive field,
// convergence phase with prediction, shake, IPR on the residual and
// linking) driven from the host, with every per-track and per-object
// computation on the GPU:
//   k_predict      predictNext (STB.cpp:233-316), one thread per track axis
//   k_in_view      AxisLimit::check + Object3D::isReconstructable
//                  (STBCommons.h:142-152, ObjectInfo.cpp:3-29)
//   k_repeat       checkRepeat (STB.cpp:698-741), objects x track ends
//   k_link_short   linkShortTrack (STB.cpp:743-872) with brute-force
//                  neighbour and nearest-object searches
//   k_link_field   buildTrackFromPredField's getDisp + findNN
//                  (STB.cpp:160-231, PredField.cpp:91-165)
//   k_lin_fit      checkLinearFit (STB.cpp:874-924)
//   k_pf_node      calPredField per grid node (PredField.cpp:21-86,
//                  :264-371): one wave per node, displacement histogram in
//                  LDS, peak bin and 1D Gaussian peak fit
//   k_pf_smooth    smoothDispField (PredField.cpp:374-485), one pass per
//                  direction
// Prediction shake + residual and IPR use olpt_shake / olpt_ipr_run.  The
// track lists themselves (four ordered lists, as the reference's deques)
// live on the host: per frame they change by order-preserving partitions,
// appends and moves whose order decides later results, and their full
// histories are what the track files hold.
//
// Brute-force searches replace the reference's k-d trees.  Both return
// the nearest object within the radius (strict <); sums over neighbour
// tracks run in track order here and in tree order there, which changes
// the last bits of the neighbour velocity only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <dirent.h>
#include <sys/stat.h>
#include <vector>

#include "olpt_c.h"
#include "olpt_device.hpp"
#include "olpt_kernels.hpp"

using namespace olpt;

namespace {

// STBCommons.h:10,16,27-32
constexpr double kSmallNumber = 1e-8;
constexpr double kLogSmall = 1e-4;
constexpr int kWienerMaxIter = 5000;
constexpr int kUnlinked = -1;
constexpr double kMaxErrLinearFit = 5e-2;
constexpr int kLenLongTrack = 7;
constexpr int kPredWin = 6;   // predictNext uses at most 6 samples
constexpr int kFitWin = 4;    // checkLinearFit uses at most 4 samples
constexpr int kPfListCap = 256;

#define STB_TRY(x)                                                              \
  do {                                                                          \
    hipError_t _e = (x);                                                        \
    if (_e != hipSuccess)                                                       \
      return set_error(OLPT_E_GPU, (std::string(#x ": ") + hipGetErrorString(_e)).c_str()); \
  } while (0)
#define STB_CHECK(x)                                                            \
  do {                                                                          \
    olpt_status _s = (x);                                                       \
    if (_s) return _s;                                                          \
  } while (0)

int blocks_for(long n, int t = 256, int cap = 65535) {
  long b = (n + t - 1) / t;
  return (int)std::max(1L, std::min<long>(b, cap));
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// predictNext along one axis: normalised LMS fit of y_t from the previous
// `order` samples, iterated until |error| <= 1e-8 (at most 5000 steps),
// then the one-step-ahead prediction (STB.cpp:250-313).
__global__ void k_predict(const double *__restrict__ win, const int32_t *__restrict__ len,
                          int n, double *__restrict__ out, int32_t *__restrict__ ok) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= 3L * n) return;
  const int i = (int)(gid / 3), axis = (int)(gid % 3);
  const int m = len[i];
  if (m < 3) {
    out[3L * i + axis] = 0.0;
    if (axis == 0) ok[i] = 0;
    return;
  }
  if (axis == 0) ok[i] = 1;
  const int order = m - 1;
  double s[kPredWin], f[kPredWin - 1];
#pragma unroll
  for (int j = 0; j < kPredWin; ++j)
    s[j] = j <= order ? win[((long)i * kPredWin + j) * 3 + axis] : 0.0;
  double yt = 0.0;
#pragma unroll
  for (int j = 0; j < kPredWin; ++j)
    if (j == order) yt = s[j];
  bool shifted = false;
  if (fabs(yt) < 1.0) {
    shifted = true;
#pragma unroll
    for (int j = 0; j < kPredWin; ++j)
      if (j <= order) s[j] += 10.0;
  }
  double yp = 0.0;  // s[order], s[order - 1] after the shift
#pragma unroll
  for (int j = 0; j < kPredWin; ++j) {
    if (j == order) yt = s[j];
    if (j == order - 1) yp = s[j];
  }
  double denom = 0.0;
#pragma unroll
  for (int j = 0; j < kPredWin - 1; ++j)
    if (j < order) denom += s[j] * s[j];
  const double step = denom > 1e-12 ? 1.0 / denom : 0.0;
#pragma unroll
  for (int j = 0; j < kPredWin - 1; ++j) f[j] = 0.0;
  double pred = 0.0;
  double err = yt - pred;
  int it = 0;
  while (fabs(err) > kSmallNumber && it < kWienerMaxIter && step > 0.0) {
    pred = 0.0;
#pragma unroll
    for (int j = 0; j < kPredWin - 1; ++j)
      if (j < order) {
        f[j] += step * s[j] * err;
        pred += f[j] * s[j];
      }
    err = yt - pred;
    ++it;
  }
  if (order >= 1 && step == 0.0) {
    pred = yt + (yt - yp);
  } else {
    pred = 0.0;
#pragma unroll
    for (int j = 0; j < kPredWin - 1; ++j)
      if (j < order) pred += f[j] * s[j + 1];
  }
  if (shifted) pred -= 10.0;
  out[3L * i + axis] = pred;
}

// Prediction inside the view volume and visible (inside the image) in at
// least 2 cameras; the reference compares both image coordinates with the
// column count (ObjectInfo.cpp:15-16).
__global__ void k_in_view(const double *__restrict__ xyz, int n, const CamT<double> *__restrict__ cams,
                          int n_cam, double x0, double x1, double y0, double y1, double z0,
                          double z1, int8_t *__restrict__ ok) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double X[3] = {xyz[3L * i], xyz[3L * i + 1], xyz[3L * i + 2]};
  bool in = !(X[0] > x1 || X[0] < x0 || X[1] > y1 || X[1] < y0 || X[2] > z1 || X[2] < z0);
  if (in) {
    int n_out = 0;
    for (int k = 0; k < n_cam; ++k) {
      double u, v;
      project(cams[k], X, u, v);
      const double nc = (double)cams[k].n_col;
      if (u < 0 || u >= nc || v < 0 || v >= nc) {
        ++n_out;
        if (n_cam - n_out < 2) {
          in = false;
          break;
        }
      }
    }
  }
  ok[i] = in ? 1 : 0;
}

// |p - q| <= tol for some track end q (Pt3D::norm: sum of squares in axis
// order, Matrix.hpp:255-261).  Grid: x over objects, y over chunks of
// track ends staged through LDS; a hit sets rep[i] = 1 (rep cleared first).
constexpr int kRepChunk = 1024;
__global__ void __launch_bounds__(256)
k_repeat(const double *__restrict__ obj, int n, const double *__restrict__ ends, int m, double tol,
         int32_t *__restrict__ rep) {
  __shared__ double s[kRepChunk * 3];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int base = blockIdx.y * kRepChunk;
  const int cnt = min(kRepChunk, m - base);
  for (int t = threadIdx.x; t < cnt * 3; t += blockDim.x) s[t] = ends[3L * base + t];
  __syncthreads();
  if (i >= n) return;
  const double p0 = obj[3L * i], p1 = obj[3L * i + 1], p2 = obj[3L * i + 2];
  for (int t = 0; t < cnt; ++t) {
    const double d0 = p0 - s[3 * t], d1 = p1 - s[3 * t + 1], d2 = p2 - s[3 * t + 2];
    double r = 0.0;
    r += d0 * d0;
    r += d1 * d1;
    r += d2 * d2;
    if (sqrt(r) <= tol) {
      rep[i] = 1;
      return;
    }
  }
}

// Block-wide sum / nearest (fixed reduction tree: deterministic).
constexpr int kLinkThreads = 256;
__device__ double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x / 64;
  __syncthreads();
  if (threadIdx.x % 64 == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < kLinkThreads / 64; ++k) t += sh[k];
  return t;
}

// Nearest object to q with squared distance < r2 (nanoflann L2_Simple:
// ((0 + dx^2) + dy^2) + dz^2 with d = q - p), or kUnlinked; ties go to the
// lower index.  Called by all threads of the block.
__device__ int block_nearest(const double *__restrict__ obj, int n, const double q[3], double r2,
                             double *shd, int *shi) {
  double bd = r2;
  int best = kUnlinked;
  for (int j = threadIdx.x; j < n; j += kLinkThreads) {
    const double d0 = q[0] - obj[3L * j], d1 = q[1] - obj[3L * j + 1], d2 = q[2] - obj[3L * j + 2];
    double d = 0.0;
    d += d0 * d0;
    d += d1 * d1;
    d += d2 * d2;
    if (d < bd) {
      bd = d;
      best = j;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double od = __shfl_xor(bd, off, 64);
    const int oi = __shfl_xor(best, off, 64);
    if (oi != kUnlinked && (best == kUnlinked || od < bd || (od == bd && oi < best))) {
      bd = od;
      best = oi;
    }
  }
  const int w = threadIdx.x / 64;
  __syncthreads();
  if (threadIdx.x % 64 == 0) {
    shd[w] = bd;
    shi[w] = best;
  }
  __syncthreads();
  bd = r2;
  best = kUnlinked;
  for (int k = 0; k < kLinkThreads / 64; ++k)
    if (shi[k] != kUnlinked && (best == kUnlinked || shd[k] < bd || (shd[k] == bd && shi[k] < best))) {
      bd = shd[k];
      best = shi[k];
    }
  return best;
}

// Nearest object, one thread (initial-phase linking, one thread per track).
__device__ int nearest(const double *__restrict__ obj, int n, const double q[3], double r2) {
  int best = kUnlinked;
  double bd = r2;
  for (int j = 0; j < n; ++j) {
    const double d0 = q[0] - obj[3L * j], d1 = q[1] - obj[3L * j + 1], d2 = q[2] - obj[3L * j + 2];
    double d = 0.0;
    d += d0 * d0;
    d += d1 * d1;
    d += d2 * d2;
    if (d < bd) {
      bd = d;
      best = j;
    }
  }
  return best;
}

// linkShortTrack (STB.cpp:743-872), one workgroup per short track.  nb_pt:
// each long-active track's point before its last one, rounded to float as
// the reference's track cloud hands it to the k-d tree (Track.h:56-57);
// nb_disp: its last displacement in FP64; nb_len: its length.
__global__ void __launch_bounds__(kLinkThreads)
k_link_short(const double *__restrict__ last, int n_sa, const double *__restrict__ nb_pt,
             const double *__restrict__ nb_disp, const int32_t *__restrict__ nb_len, int n_la,
             const double *__restrict__ obj, int n_obj, double base_r, int n_iter,
             int32_t *__restrict__ link) {
  __shared__ double shd[kLinkThreads / 64];
  __shared__ int shi[kLinkThreads / 64];
  const int i = blockIdx.x;
  if (i >= n_sa) return;
  const double p[3] = {last[3L * i], last[3L * i + 1], last[3L * i + 2]};
  int id = kUnlinked;
  if (base_r > 0.0) {
    double factor = 1.0;
    for (int step = 0; step < n_iter; ++step) {
      factor = step == 0 ? 1.0 : factor * 1.1;
      const double r_track = (3.0 * base_r) * factor;
      const double r2 = r_track * r_track;
      double c_loc = 0.0, s_loc = 0.0;
      for (int j = threadIdx.x; j < n_la; j += kLinkThreads) {
        const double d0 = p[0] - nb_pt[3L * j], d1 = p[1] - nb_pt[3L * j + 1],
                     d2 = p[2] - nb_pt[3L * j + 2];
        double d = 0.0;
        d += d0 * d0;
        d += d1 * d1;
        d += d2 * d2;
        if (d < r2) {
          s_loc += d;
          c_loc += 1.0;
        }
      }
      const double cnt = block_sum(c_loc, shd);
      const double sum_d2 = block_sum(s_loc, shd);
      double vel[3] = {0.0, 0.0, 0.0};
      if (cnt > 0.0) {
        const double mean_d2 = sum_d2 / cnt;
        double L2 = (0.8 * 0.8) * mean_d2;
        const double R = base_r * factor;
        const double lo = (0.1 * R) * (0.1 * R), hi = (0.8 * R) * (0.8 * R);
        L2 = L2 < lo ? lo : (hi < L2 ? hi : L2);
        const double beta2 = 1.0 / (L2 + 1e-12);
        double sw = 0.0, sd0 = 0.0, sd1 = 0.0, sd2 = 0.0;
        for (int j = threadIdx.x; j < n_la; j += kLinkThreads) {
          const double d0 = p[0] - nb_pt[3L * j], d1 = p[1] - nb_pt[3L * j + 1],
                       d2 = p[2] - nb_pt[3L * j + 2];
          double d = 0.0;
          d += d0 * d0;
          d += d1 * d1;
          d += d2 * d2;
          if (!(d < r2) || nb_len[j] < 2) continue;
          const double w = 1.0 / (1.0 + d * beta2);
          sw += w;
          sd0 += w * nb_disp[3L * j];
          sd1 += w * nb_disp[3L * j + 1];
          sd2 += w * nb_disp[3L * j + 2];
        }
        sw = block_sum(sw, shd);
        sd0 = block_sum(sd0, shd);
        sd1 = block_sum(sd1, shd);
        sd2 = block_sum(sd2, shd);
        if (sw > 0.0) {
          const double inv = 1.0 / sw;
          vel[0] = sd0 * inv;
          vel[1] = sd1 * inv;
          vel[2] = sd2 * inv;
        }
      }
      const double est[3] = {p[0] + vel[0], p[1] + vel[1], p[2] + vel[2]};
      const double r_obj = base_r * factor;
      id = block_nearest(obj, n_obj, est, r_obj * r_obj, shd, shi);
      if (id != kUnlinked) break;
      if (cnt == 0.0) {
        id = block_nearest(obj, n_obj, p, r_obj * r_obj, shd, shi);
        if (id != kUnlinked) break;
      }
    }
  }
  if (threadIdx.x == 0) link[i] = id;
}

// Predictive-field grid and bins (PredField::setGrid, PredField.cpp:168-194).
struct PfGrid {
  double x0, x1, y0, y1, z0, z1;
  int nx, ny, nz;
  double dx, dy, dz, r;
  int nb[3];
  double m[3], c[3];
};

__device__ __forceinline__ double lin(double lo, double d, int i) { return lo + d * i; }

// PredField::getDisp (PredField.cpp:91-165) + myMATH::triLinearInterp
// (myMATH.cpp:131-214).
__device__ void pf_disp(const PfGrid &g, const double *__restrict__ field, const double q[3],
                        double out[3]) {
  const double px = fmin(fmax(q[0], g.x0), g.x1);
  const double py = fmin(fmax(q[1], g.y0), g.y1);
  const double pz = fmin(fmax(q[2], g.z0), g.z1);
  int xi = max(0, (int)floor((px - g.x0) / g.dx));
  int yi = max(0, (int)floor((py - g.y0) / g.dy));
  int zi = max(0, (int)floor((pz - g.z0) / g.dz));
  xi = min(xi, g.nx - 2);
  yi = min(yi, g.ny - 2);
  zi = min(zi, g.nz - 2);
  const double gx0 = lin(g.x0, g.dx, xi), gx1 = lin(g.x0, g.dx, xi + 1);
  const double gy0 = lin(g.y0, g.dy, yi), gy1 = lin(g.y0, g.dy, yi + 1);
  const double gz0 = lin(g.z0, g.dz, zi), gz1 = lin(g.z0, g.dz, zi + 1);
  auto id = [&](int a, int b, int c) { return ((long)a * g.ny + b) * g.nz + c; };
  const long i000 = id(xi, yi, zi), i100 = id(xi + 1, yi, zi), i101 = id(xi + 1, yi, zi + 1),
             i001 = id(xi, yi, zi + 1), i010 = id(xi, yi + 1, zi), i110 = id(xi + 1, yi + 1, zi),
             i111 = id(xi + 1, yi + 1, zi + 1), i011 = id(xi, yi + 1, zi + 1);
  double xd = (px - gx0) / (gx1 - gx0);
  double yd = (py - gy0) / (gy1 - gy0);
  double zd = (pz - gz0) / (gz1 - gz0);
  xd = fmax(0.0, fmin(1.0, xd));
  yd = fmax(0.0, fmin(1.0, yd));
  zd = fmax(0.0, fmin(1.0, zd));
  for (int a = 0; a < 3; ++a) {
    const double c000 = field[3 * i000 + a], c100 = field[3 * i100 + a],
                 c101 = field[3 * i101 + a], c001 = field[3 * i001 + a],
                 c010 = field[3 * i010 + a], c110 = field[3 * i110 + a],
                 c111 = field[3 * i111 + a], c011 = field[3 * i011 + a];
    const double c00 = c000 * (1 - xd) + c100 * xd;
    const double c01 = c001 * (1 - xd) + c101 * xd;
    const double c10 = c010 * (1 - xd) + c110 * xd;
    const double c11 = c011 * (1 - xd) + c111 * xd;
    const double c0 = c00 * (1 - yd) + c10 * yd;
    const double c1 = c01 * (1 - yd) + c11 * yd;
    out[a] = c0 * (1 - zd) + c1 * zd;
  }
}

// buildTrackFromPredField's link (STB.cpp:166-181): nearest current
// object to last + getDisp(last) within the search radius.
__global__ void k_link_field(const double *__restrict__ last, int n, PfGrid g,
                             const double *__restrict__ field, const double *__restrict__ obj,
                             int n_obj, double radius, int32_t *__restrict__ link) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {last[3L * i], last[3L * i + 1], last[3L * i + 2]};
  double v[3];
  pf_disp(g, field, p, v);
  const double q[3] = {p[0] + v[0], p[1] + v[1], p[2] + v[2]};
  link[i] = nearest(obj, n_obj, q, radius * radius);
}

// displacements at query points (the standalone PredField::getDisp)
__global__ void k_pf_query(PfGrid g, const double *__restrict__ field, const double *__restrict__ q,
                           int n, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {q[3L * i], q[3L * i + 1], q[3L * i + 2]};
  double v[3];
  pf_disp(g, field, p, v);
  for (int a = 0; a < 3; ++a) out[3L * i + a] = v[a];
}

// checkLinearFit: kern [2][n_pts] = inverse(X'X) X' with X = [1, i]
// (computed on the host in the reference's operation order).
__global__ void k_lin_fit(const double *__restrict__ win, const int32_t *__restrict__ len, int n,
                          const double *__restrict__ kern, int n_pts, int32_t *__restrict__ ok) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (len[i] < n_pts) {
    ok[i] = 0;
    return;
  }
  const double *w = win + (long)i * kFitWin * 3;
  double est[3];
  for (int a = 0; a < 3; ++a) {
    double t0 = 0.0, t1 = 0.0;
    for (int j = 0; j < n_pts; ++j) t0 += kern[j] * w[3 * j + a];
    for (int j = 0; j < n_pts; ++j) t1 += kern[n_pts + j] * w[3 * j + a];
    est[a] = t0 + (double)(n_pts - 1) * t1;
  }
  const double *l = w + 3 * (n_pts - 1);
  double r = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double d = est[a] - l[a];
    r += d * d;
  }
  ok[i] = sqrt(r) > kMaxErrLinearFit ? 0 : 1;
}

// Object indices within the open cube |d| < r of the node, ascending
// (PredField::findNeighbor, PredField.cpp:197-224), gathered from the
// cells of a uniform cell grid (cell_start/cell_obj: CSR of object ids per
// cell, ids ascending within a cell is not required: the list is sorted).
struct PfCells {
  int n[3];
  double lo[3], h;
  const int32_t *start;  // [ncell + 1]
  const int32_t *ids;    // [n_obj]
};

__device__ int pf_gather(const PfCells &cl, const double *__restrict__ obj, const double node[3],
                         double r, int32_t *list, int cap, bool &over) {
  int c0[3], c1[3];
  for (int a = 0; a < 3; ++a) {
    c0[a] = (int)floor((node[a] - r - cl.lo[a]) / cl.h) - 1;
    c1[a] = (int)floor((node[a] + r - cl.lo[a]) / cl.h) + 1;
    c0[a] = max(0, min(cl.n[a] - 1, c0[a]));
    c1[a] = max(0, min(cl.n[a] - 1, c1[a]));
  }
  int cnt = 0;
  for (int x = c0[0]; x <= c1[0]; ++x)
    for (int y = c0[1]; y <= c1[1]; ++y)
      for (int z = c0[2]; z <= c1[2]; ++z) {
        const long c = ((long)x * cl.n[1] + y) * cl.n[2] + z;
        for (int e = cl.start[c]; e < cl.start[c + 1]; ++e) {
          const int j = cl.ids[e];
          const double dx = obj[3L * j] - node[0], dy = obj[3L * j + 1] - node[1],
                       dz = obj[3L * j + 2] - node[2];
          if (dx < r && dx > -r && dy < r && dy > -r && dz < r && dz > -r) {
            if (cnt < cap) list[cnt] = j;
            else over = true;
            ++cnt;
          }
        }
      }
  cnt = min(cnt, cap);
  for (int a = 1; a < cnt; ++a) {  // insertion sort, lists are short
    const int v = list[a];
    int b = a - 1;
    while (b >= 0 && list[b] > v) {
      list[b + 1] = list[b];
      --b;
    }
    list[b + 1] = v;
  }
  return cnt;
}

// fitPeakBinLocGauss (PredField.cpp:345-371)
__device__ double pf_fit(double y1, double v1, double y2, double v2, double y3, double v3) {
  const double l1 = v1 < kLogSmall ? log(kLogSmall) : log(v1);
  const double l2 = v2 < kLogSmall ? log(kLogSmall) : log(v2);
  const double l3 = v3 < kLogSmall ? log(kLogSmall) : log(v3);
  double yc = -0.5 * ((l1 * double((y2 * y2) - (y3 * y3))) - (l2 * double((y1 * y1) - (y3 * y3))) +
                      (l3 * double((y1 * y1) - (y2 * y2)))) /
              ((l1 * double(y3 - y2)) - (l3 * double(y1 - y2)) + (l2 * double(y1 - y3)));
  if (!isfinite(yc)) yc = y2;
  return yc;
}

// One wave per grid node: neighbour lists of both frames, the pair
// displacement histogram (27-bin Gaussian splat per pair, PredField.cpp:
// 240-275; lane l owns offset l), its first maximum (std::max_element) and
// the per-axis peak fit (:278-343).  Histogram and lists sit in LDS: one
// wave's LDS accesses complete in program order, so two pairs whose splats
// share a bin add to it in pair order, as the reference's loop does.  Every
// wave of a block reaches every barrier.
constexpr int kPfWaves = 2;
constexpr int kPfMaxBins = 3375;  // 15^3: histogram + lists stay below 64 KB of LDS
__global__ void __launch_bounds__(64 * kPfWaves)
k_pf_node(PfGrid g, PfCells cp, PfCells cc, const double *__restrict__ prev,
          const double *__restrict__ curr, double *__restrict__ field,
          int32_t *__restrict__ overflow) {
  extern __shared__ double s_pdf[];  // [kPfWaves][nbin]
  __shared__ int32_t s_list[kPfWaves][2][kPfListCap];
  __shared__ int s_n[kPfWaves][2];
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long node = (long)blockIdx.x * kPfWaves + wv;
  const long ng = (long)g.nx * g.ny * g.nz;
  const bool valid = node < ng;
  const int nbin = g.nb[0] * g.nb[1] * g.nb[2];
  double *pdf = s_pdf + (long)wv * nbin;
  int32_t *lp = s_list[wv][0], *lc = s_list[wv][1];
  int xi = 0, yi = 0, zi = 0;
  if (valid) {
    xi = (int)(node / ((long)g.ny * g.nz));
    yi = (int)((node / g.nz) % g.ny);
    zi = (int)(node % g.nz);
  }
  const double nd[3] = {lin(g.x0, g.dx, xi), lin(g.y0, g.dy, yi), lin(g.z0, g.dz, zi)};
  if (lane == 0) {
    bool over = false;
    s_n[wv][0] = valid ? pf_gather(cp, prev, nd, g.r, lp, kPfListCap, over) : 0;
    s_n[wv][1] = valid ? pf_gather(cc, curr, nd, g.r, lc, kPfListCap, over) : 0;
    if (over) overflow[0] = 1;
  }
  for (int b = lane; b < nbin; b += 64) pdf[b] = 0.0;
  __syncthreads();
  const int np = s_n[wv][0], nc = s_n[wv][1];
  const int oi = lane / 9 - 1, oj = (lane / 3) % 3 - 1, ok = lane % 3 - 1;
  for (int j = 0; j < nc; ++j) {
    const int ic = lc[j];
    for (int k = 0; k < np; ++k) {
      const int ip = lp[k];
      const double d0 = curr[3L * ic] - prev[3L * ip];
      const double d1 = curr[3L * ic + 1] - prev[3L * ip + 1];
      const double d2 = curr[3L * ic + 2] - prev[3L * ip + 2];
      const double xb = g.m[0] * d0 + g.c[0];
      const double yb = g.m[1] * d1 + g.c[1];
      const double zb = g.m[2] * d2 + g.c[2];
      const int xr = (int)round(xb), yr = (int)round(yb), zr = (int)round(zb);
      if (lane < 27) {
        const int bi = xr + oi, bj = yr + oj, bk = zr + ok;
        if (bi >= max(0, xr - 1) && bi <= min(g.nb[0] - 1, xr + 1) && bj >= max(0, yr - 1) &&
            bj <= min(g.nb[1] - 1, yr + 1) && bk >= max(0, zr - 1) &&
            bk <= min(g.nb[2] - 1, zr + 1)) {
          const double xx = bi - xb, yy = bj - yb, zz = bk - zb;
          const int bin = (bi * g.nb[1] + bj) * g.nb[2] + bk;
          pdf[bin] += exp(-(xx * xx + yy * yy + zz * zz));
        }
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  if (np == 0 || nc == 0) {
    if (lane < 3) field[3 * node + lane] = 0.0;
    return;
  }
  // first maximum
  double bv = -1.0;
  int bidx = 0x7fffffff;
  for (int b = lane; b < nbin; b += 64) {
    const double v = pdf[b];
    if (v > bv) {
      bv = v;
      bidx = b;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int oi2 = __shfl_xor(bidx, off, 64);
    if (ov > bv || (ov == bv && oi2 < bidx)) {
      bv = ov;
      bidx = oi2;
    }
  }
  if (lane != 0) return;
  int pk[3] = {-1, -1, -1};
  if (bv > 1e-1) {
    pk[0] = bidx / (g.nb[1] * g.nb[2]);
    pk[1] = (bidx - pk[0] * g.nb[1] * g.nb[2]) / g.nb[2];
    pk[2] = bidx % g.nb[2];
  }
  for (int a = 0; a < 3; ++a) {
    double v;
    if (pk[a] < 0) {
      v = 0.0;
    } else if (pk[a] == 0) {
      v = -2 * g.r;
    } else if (pk[a] == g.nb[a] - 1) {
      v = 2 * g.r;
    } else {
      int lo[3] = {pk[0], pk[1], pk[2]}, hi[3] = {pk[0], pk[1], pk[2]};
      lo[a] -= 1;
      hi[a] += 1;
      auto at = [&](const int *t) { return pdf[(t[0] * g.nb[1] + t[1]) * g.nb[2] + t[2]]; };
      const double loc = pf_fit(pk[a] - 1, at(lo), pk[a], at(pk), pk[a] + 1, at(hi));
      v = (loc - g.c[a]) / g.m[a];
    }
    field[3 * node + a] = v;
  }
}

// One direction of applyGaussian (PredField.cpp:374-466) for all three
// components; kern [3][2*rad_max+1] (component kernels), rad [3].
__global__ void k_pf_smooth(PfGrid g, const double *__restrict__ in, double *__restrict__ out,
                            int dir, const double *__restrict__ kern, int kstride, int3 rad) {
  const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ng = (long)g.nx * g.ny * g.nz;
  if (node >= ng) return;
  const int x = (int)(node / ((long)g.ny * g.nz));
  const int y = (int)((node / g.nz) % g.ny);
  const int z = (int)(node % g.nz);
  const int nn = dir == 0 ? g.nx : (dir == 1 ? g.ny : g.nz);
  const int pos = dir == 0 ? x : (dir == 1 ? y : z);
  for (int a = 0; a < 3; ++a) {
    const int r = a == 0 ? rad.x : (a == 1 ? rad.y : rad.z);
    double s = 0.0;
    for (int k = -r; k <= r; ++k) {
      const int idx = pos + k;
      if (idx >= 0 && idx < nn) {
        const int sx = dir == 0 ? idx : x, sy = dir == 1 ? idx : y, sz = dir == 2 ? idx : z;
        const long src = ((long)sx * g.ny + sy) * g.nz + sz;
        s += in[3 * src + a] * kern[a * kstride + k + r];
      }
    }
    out[3 * node + a] = s;
  }
}

// cell index of points (uniform grid of cell size h over [lo, lo + n h))
__global__ void k_cell_of(const double *__restrict__ p, int n, PfCells cl, int32_t *__restrict__ cell,
                          int32_t *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long c = 0;
  for (int a = 0; a < 3; ++a) {
    int q = (int)floor((p[3L * i + a] - cl.lo[a]) / cl.h);
    q = max(0, min(cl.n[a] - 1, q));
    c = c * cl.n[a] + q;
  }
  cell[i] = (int32_t)c;
  atomicAdd(&cnt[c], 1);
}

__global__ void k_cell_put(int n, const int32_t *__restrict__ cell, int32_t *__restrict__ fill,
                           int32_t *__restrict__ ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int pos = atomicAdd(&fill[cell[i]], 1);
  ids[pos] = i;
}

__global__ void k_fill_i32_stb(int32_t *p, long n, int32_t v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T> T *as() { return (T *)p; }
  olpt_status need(size_t b) {
    if (b <= bytes && p) return OLPT_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    const size_t nb = std::max<size_t>(b, 256);
    if (hipMalloc(&p, nb) != hipSuccess) {
      p = nullptr;
      return set_error(OLPT_E_ALLOC, "hipMalloc failed in the tracking module");
    }
    bytes = nb;
    return OLPT_OK;
  }
};

// myMATH::internal::gaussForward / gaussBackward (myMATH.h:183-289) for a
// small n x n matrix (row-major), as the reference applies them.
void gauss_inverse_n(const std::vector<double> &A, int n, std::vector<double> &inv) {
  std::vector<double> u(A), b(n * n, 0.0);
  for (int i = 0; i < n; ++i) b[i * n + i] = 1.0;
  std::vector<int> idx(n);
  std::vector<double> rmax(n);
  for (int x = 0; x < n; ++x) {
    idx[x] = x;
    rmax[x] = *std::max_element(u.begin() + x * n, u.begin() + x * n + n);
  }
  for (int y = 0; y < n - 1; ++y) {
    double mp = 0;
    int t = 0;
    for (int x = y; x < n; ++x) {
      const double pv = std::fabs(u[idx[x] * n + y] / rmax[idx[x]]);
      if (pv > mp) {
        mp = pv;
        t = x;
      }
    }
    std::swap(idx[t], idx[y]);
    for (int x = y + 1; x < n; ++x) {
      const double m = u[idx[x] * n + y] / u[idx[y] * n + y];
      u[idx[x] * n + y] = m;
      for (int k = y + 1; k < n; ++k) u[idx[x] * n + k] -= m * u[idx[y] * n + k];
    }
  }
  inv = b;
  for (int c = 0; c < n; ++c) {
    for (int y = 0; y < n - 1; ++y)
      for (int x = y + 1; x < n; ++x) b[idx[x] * n + c] -= u[idx[x] * n + y] * b[idx[y] * n + c];
    inv[(n - 1) * n + c] = b[idx[n - 1] * n + c] / u[idx[n - 1] * n + n - 1];
    for (int x = n - 2; x > -1; --x) {
      double z = b[idx[x] * n + c];
      for (int y = x + 1; y < n; ++y) z -= u[idx[x] * n + y] * inv[y * n + c];
      inv[x * n + c] = z / u[idx[x] * n + x];
    }
  }
}

// checkLinearFit's kernel inverse(X'X) X' for X = [1, i], i < n_pts, with
// Matrix::operator* summation order (Matrix.hpp:392-408).
std::vector<double> linfit_kernel(int n_pts) {
  std::vector<double> X(n_pts * 2), Xt(2 * n_pts), XtX(4, 0.0), inv, K(2 * n_pts, 0.0);
  for (int i = 0; i < n_pts; ++i) {
    X[i * 2] = 1;
    X[i * 2 + 1] = i;
    Xt[i] = 1;
    Xt[n_pts + i] = i;
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < n_pts; ++k) XtX[a * 2 + b] += Xt[a * n_pts + k] * X[k * 2 + b];
  gauss_inverse_n(XtX, 2, inv);
  for (int a = 0; a < 2; ++a)
    for (int j = 0; j < n_pts; ++j)
      for (int k = 0; k < 2; ++k) K[a * n_pts + j] += inv[a * 2 + k] * Xt[k * n_pts + j];
  return K;
}

// createGaussianKernel (myMATH.cpp:537-557)
std::vector<double> gauss_kernel(int radius, double sigma) {
  const int size = 2 * radius + 1;
  std::vector<double> k(size);
  double sum = 0.0;
  for (int i = 0; i < size; ++i) {
    const int x = i - radius;
    k[i] = std::exp(-(x * x) / (2 * sigma * sigma));
    sum += k[i];
  }
  for (int i = 0; i < size; ++i) k[i] /= sum;
  return k;
}

PfGrid make_grid(const olpt_stb_params &sp) {
  PfGrid g;
  g.x0 = sp.limit[0];
  g.x1 = sp.limit[1];
  g.y0 = sp.limit[2];
  g.y1 = sp.limit[3];
  g.z0 = sp.limit[4];
  g.z1 = sp.limit[5];
  g.nx = sp.pf_nx;
  g.ny = sp.pf_ny;
  g.nz = sp.pf_nz;
  g.dx = (g.x1 - g.x0) / (g.nx - 1);
  g.dy = (g.y1 - g.y0) / (g.ny - 1);
  g.dz = (g.z1 - g.z0) / (g.nz - 1);
  g.r = sp.pf_r_mm;
  for (int a = 0; a < 3; ++a) {
    g.nb[a] = sp.pf_nbin[a];
    g.m[a] = (g.nb[a] - 1) / (4 * g.r);
    g.c[a] = (double)((g.nb[a] - 1) / 2);  // integer division, as the reference
  }
  return g;
}

bool pf_params_ok(const olpt_stb_params &sp) {
  return sp.pf_nx >= 2 && sp.pf_ny >= 2 && sp.pf_nz >= 2 && sp.pf_r_mm > 0 &&
         sp.pf_nbin[0] >= 2 && sp.pf_nbin[1] >= 2 && sp.pf_nbin[2] >= 2 &&
         (long)sp.pf_nbin[0] * sp.pf_nbin[1] * sp.pf_nbin[2] <= kPfMaxBins;
}

// Predictive field on the device; prev/curr device arrays.
struct PfWork {
  DBuf cell_p, cell_c, cnt_p, cnt_c, start_p, start_c, ids_p, ids_c, field, tmp, kern, over;
};

olpt_status cells_build(const double *d_pts, int n, PfCells &cl, DBuf &cell, DBuf &cnt,
                        DBuf &start, DBuf &ids, hipStream_t st) {
  const long nc = (long)cl.n[0] * cl.n[1] * cl.n[2];
  STB_CHECK(cell.need(sizeof(int32_t) * std::max(1, n)));
  STB_CHECK(cnt.need(sizeof(int32_t) * (nc + 1)));
  STB_CHECK(start.need(sizeof(int32_t) * (nc + 1)));
  STB_CHECK(ids.need(sizeof(int32_t) * std::max(1, n)));
  k_fill_i32_stb<<<blocks_for(nc + 1), 256, 0, st>>>(cnt.as<int32_t>(), nc + 1, 0);
  if (n > 0) k_cell_of<<<blocks_for(n), 256, 0, st>>>(d_pts, n, cl, cell.as<int32_t>(), cnt.as<int32_t>());
  // exclusive scan on the host: the cell grid is small (a few thousand cells)
  std::vector<int32_t> h(nc + 1);
  STB_TRY(hipMemcpyAsync(h.data(), cnt.p, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, st));
  STB_TRY(hipStreamSynchronize(st));
  int32_t acc = 0;
  for (long i = 0; i < nc; ++i) {
    const int32_t v = h[i];
    h[i] = acc;
    acc += v;
  }
  h[nc] = acc;
  STB_TRY(hipMemcpyAsync(start.p, h.data(), sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice, st));
  STB_TRY(hipMemcpyAsync(cnt.p, h.data(), sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice, st));
  if (n > 0) k_cell_put<<<blocks_for(n), 256, 0, st>>>(n, cell.as<int32_t>(), cnt.as<int32_t>(), ids.as<int32_t>());
  cl.start = start.as<int32_t>();
  cl.ids = ids.as<int32_t>();
  STB_TRY(hipGetLastError());
  return OLPT_OK;
}

olpt_status pred_field(const olpt_stb_params &sp, const double *d_prev, int n_prev,
                       const double *d_curr, int n_curr, PfWork &w, hipStream_t st) {
  const PfGrid g = make_grid(sp);
  const long ng = (long)g.nx * g.ny * g.nz;
  const int nbin = g.nb[0] * g.nb[1] * g.nb[2];
  PfCells cp, cc;
  const double lo[3] = {g.x0, g.y0, g.z0}, hi[3] = {g.x1, g.y1, g.z1};
  // cells of at least 2r (a node's cube spans at most 2 cells per axis),
  // at most 64 per axis; points outside fall into the border cells
  double h = 2.0 * g.r;
  for (int a = 0; a < 3; ++a) h = std::max(h, (hi[a] - lo[a] + 4.0 * g.r) / 63.0);
  for (int a = 0; a < 3; ++a) {
    cp.lo[a] = lo[a] - 2.0 * g.r;
    cp.n[a] = std::max(1, std::min(64, (int)std::ceil((hi[a] - lo[a] + 4.0 * g.r) / h) + 1));
  }
  cp.h = h;
  cc = cp;
  STB_CHECK(cells_build(d_prev, n_prev, cp, w.cell_p, w.cnt_p, w.start_p, w.ids_p, st));
  STB_CHECK(cells_build(d_curr, n_curr, cc, w.cell_c, w.cnt_c, w.start_c, w.ids_c, st));
  const long nblk = (ng + kPfWaves - 1) / kPfWaves;
  STB_CHECK(w.field.need(sizeof(double) * 3 * ng));
  STB_CHECK(w.tmp.need(sizeof(double) * 3 * ng));
  STB_CHECK(w.over.need(sizeof(int32_t)));
  k_fill_i32_stb<<<1, 64, 0, st>>>(w.over.as<int32_t>(), 1, 0);
  k_pf_node<<<(unsigned)nblk, 64 * kPfWaves, sizeof(double) * kPfWaves * nbin, st>>>(
      g, cp, cc, d_prev, d_curr, w.field.as<double>(), w.over.as<int32_t>());
  STB_TRY(hipGetLastError());
  int32_t over = 0;
  STB_TRY(hipMemcpyAsync(&over, w.over.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  STB_TRY(hipStreamSynchronize(st));
  if (over)
    return set_error(OLPT_E_CAPACITY, "predictive field: more than 256 objects near one grid node");
  if (sp.pf_smooth) {
    int r[3];
    std::vector<std::vector<double>> ks(3);
    int kmax = 0;
    for (int a = 0; a < 3; ++a) {
      r[a] = (int)std::round(3 * sp.pf_sigma[a]);
      ks[a] = gauss_kernel(r[a], sp.pf_sigma[a]);
      kmax = std::max(kmax, 2 * r[a] + 1);
    }
    std::vector<double> hk(3 * kmax, 0.0);
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < 2 * r[a] + 1; ++i) hk[a * kmax + i] = ks[a][i];
    STB_CHECK(w.kern.need(sizeof(double) * hk.size()));
    STB_TRY(hipMemcpyAsync(w.kern.p, hk.data(), sizeof(double) * hk.size(), hipMemcpyHostToDevice, st));
    double *a = w.field.as<double>(), *b = w.tmp.as<double>();
    for (int dir = 0; dir < 3; ++dir) {
      k_pf_smooth<<<blocks_for(ng, 256, 1 << 30), 256, 0, st>>>(g, a, b, dir, w.kern.as<double>(),
                                                                 kmax, make_int3(r[0], r[1], r[2]));
      std::swap(a, b);
    }
    if (a != w.field.as<double>())
      STB_TRY(hipMemcpyAsync(w.field.p, a, sizeof(double) * 3 * ng, hipMemcpyDeviceToDevice, st));
    STB_TRY(hipGetLastError());
  }
  return OLPT_OK;
}

int mkdirs(const std::string &path) {
  std::string cur;
  for (size_t i = 0; i < path.size(); ++i) {
    cur.push_back(path[i]);
    if (path[i] == '/' || i + 1 == path.size()) {
      if (cur == "/" || cur == "./") continue;
      struct stat sb;
      if (stat(cur.c_str(), &sb) != 0 && mkdir(cur.c_str(), 0755) != 0) return 1;
    }
  }
  return 0;
}

} // namespace

// ---------------------------------------------------------------------------
// tracking state
// ---------------------------------------------------------------------------
struct StbPt {
  double x, y, z;
};
struct StbTrack {
  std::vector<StbPt> p;
  std::vector<int> t;
  size_t size() const { return t.size(); }
  void add(StbPt q, int frame) {
    p.push_back(q);
    t.push_back(frame);
  }
};

struct olpt_stb {
  olpt_ctx *ctx = nullptr;
  CtxInfo ci;
  olpt_stb_params sp;
  olpt_ipr_params ip;
  std::string out;
  std::vector<CamT<double>> hcam;
  std::vector<StbTrack> lists[4];  // OLPT_TRACK_*
  std::vector<std::vector<StbPt>> cand;
  std::vector<std::vector<char>> cand_tracked;
  // device buffers
  DBuf win, len, pred, okb, in_view, xyz_sh, flags, score, res, obj_xyz, obj_n, uv, ends, rep,
      last, nb_pt, nb_disp, nb_len, link, kern, cand_d;
  PfWork pf;
  int ipr_cap = 0;
  // learned shake proposal for the IPR calls (olpt_stb_set_ipr_model)
  const olpt_mlp *ipr_mlp = nullptr;
  double ipr_mlp_vox = 0.0, ipr_mlp_step = 0.0;
  olpt_ipr_params ip_ipr{};
};

namespace {

void project_host(const olpt_stb *s, int k, const StbPt &q, double &u, double &v) {
  const double X[3] = {q.x, q.y, q.z};
  project(s->hcam[k], X, u, v);
}

// STB::saveTracks (STB.cpp:967-1004) + Track::saveTrack + Tracer3D::
// saveObject3D (ObjectInfo.cpp:76-94): fixed, 6 decimals.
olpt_status save_list(const olpt_stb *s, const std::string &file, const std::vector<StbTrack> &l) {
  FILE *f = std::fopen(file.c_str(), "w");
  if (!f) return set_error(OLPT_E_IO, ("Cannot open file:" + file).c_str());
  std::fprintf(f, "TrackID,FrameID,WorldX,WorldY,WorldZ");
  for (int k = 0; k < s->ci.n_cam; ++k) std::fprintf(f, ",cam%d_x(col),cam%d_y(row)", k, k);
  std::fprintf(f, "\n");
  for (size_t i = 0; i < l.size(); ++i)
    for (size_t j = 0; j < l[i].size(); ++j) {
      const StbPt &q = l[i].p[j];
      std::fprintf(f, "%d,%d,%.6f,%.6f,%.6f", (int)i, l[i].t[j], q.x, q.y, q.z);
      for (int k = 0; k < s->ci.n_cam; ++k) {
        double u, v;
        project_host(s, k, q, u, v);
        std::fprintf(f, ",%.6f,%.6f", u, v);
      }
      std::fprintf(f, "\n");
    }
  std::fclose(f);
  return OLPT_OK;
}

olpt_status save_all(olpt_stb *s, const std::string &folder, int frame) {
  if (mkdirs(folder)) return set_error(OLPT_E_IO, ("cannot create folder " + folder).c_str());
  std::string dir = folder;
  if (!dir.empty() && dir.back() != '/') dir.push_back('/');
  const std::string f = std::to_string(frame);
  // remove every older active snapshot (STB.cpp:931-953)
  if (DIR *d = opendir(dir.c_str())) {
    std::vector<std::string> rm;
    while (dirent *e = readdir(d)) {
      const std::string name = e->d_name;
      for (const std::string pre : {"LongTrackActive_", "ShortTrackActive_"})
        if (name.size() >= pre.size() + 4 && name.compare(0, pre.size(), pre) == 0 &&
            name.rfind(".csv") == name.size() - 4)
          rm.push_back(dir + name);
    }
    closedir(d);
    for (const auto &p : rm) (void)std::remove(p.c_str());
  }
  STB_CHECK(save_list(s, dir + "LongTrackActive_" + f + ".csv", s->lists[OLPT_TRACK_LONG_ACTIVE]));
  STB_CHECK(save_list(s, dir + "ShortTrackActive_" + f + ".csv", s->lists[OLPT_TRACK_SHORT_ACTIVE]));
  STB_CHECK(save_list(s, dir + "ExitTrack_" + f + ".csv", s->lists[OLPT_TRACK_EXIT]));
  STB_CHECK(save_list(s, dir + "LongTrackInactive_" + f + ".csv", s->lists[OLPT_TRACK_LONG_INACTIVE]));
  return OLPT_OK;
}

// Matrix::write (Matrix.hpp:266-292): scientific, 8 digits
olpt_status save_field(const std::string &file, const std::vector<double> &f) {
  FILE *fp = std::fopen(file.c_str(), "w");
  if (!fp) return set_error(OLPT_E_IO, ("Cannot open file:" + file).c_str());
  for (size_t i = 0; i + 2 < f.size(); i += 3)
    std::fprintf(fp, "%.8e,%.8e,%.8e\n", f[i], f[i + 1], f[i + 2]);
  std::fclose(fp);
  return OLPT_OK;
}

olpt_status upload_pts(DBuf &b, const std::vector<StbPt> &v, hipStream_t st) {
  STB_CHECK(b.need(sizeof(StbPt) * std::max<size_t>(1, v.size())));
  if (!v.empty())
    STB_TRY(hipMemcpyAsync(b.p, v.data(), sizeof(StbPt) * v.size(), hipMemcpyHostToDevice, st));
  return OLPT_OK;
}

// IPR on d_imgs (dtype) -> host objects
olpt_status run_ipr(olpt_stb *s, const void *d_imgs, int dtype, std::vector<StbPt> &objs) {
  hipStream_t st = s->ci.st;
  if (!s->ipr_cap) s->ipr_cap = std::max(1 << 16, 4 * s->ci.cap_obj);
  STB_CHECK(s->obj_xyz.need(sizeof(double) * 3 * (size_t)s->ipr_cap));
  STB_CHECK(s->obj_n.need(sizeof(int32_t)));
  if (s->ipr_mlp) {
    // the proposal serves this IPR call only: the shake of the predictions
    // keeps the tracer configuration's plain schedule
    STB_CHECK(olpt_set_shake_model(s->ctx, s->ipr_mlp, s->ipr_mlp_vox, s->ipr_mlp_step));
    const olpt_status r = olpt_ipr_run(s->ctx, d_imgs, dtype, 1, &s->ip_ipr,
                                       s->obj_xyz.as<double>(), nullptr, s->ipr_cap,
                                       s->obj_n.as<int32_t>(), nullptr, 0);
    STB_CHECK(olpt_set_shake_model(s->ctx, nullptr, 0.0, 0.0));
    STB_CHECK(r);
  } else {
    STB_CHECK(olpt_ipr_run(s->ctx, d_imgs, dtype, 1, &s->ip, s->obj_xyz.as<double>(), nullptr,
                           s->ipr_cap, s->obj_n.as<int32_t>(), nullptr, 0));
  }
  int32_t n = 0;
  STB_TRY(hipMemcpyAsync(&n, s->obj_n.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  STB_TRY(hipStreamSynchronize(st));
  n = std::min(n, s->ipr_cap);
  objs.resize(n);
  if (n)
    STB_TRY(hipMemcpyAsync(objs.data(), s->obj_xyz.p, sizeof(StbPt) * n, hipMemcpyDeviceToHost, st));
  STB_TRY(hipStreamSynchronize(st));
  return OLPT_OK;
}

// STB::buildTrackFromPredField (STB.cpp:160-231)
olpt_status build_from_field(olpt_stb *s, int frame_id, bool with_field) {
  hipStream_t st = s->ci.st;
  const int id = frame_id - s->sp.frame_start;
  std::vector<StbTrack> &sa = s->lists[OLPT_TRACK_SHORT_ACTIVE];
  std::vector<StbPt> &cand = s->cand[id];
  std::vector<char> &tracked = s->cand_tracked[id];
  if (id > 0 && with_field) {
    const int n_sa = (int)sa.size();
    std::vector<StbPt> last(n_sa);
    for (int j = 0; j < n_sa; ++j) last[j] = sa[j].p.back();
    std::vector<int32_t> link(n_sa, kUnlinked);
    if (n_sa > 0) {
      STB_CHECK(upload_pts(s->last, last, st));
      STB_CHECK(upload_pts(s->cand_d, cand, st));
      STB_CHECK(s->link.need(sizeof(int32_t) * n_sa));
      k_link_field<<<blocks_for(n_sa), 256, 0, st>>>(
          s->last.as<double>(), n_sa, make_grid(s->sp), s->pf.field.as<double>(),
          s->cand_d.as<double>(), (int)cand.size(), s->sp.radius_search_obj_mm,
          s->link.as<int32_t>());
      STB_TRY(hipGetLastError());
      STB_TRY(hipMemcpyAsync(link.data(), s->link.p, sizeof(int32_t) * n_sa, hipMemcpyDeviceToHost, st));
      STB_TRY(hipStreamSynchronize(st));
    }
    // reverse order: erase the unlinked (STB.cpp:183-204)
    std::vector<StbTrack> kept;
    kept.reserve(n_sa);
    for (int j = 0; j < n_sa; ++j) {
      if (link[j] != kUnlinked) {
        tracked[link[j]] = 1;
        sa[j].add(cand[link[j]], frame_id);
        kept.push_back(std::move(sa[j]));
      }
    }
    sa.swap(kept);
  }
  for (size_t i = 0; i < cand.size(); ++i)
    if (!tracked[i]) {
      tracked[i] = 1;
      StbTrack t;
      t.add(cand[i], frame_id);
      sa.push_back(std::move(t));
    }
  return OLPT_OK;
}

// STB::runInitPhase (STB.cpp:91-158)
olpt_status init_phase(olpt_stb *s, int frame, const void *d_imgs, int dtype) {
  std::vector<StbPt> objs;
  STB_CHECK(run_ipr(s, d_imgs, dtype, objs));
  s->cand.push_back(std::move(objs));
  s->cand_tracked.emplace_back(s->cand.back().size(), 0);
  const int n_init = s->sp.n_initial_frames;
  if (frame - s->sp.frame_start != n_init - 1) return OLPT_OK;
  hipStream_t st = s->ci.st;
  for (int i = 0; i < n_init; ++i) {
    const int fid = i + s->sp.frame_start;
    if (i == 0) {
      STB_CHECK(build_from_field(s, fid, false));
    } else {
      DBuf dp, dc;
      STB_CHECK(upload_pts(dp, s->cand[i - 1], st));
      STB_CHECK(upload_pts(dc, s->cand[i], st));
      STB_CHECK(pred_field(s->sp, dp.as<double>(), (int)s->cand[i - 1].size(), dc.as<double>(),
                           (int)s->cand[i].size(), s->pf, st));
      if (!s->out.empty()) {
        const PfGrid g = make_grid(s->sp);
        std::vector<double> f(3L * g.nx * g.ny * g.nz);
        STB_TRY(hipMemcpyAsync(f.data(), s->pf.field.p, sizeof(double) * f.size(),
                               hipMemcpyDeviceToHost, st));
        STB_TRY(hipStreamSynchronize(st));
        STB_CHECK(save_field(s->out + "PredField_" + std::to_string(fid) + "_" +
                                 std::to_string(fid + 1) + ".csv", f));
      }
      STB_CHECK(build_from_field(s, fid, true));
    }
  }
  // tracks of at least n_init points become long tracks (STB.cpp:133-147)
  std::vector<StbTrack> &sa = s->lists[OLPT_TRACK_SHORT_ACTIVE];
  std::vector<StbTrack> &la = s->lists[OLPT_TRACK_LONG_ACTIVE];
  std::vector<StbTrack> keep;
  for (auto &t : sa) {
    if (t.p.size() >= (size_t)n_init) la.push_back(std::move(t));
    else keep.push_back(std::move(t));
  }
  sa.swap(keep);
  if (!s->out.empty()) STB_CHECK(save_all(s, s->out + "InitialTrack/", frame));
  return OLPT_OK;
}

// last min(len, W) points of each track, oldest first -> [n][W][3] + lens
olpt_status upload_windows(olpt_stb *s, const std::vector<StbTrack> &l, int W, bool len_full,
                           hipStream_t st) {
  const size_t n = l.size();
  std::vector<double> w(n * W * 3, 0.0);
  std::vector<int32_t> ln(n);
  for (size_t i = 0; i < n; ++i) {
    const size_t L = l[i].p.size();
    const size_t m = std::min<size_t>(L, W);
    for (size_t j = 0; j < m; ++j) {
      const StbPt &q = l[i].p[L - m + j];
      w[(i * W + j) * 3] = q.x;
      w[(i * W + j) * 3 + 1] = q.y;
      w[(i * W + j) * 3 + 2] = q.z;
    }
    ln[i] = len_full ? (int32_t)l[i].t.size() : (int32_t)m;
  }
  STB_CHECK(s->win.need(sizeof(double) * std::max<size_t>(1, w.size())));
  STB_CHECK(s->len.need(sizeof(int32_t) * std::max<size_t>(1, n)));
  if (n) {
    STB_TRY(hipMemcpyAsync(s->win.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, st));
    STB_TRY(hipMemcpyAsync(s->len.p, ln.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  }
  return OLPT_OK;
}

// STB::runConvPhase (STB.cpp:337-695), volume self-calibration excluded
olpt_status conv_phase(olpt_stb *s, int frame, const void *d_imgs, int dtype) {
  hipStream_t st = s->ci.st;
  const CtxInfo &ci = s->ci;
  std::vector<StbTrack> &sa = s->lists[OLPT_TRACK_SHORT_ACTIVE];
  std::vector<StbTrack> &la = s->lists[OLPT_TRACK_LONG_ACTIVE];
  std::vector<StbTrack> &li = s->lists[OLPT_TRACK_LONG_INACTIVE];
  std::vector<StbTrack> &ex = s->lists[OLPT_TRACK_EXIT];
  const int n_sa = (int)sa.size();
  // 1. prediction of every active long track, then view check
  int n_la = (int)la.size();
  std::vector<StbPt> pred(n_la);
  if (n_la > 0) {
    STB_CHECK(upload_windows(s, la, kPredWin, false, st));
    STB_CHECK(s->pred.need(sizeof(double) * 3 * n_la));
    STB_CHECK(s->okb.need(sizeof(int32_t) * n_la));
    STB_CHECK(s->in_view.need((size_t)n_la));
    k_predict<<<blocks_for(3L * n_la), 256, 0, st>>>(s->win.as<double>(), s->len.as<int32_t>(), n_la,
                                                     s->pred.as<double>(), s->okb.as<int32_t>());
    k_in_view<<<blocks_for(n_la), 256, 0, st>>>(s->pred.as<double>(), n_la, ci.d_cam64, ci.n_cam,
                                                s->sp.limit[0], s->sp.limit[1], s->sp.limit[2],
                                                s->sp.limit[3], s->sp.limit[4], s->sp.limit[5],
                                                s->in_view.as<int8_t>());
    STB_TRY(hipGetLastError());
    std::vector<int8_t> inr(n_la);
    STB_TRY(hipMemcpyAsync(pred.data(), s->pred.p, sizeof(StbPt) * n_la, hipMemcpyDeviceToHost, st));
    STB_TRY(hipMemcpyAsync(inr.data(), s->in_view.p, (size_t)n_la, hipMemcpyDeviceToHost, st));
    STB_TRY(hipStreamSynchronize(st));
    std::vector<StbTrack> keep;
    std::vector<StbPt> kpred;
    for (int i = 0; i < n_la; ++i) {
      if (!inr[i]) {
        if (la[i].t.size() >= (size_t)kLenLongTrack) ex.push_back(std::move(la[i]));
      } else {
        keep.push_back(std::move(la[i]));
        kpred.push_back(pred[i]);
      }
    }
    la.swap(keep);
    pred.swap(kpred);
  }
  // 2. shake the predictions; residual of the kept ones
  const void *ipr_imgs = d_imgs;
  int ipr_dtype = dtype;
  const size_t tsz = ci.prec == OLPT_PREC_F64 ? 8 : 4;
  if (!pred.empty()) {
    const int np = (int)pred.size();
    if (np > ci.cap_obj)
      return set_error(OLPT_E_CAPACITY, "more predicted tracks than the context's object capacity");
    STB_CHECK(upload_pts(s->pred, pred, st));
    STB_CHECK(s->xyz_sh.need(sizeof(double) * 3 * np));
    STB_CHECK(s->flags.need((size_t)np));
    STB_CHECK(s->score.need(sizeof(double) * np));
    STB_CHECK(s->res.need(tsz * (size_t)ci.n_cam * ci.H * ci.W));
    STB_CHECK(olpt_shake(s->ctx, nullptr, d_imgs, dtype, &s->ip, s->pred.as<double>(), np,
                         s->xyz_sh.as<double>(), s->flags.as<uint8_t>(), s->score.as<double>(),
                         s->res.p));
    std::vector<StbPt> xs(np);
    std::vector<uint8_t> fl(np);
    STB_TRY(hipMemcpyAsync(xs.data(), s->xyz_sh.p, sizeof(StbPt) * np, hipMemcpyDeviceToHost, st));
    STB_TRY(hipMemcpyAsync(fl.data(), s->flags.p, (size_t)np, hipMemcpyDeviceToHost, st));
    STB_TRY(hipStreamSynchronize(st));
    std::vector<StbTrack> keep;
    for (int r = 0; r < np; ++r) {
      if (fl[r] == 0) {
        la[r].add(xs[r], frame);
        keep.push_back(std::move(la[r]));
      } else if (la[r].t.size() >= (size_t)kLenLongTrack) {
        li.push_back(std::move(la[r]));
      }
    }
    la.swap(keep);
    ipr_imgs = s->res.p;
    ipr_dtype = ci.prec == OLPT_PREC_F64 ? OLPT_F64 : OLPT_F32;
  }
  // 3. IPR on the residual
  std::vector<StbPt> objs;
  STB_CHECK(run_ipr(s, ipr_imgs, ipr_dtype, objs));
  // 4. drop objects repeating an active long track's last point
  if (!objs.empty() && !la.empty()) {
    std::vector<StbPt> ends(la.size());
    for (size_t i = 0; i < la.size(); ++i) ends[i] = la[i].p.back();
    STB_CHECK(upload_pts(s->ends, ends, st));
    STB_CHECK(upload_pts(s->cand_d, objs, st));
    const int no = (int)objs.size(), ne = (int)ends.size();
    STB_CHECK(s->rep.need(sizeof(int32_t) * no));
    k_fill_i32_stb<<<blocks_for(no), 256, 0, st>>>(s->rep.as<int32_t>(), no, 0);
    k_repeat<<<dim3((no + 255) / 256, (ne + kRepChunk - 1) / kRepChunk), 256, 0, st>>>(
        s->cand_d.as<double>(), no, s->ends.as<double>(), ne, s->ip.tol_3d_mm,
        s->rep.as<int32_t>());
    STB_TRY(hipGetLastError());
    std::vector<int32_t> rep(no);
    STB_TRY(hipMemcpyAsync(rep.data(), s->rep.p, sizeof(int32_t) * no, hipMemcpyDeviceToHost, st));
    STB_TRY(hipStreamSynchronize(st));
    std::vector<StbPt> kept;
    for (size_t i = 0; i < objs.size(); ++i)
      if (!rep[i]) kept.push_back(objs[i]);
    objs.swap(kept);
  }
  // 5. link short tracks to the objects (STB.cpp:474-577)
  const int n_obj = (int)objs.size();
  if (n_obj > 0) {
    std::vector<int32_t> link(n_sa, kUnlinked);
    if (n_sa > 0) {
      std::vector<StbPt> last(n_sa), nbp(la.size()), nbd(la.size());
      std::vector<int32_t> nbl(la.size());
      for (int i = 0; i < n_sa; ++i) last[i] = sa[i].p.back();
      for (size_t j = 0; j < la.size(); ++j) {
        const size_t L = la[j].p.size();
        const StbPt &b = la[j].p[L - 2];
        nbp[j] = StbPt{(double)(float)b.x, (double)(float)b.y, (double)(float)b.z};
        const StbPt &a = la[j].p[L - 1];
        nbd[j] = StbPt{a.x - b.x, a.y - b.y, a.z - b.z};
        nbl[j] = (int32_t)L;
      }
      STB_CHECK(upload_pts(s->last, last, st));
      STB_CHECK(upload_pts(s->nb_pt, nbp, st));
      STB_CHECK(upload_pts(s->nb_disp, nbd, st));
      STB_CHECK(s->nb_len.need(sizeof(int32_t) * std::max<size_t>(1, nbl.size())));
      if (!nbl.empty())
        STB_TRY(hipMemcpyAsync(s->nb_len.p, nbl.data(), sizeof(int32_t) * nbl.size(),
                               hipMemcpyHostToDevice, st));
      STB_CHECK(upload_pts(s->cand_d, objs, st));
      STB_CHECK(s->link.need(sizeof(int32_t) * n_sa));
      k_link_short<<<n_sa, kLinkThreads, 0, st>>>(
          s->last.as<double>(), n_sa, s->nb_pt.as<double>(), s->nb_disp.as<double>(),
          s->nb_len.as<int32_t>(), (int)la.size(), s->cand_d.as<double>(), n_obj,
          s->sp.radius_search_track_mm, 5, s->link.as<int32_t>());
      STB_TRY(hipGetLastError());
      STB_TRY(hipMemcpyAsync(link.data(), s->link.p, sizeof(int32_t) * n_sa, hipMemcpyDeviceToHost, st));
      STB_TRY(hipStreamSynchronize(st));
    }
    // first come, first served (STB.cpp:517-533)
    std::vector<int> owner(n_obj, -1);
    for (int j = 0; j < n_sa; ++j) {
      const int lid = link[j];
      if (lid == kUnlinked || lid < 0 || lid >= n_obj) continue;
      if (owner[lid] == -1) owner[lid] = j;
      else link[j] = kUnlinked;
    }
    std::vector<char> used(n_obj, 0);
    std::vector<StbTrack> keep;
    for (int i = 0; i < n_sa; ++i) {
      const int lid = link[i];
      if (lid != kUnlinked && lid >= 0 && lid < n_obj && !used[lid]) {
        used[lid] = 1;
        sa[i].add(objs[lid], frame);
        if (sa[i].t.size() >= (size_t)s->sp.n_initial_frames) la.push_back(std::move(sa[i]));
        else keep.push_back(std::move(sa[i]));
      }
    }
    sa.swap(keep);
    for (int i = 0; i < n_obj; ++i)
      if (!used[i]) {
        StbTrack t;
        t.add(objs[i], frame);
        sa.push_back(std::move(t));
      }
  }
  // 6. linear-fit pruning of the active long tracks (STB.cpp:582-622)
  const int n_la2 = (int)la.size();
  if (n_la2 > 0) {
    int n_pts = s->sp.n_initial_frames > 4 ? 4 : s->sp.n_initial_frames;
    const std::vector<double> K = linfit_kernel(n_pts);
    std::vector<double> w((size_t)n_la2 * kFitWin * 3, 0.0);
    std::vector<int32_t> ln(n_la2);
    for (int i = 0; i < n_la2; ++i) {
      const size_t L = la[i].p.size();
      ln[i] = (int32_t)la[i].t.size();
      if (L < (size_t)n_pts) continue;
      for (int j = 0; j < n_pts; ++j) {
        const StbPt &q = la[i].p[L - n_pts + j];
        w[((size_t)i * kFitWin + j) * 3] = q.x;
        w[((size_t)i * kFitWin + j) * 3 + 1] = q.y;
        w[((size_t)i * kFitWin + j) * 3 + 2] = q.z;
      }
    }
    STB_CHECK(s->win.need(sizeof(double) * w.size()));
    STB_CHECK(s->len.need(sizeof(int32_t) * n_la2));
    STB_CHECK(s->kern.need(sizeof(double) * K.size()));
    STB_CHECK(s->okb.need(sizeof(int32_t) * n_la2));
    STB_TRY(hipMemcpyAsync(s->win.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, st));
    STB_TRY(hipMemcpyAsync(s->len.p, ln.data(), sizeof(int32_t) * n_la2, hipMemcpyHostToDevice, st));
    STB_TRY(hipMemcpyAsync(s->kern.p, K.data(), sizeof(double) * K.size(), hipMemcpyHostToDevice, st));
    k_lin_fit<<<blocks_for(n_la2), 256, 0, st>>>(s->win.as<double>(), s->len.as<int32_t>(), n_la2,
                                                 s->kern.as<double>(), n_pts, s->okb.as<int32_t>());
    STB_TRY(hipGetLastError());
    std::vector<int32_t> ok(n_la2);
    STB_TRY(hipMemcpyAsync(ok.data(), s->okb.p, sizeof(int32_t) * n_la2, hipMemcpyDeviceToHost, st));
    STB_TRY(hipStreamSynchronize(st));
    std::vector<StbTrack> keep;
    for (int i = 0; i < n_la2; ++i) {
      if (!ok[i]) {
        if (la[i].p.size() >= (size_t)kLenLongTrack) li.push_back(std::move(la[i]));
      } else {
        keep.push_back(std::move(la[i]));
      }
    }
    la.swap(keep);
  }
  // 7. snapshot every 500 frames (STB.cpp:687-694)
  if (frame % 500 == 0) {
    if (!s->out.empty()) STB_CHECK(save_all(s, s->out + "ConvergeTrack/", frame));
    li.clear();
    ex.clear();
  }
  return OLPT_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

olpt_status olpt_stb_create(olpt_ctx *ctx, const olpt_stb_params *sp, const olpt_ipr_params *ip,
                            const char *output_path, olpt_stb **out) {
  if (!ctx || !sp || !ip || !out) return set_error(OLPT_E_INVALID_ARGUMENT, "null argument");
  CtxInfo ci;
  if (!ctx_info(ctx, &ci)) return set_error(OLPT_E_INVALID_ARGUMENT, "context has no cameras or workspace");
  if (sp->frame_end < sp->frame_start)
    return set_error(OLPT_E_INVALID_ARGUMENT, "Invalid frame range in config");
  if (!pf_params_ok(*sp)) return set_error(OLPT_E_INVALID_ARGUMENT, "bad predictive-field parameters");
  auto *s = new olpt_stb();
  s->ctx = ctx;
  s->ci = ci;
  s->sp = *sp;
  if (s->sp.n_initial_frames < 4) s->sp.n_initial_frames = 4;  // Config.cpp:197-199
  s->ip = *ip;
  for (int k = 0; k < ci.n_cam; ++k) s->hcam.push_back(make_cam<double>(ci.cams[k]));
  if (output_path && *output_path) {
    s->out = output_path;
    if (s->out.back() != '/' && s->out.back() != '\\') s->out.push_back('/');
    if (mkdirs(s->out) || mkdirs(s->out + "InitialTrack/") || mkdirs(s->out + "ConvergeTrack/")) {
      delete s;
      return set_error(OLPT_E_IO, "cannot create the output folders");
    }
  }
  *out = s;
  return OLPT_OK;
}

olpt_status olpt_stb_set_ipr_model(olpt_stb *s, const olpt_mlp *mlp, double voxel_mm,
                                   double max_step_vox, const olpt_ipr_params *ipr) {
  if (!s) return set_error(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (mlp && !ipr) return set_error(OLPT_E_INVALID_ARGUMENT, "null IPR parameters");
  if (mlp && s->ci.n_cam != 4)
    return set_error(OLPT_E_UNSUPPORTED, "the learned shake proposal takes exactly 4 cameras");
  if (mlp && !(voxel_mm > 0)) return set_error(OLPT_E_INVALID_ARGUMENT, "voxel_mm must be positive");
  s->ipr_mlp = mlp;
  s->ipr_mlp_vox = voxel_mm;
  s->ipr_mlp_step = max_step_vox;
  if (mlp) s->ip_ipr = *ipr;
  return OLPT_OK;
}

olpt_status olpt_stb_destroy(olpt_stb *s) {
  delete s;
  return OLPT_OK;
}

olpt_status olpt_stb_process_frame(olpt_stb *s, int frame_id, const void *d_imgs, int dtype) {
  if (!s || !d_imgs) return set_error(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (frame_id < s->sp.frame_start || frame_id > s->sp.frame_end)
    return set_error(OLPT_E_OUT_OF_RANGE, "Frame number is out of range");
  if (!ctx_info(s->ctx, &s->ci)) return set_error(OLPT_E_INVALID_ARGUMENT, "context not ready");
  olpt_status st;
  if (frame_id - s->sp.frame_start < s->sp.n_initial_frames) {
    if ((size_t)(frame_id - s->sp.frame_start) != s->cand.size())
      return set_error(OLPT_E_INVALID_ARGUMENT, "initial-phase frames must be processed in order");
    st = init_phase(s, frame_id, d_imgs, dtype);
  } else {
    st = conv_phase(s, frame_id, d_imgs, dtype);
  }
  if (st) return st;
  if (frame_id == s->sp.frame_end && frame_id % 500 != 0 && !s->out.empty())
    STB_CHECK(save_all(s, s->out + "ConvergeTrack/", frame_id));
  return OLPT_OK;
}

olpt_status olpt_stb_track_count(olpt_stb *s, int which, int32_t *n_tracks, int32_t *n_points) {
  if (!s || which < 0 || which > 3 || !n_tracks || !n_points)
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  *n_tracks = (int32_t)s->lists[which].size();
  long np = 0;
  for (const auto &t : s->lists[which]) np += (long)t.size();
  *n_points = (int32_t)np;
  return OLPT_OK;
}

olpt_status olpt_stb_get_tracks(olpt_stb *s, int which, int32_t cap, int32_t *track, int32_t *frame,
                                double *xyz) {
  if (!s || which < 0 || which > 3 || !track || !frame || !xyz)
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  long r = 0;
  const auto &l = s->lists[which];
  for (size_t i = 0; i < l.size(); ++i)
    for (size_t j = 0; j < l[i].size(); ++j) {
      if (r >= cap) return set_error(OLPT_E_CAPACITY, "track rows exceed the output capacity");
      track[r] = (int32_t)i;
      frame[r] = l[i].t[j];
      xyz[3 * r] = l[i].p[j].x;
      xyz[3 * r + 1] = l[i].p[j].y;
      xyz[3 * r + 2] = l[i].p[j].z;
      ++r;
    }
  return OLPT_OK;
}

olpt_status olpt_stb_set_tracks(olpt_stb *s, int which, int32_t n_rows, const int32_t *track,
                                const int32_t *frame, const double *xyz) {
  if (!s || which < 0 || which > 3 || n_rows < 0 || (n_rows && (!track || !frame || !xyz)))
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  std::vector<StbTrack> l;
  for (int r = 0; r < n_rows; ++r) {
    const int tid = track[r];
    if (tid < 0) continue;
    while ((int)l.size() <= tid) l.emplace_back();
    l[tid].add(StbPt{xyz[3 * r], xyz[3 * r + 1], xyz[3 * r + 2]}, frame[r]);
  }
  // track ids missing from the rows would leave empty tracks, which the
  // reference cannot process either; they are not kept
  l.erase(std::remove_if(l.begin(), l.end(), [](const StbTrack &t) { return t.size() == 0; }),
          l.end());
  s->lists[which].swap(l);
  return OLPT_OK;
}

olpt_status olpt_stb_save_tracks(olpt_stb *s, const char *folder, int frame) {
  if (!s || !folder) return set_error(OLPT_E_INVALID_ARGUMENT, "null argument");
  return save_all(s, folder, frame);
}

olpt_status olpt_stb_load_tracks(olpt_stb *s, const char *folder, int frame) {
  if (!s || !folder) return set_error(OLPT_E_INVALID_ARGUMENT, "null argument");
  std::string dir = folder;
  const struct {
    const char *pre;
    int which;
  } items[] = {{"LongTrackActive_", OLPT_TRACK_LONG_ACTIVE},
               {"ShortTrackActive_", OLPT_TRACK_SHORT_ACTIVE}};
  for (const auto &it : items) {
    const std::string path = dir + it.pre + std::to_string(frame) + ".csv";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return set_error(OLPT_E_IO, ("LoadTracks: missing file:" + path).c_str());
    std::vector<int32_t> tid, fr;
    std::vector<double> xyz;
    char line[8192];
    bool header = true;
    while (std::fgets(line, sizeof(line), f)) {
      if (header) {
        header = false;
        continue;
      }
      int t, fi;
      double x, y, z;
      if (std::sscanf(line, "%d,%d,%lf,%lf,%lf", &t, &fi, &x, &y, &z) != 5) continue;
      tid.push_back(t);
      fr.push_back(fi);
      xyz.push_back(x);
      xyz.push_back(y);
      xyz.push_back(z);
    }
    std::fclose(f);
    STB_CHECK(olpt_stb_set_tracks(s, it.which, (int32_t)tid.size(), tid.data(), fr.data(), xyz.data()));
  }
  return OLPT_OK;
}

olpt_status olpt_stb_predict(olpt_ctx *ctx, const double *d_pos, const int32_t *d_len, int n_tr,
                             double *d_out, int32_t *d_ok) {
  CtxInfo ci;
  if (!ctx_info(ctx, &ci)) return set_error(OLPT_E_INVALID_ARGUMENT, "context not ready");
  if (n_tr < 0 || (n_tr && (!d_pos || !d_len || !d_out || !d_ok)))
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  if (n_tr == 0) return OLPT_OK;
  k_predict<<<blocks_for(3L * n_tr), 256, 0, ci.st>>>(d_pos, d_len, n_tr, d_out, d_ok);
  STB_TRY(hipGetLastError());
  STB_TRY(hipStreamSynchronize(ci.st));
  return OLPT_OK;
}

olpt_status olpt_stb_linear_fit(olpt_ctx *ctx, int n_initial_frames, const double *d_pos,
                                const int32_t *d_len, int n_tr, int32_t *d_ok) {
  CtxInfo ci;
  if (!ctx_info(ctx, &ci)) return set_error(OLPT_E_INVALID_ARGUMENT, "context not ready");
  if (n_tr < 0 || (n_tr && (!d_pos || !d_len || !d_ok)))
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  if (n_tr == 0) return OLPT_OK;
  const int n_pts = n_initial_frames > 4 ? 4 : n_initial_frames;
  if (n_pts < 2) return set_error(OLPT_E_INVALID_ARGUMENT, "n_initial_frames must be at least 2");
  const std::vector<double> K = linfit_kernel(n_pts);
  DBuf kb;
  STB_CHECK(kb.need(sizeof(double) * K.size()));
  STB_TRY(hipMemcpyAsync(kb.p, K.data(), sizeof(double) * K.size(), hipMemcpyHostToDevice, ci.st));
  k_lin_fit<<<blocks_for(n_tr), 256, 0, ci.st>>>(d_pos, d_len, n_tr, kb.as<double>(), n_pts, d_ok);
  STB_TRY(hipGetLastError());
  STB_TRY(hipStreamSynchronize(ci.st));
  return OLPT_OK;
}

olpt_status olpt_pred_field(olpt_ctx *ctx, const olpt_stb_params *sp, const double *d_prev,
                            int n_prev, const double *d_curr, int n_curr, double *d_field,
                            const double *d_q, int n_q, double *d_disp) {
  CtxInfo ci;
  if (!ctx_info(ctx, &ci)) return set_error(OLPT_E_INVALID_ARGUMENT, "context not ready");
  if (!sp || n_prev < 0 || n_curr < 0 || n_q < 0 || (n_prev && !d_prev) || (n_curr && !d_curr) ||
      (n_q && (!d_q || !d_disp)))
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  if (!pf_params_ok(*sp)) return set_error(OLPT_E_INVALID_ARGUMENT, "bad predictive-field parameters");
  PfWork w;
  STB_CHECK(pred_field(*sp, d_prev, n_prev, d_curr, n_curr, w, ci.st));
  const PfGrid g = make_grid(*sp);
  const long ng = (long)g.nx * g.ny * g.nz;
  if (d_field)
    STB_TRY(hipMemcpyAsync(d_field, w.field.p, sizeof(double) * 3 * ng, hipMemcpyDeviceToDevice, ci.st));
  if (n_q > 0) k_pf_query<<<blocks_for(n_q), 256, 0, ci.st>>>(g, w.field.as<double>(), d_q, n_q, d_disp);
  STB_TRY(hipGetLastError());
  STB_TRY(hipStreamSynchronize(ci.st));
  return OLPT_OK;
}

olpt_status olpt_estimate_otf(olpt_ctx *ctx, const void *d_frames, int dtype, int n_img,
                              const olpt_ipr_params *p, double *abca, int32_t *ok) {
  CtxInfo ci;
  if (!ctx_info(ctx, &ci)) return set_error(OLPT_E_INVALID_ARGUMENT, "context not ready");
  if (!d_frames || !p || !abca || !ok || n_img < 1)
    return set_error(OLPT_E_INVALID_ARGUMENT, "bad argument");
  const size_t es = dtype == OLPT_U8 ? 1 : dtype == OLPT_U16 ? 2 : dtype == OLPT_F32 ? 4 : 8;
  if (dtype < OLPT_U8 || dtype > OLPT_F64) return set_error(OLPT_E_UNSUPPORTED, "unsupported image dtype");
  const int H = ci.H, W = ci.W;
  const long HW = (long)H * W;
  const int cap = 1 << 20;
  DBuf xy;
  STB_CHECK(xy.need(sizeof(double) * 2 * cap));
  std::vector<double> img(HW);
  std::vector<unsigned char> raw(HW * es);
  const double r_pt = p->radius_obj;
  for (int k = 0; k < ci.n_cam; ++k) {
    ok[k] = 1;
    std::vector<std::vector<double>> pts(n_img);
    std::vector<std::vector<double>> imgs(n_img);
    for (int j = 0; j < n_img; ++j) {
      const char *src = (const char *)d_frames + ((size_t)j * ci.n_cam + k) * HW * es;
      int32_t n = 0;
      STB_CHECK(olpt_find_peaks(ctx, src, dtype, H, W, p, xy.as<double>(), cap, &n));
      std::vector<double> h(2 * (size_t)n);
      if (n) STB_TRY(hipMemcpy(h.data(), xy.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
      // at most 1000 tracers per image, std::shuffle with seed 123 (OTF.cpp:38-56)
      if (n > 1000) {
        std::vector<int> idx(n);
        for (int i = 0; i < n; ++i) idx[i] = i;
        std::shuffle(idx.begin(), idx.end(), std::default_random_engine(123));
        std::vector<double> h2(2000);
        for (int i = 0; i < 1000; ++i) {
          h2[2 * i] = h[2 * idx[i]];
          h2[2 * i + 1] = h[2 * idx[i] + 1];
        }
        h.swap(h2);
      } else if (n == 0) {
        ok[k] = 0;
        break;
      }
      pts[j] = std::move(h);
      STB_TRY(hipMemcpy(raw.data(), src, HW * es, hipMemcpyDeviceToHost));
      imgs[j].resize(HW);
      for (long q = 0; q < HW; ++q) {
        switch (dtype) {
        case OLPT_U8: imgs[j][q] = raw[q]; break;
        case OLPT_U16: imgs[j][q] = ((const uint16_t *)raw.data())[q]; break;
        case OLPT_F32: imgs[j][q] = ((const float *)raw.data())[q]; break;
        default: imgs[j][q] = ((const double *)raw.data())[q]; break;
        }
      }
    }
    if (!ok[k]) continue;
    // OTF.cpp:58-131
    std::vector<double> I, a_list, c0, c1;
    for (int i = 0; i < n_img; ++i) {
      const std::vector<double> &im = imgs[i];
      for (size_t j = 0; j < pts[i].size() / 2; ++j) {
        const double xc = pts[i][2 * j], yc = pts[i][2 * j + 1];
        const int yc0 = std::clamp<int>(int(std::round(yc)), 0, H - 1);
        const int xc0 = std::clamp<int>(int(std::round(xc)), 0, W - 1);
        a_list.push_back(im[(long)yc0 * W + xc0]);
        const int rmin = std::max(0, int(std::floor(yc - r_pt)));
        const int rmax = std::min(H, int(std::ceil(yc + r_pt + 1)));
        const int cmin = std::max(0, int(std::floor(xc - r_pt)));
        const int cmax = std::min(W, int(std::ceil(xc + r_pt + 1)));
        for (int row = rmin; row < rmax; ++row)
          for (int col = cmin; col < cmax; ++col) {
            I.push_back(im[(long)row * W + col]);
            c0.push_back(std::pow(col - xc, 2));
            c1.push_back(std::pow(row - yc, 2));
          }
      }
    }
    const double a_max = *std::max_element(a_list.begin(), a_list.end());
    const int n_a = (int)a_list.size();
    double a_mean = 0, a_std = 0;
    for (int i = 0; i < n_a; ++i) a_mean += a_list[i];
    a_mean /= n_a;
    for (int i = 0; i < n_a; ++i) a_std += std::pow(a_list[i] - a_mean, 2);
    a_std = std::sqrt(a_std / n_a);
    const double a = std::min(a_mean + 2 * a_std, a_max);
    const size_t nco = I.size();
    std::vector<double> lg(nco);
    for (size_t i = 0; i < nco; ++i) {
      const double logI = I[i] < kLogSmall ? std::log(kLogSmall) : std::log(I[i]);
      lg[i] = -logI + std::log(a);
    }
    // inverse(X'X) X' y in Matrix::operator* order
    std::vector<double> XtX(4, 0.0), inv;
    for (size_t q = 0; q < nco; ++q) XtX[0] += c0[q] * c0[q];
    for (size_t q = 0; q < nco; ++q) XtX[1] += c0[q] * c1[q];
    for (size_t q = 0; q < nco; ++q) XtX[2] += c1[q] * c0[q];
    for (size_t q = 0; q < nco; ++q) XtX[3] += c1[q] * c1[q];
    gauss_inverse_n(XtX, 2, inv);
    double est[2] = {0.0, 0.0};
    for (int r = 0; r < 2; ++r)
      for (size_t q = 0; q < nco; ++q) {
        double t = 0.0;
        t += inv[r * 2] * c0[q];
        t += inv[r * 2 + 1] * c1[q];
        est[r] += t * lg[q];
      }
    abca[4 * k] = a;
    abca[4 * k + 1] = est[0];
    abca[4 * k + 2] = est[1];
    abca[4 * k + 3] = 0.0;
  }
  return OLPT_OK;
}

} // extern "C"
