// olpt_stereo.hip - epipolar stereo matching (StereoMatch, StereoMatch.cpp).
//
// Stages (all counts device-resident, fixed grids):
//   k_los         one FP64 line of sight per 2D point (Camera::lineOfSight is
//                 deterministic, so it is computed once instead of per visit)
//   k_cell_*      IDMap as a CSR cell index per camera (StereoMatch.cpp:
//                 1413-1438); within-cell order is irrelevant to the result
//   k_build       one wave64 per reference observation, persistent waves with
//                 dynamic work fetch.  Band enumeration is wave-cooperative:
//                 lanes own cell rows of the strip intersection
//                 (computeStripIntersection, :1475-1566) and append accepted
//                 points with ballot/mbcnt compaction into LDS; the back-
//                 projection, triangulation, camera selection and checkMatch
//                 tests then run lane-parallel over the compacted list.
//                 The DFS over build cameras (buildMatch, :184-444) walks the
//                 LDS level lists; the emitted set does not depend on order.
//   k_key_* / k_bucket_sort  lexicographic sort of the id vectors
//                 (:153-156): counting sort on the first active camera's id,
//                 then insertion sort inside each (small) bucket.
//   k_flat / k_pct / k_mis_round / k_replace_check / k_tri  pruneMatch
//                 (:631-980) + triangulateMatch (:985-1050).  The greedy
//                 packing in (pct, e_check, index) order equals the greedy
//                 maximal independent set of the conflict graph under that
//                 priority, computed in rounds: a candidate is selected once
//                 every higher-priority conflicting candidate is rejected and
//                 rejected once one of them is selected.  The reference's
//                 replacement branch is checked afterwards for every rejected
//                 candidate (flags[3] counts cases where it would have fired;
//                 with the sort order it cannot).
#include "olpt_kernels.hpp"
#include "olpt_scan.hpp"

namespace olpt {

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// --------------------------------------------------------------------------
// small kernels
// --------------------------------------------------------------------------
__global__ void k_zero_i32(int32_t *p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0;
}

static void zero_i32(int32_t *p, long n, hipStream_t st) {
  if (n <= 0) return;
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_zero_i32<<<(unsigned)blocks, 256, 0, st>>>(p, n);
}

__global__ void k_los(StereoBuffers b, StereoParams p) {
  const int cam = blockIdx.y;
  if (!p.active[cam]) return;
  const int n = b.npts[cam];
  const CamT<double> c = b.cams[cam];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const double *q = b.xy + ((long)cam * b.cap + i) * 2;
    double u[3];
    line_of_sight(c, q[0], q[1], u);
    double *o = b.los + ((long)cam * b.cap + i) * 4;
    o[0] = u[0];
    o[1] = u[1];
    o[2] = u[2];
    o[3] = 0.0;
  }
}

__global__ void k_cell_count(StereoBuffers b, StereoParams p) {
  const int cam = blockIdx.y;
  if (!p.active[cam]) return;
  const int n = b.npts[cam];
  const long ncell = (long)b.rows_cell * b.cols_cell;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const double *q = b.xy + ((long)cam * b.cap + i) * 2;
    const int cx = (int)floor(q[0] / b.cell_px);
    const int cy = (int)floor(q[1] / b.cell_px);
    int cell = -1, rank = 0;
    if (!(cx < 0 || cy < 0 || cx >= b.cols_cell || cy >= b.rows_cell)) {
      cell = cy * b.cols_cell + cx;
      rank = atomicAdd(&b.cell_cnt[cam * ncell + cell], 1);
    }
    b.pt_cell[(long)cam * b.cap + i] = cell;
    b.pt_rank[(long)cam * b.cap + i] = rank;
  }
}

__global__ void k_cell_fill(StereoBuffers b, StereoParams p) {
  const int cam = blockIdx.y;
  if (!p.active[cam]) return;
  const int n = b.npts[cam];
  const long ncell = (long)b.rows_cell * b.cols_cell;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const int cell = b.pt_cell[(long)cam * b.cap + i];
    if (cell < 0) continue;
    const int pos = b.cell_start[cam * ncell + cell] + b.pt_rank[(long)cam * b.cap + i];
    b.cell_pid[(long)cam * b.cap + pos] = i;
  }
}

// --------------------------------------------------------------------------
// geometry helpers (FP64, reference operation order)
// --------------------------------------------------------------------------
struct Line2 {
  double ax, ay;   // point (projection of the LOS origin)
  double ux, uy;   // unit direction (createUnitVector)
};

// StereoMatch::makeLine2DFromLOS3D (StereoMatch.cpp:1053-1066)
__device__ __forceinline__ Line2 make_line2d(const CamT<double> &c,
                                             const double pt[3],
                                             const double u[3]) {
  double a[2], bq[2];
  project(c, pt, a[0], a[1]);
  const double q[3] = {pt[0] + u[0], pt[1] + u[1], pt[2] + u[2]};
  project(c, q, bq[0], bq[1]);
  const double d0 = bq[0] - a[0], d1 = bq[1] - a[1];
  const double n = sqrt(d0 * d0 + d1 * d1);
  return Line2{a[0], a[1], d0 / n, d1 / n};
}

// IDMap::normalized (StereoMatch.cpp:1442-1451)
__device__ __forceinline__ void normalized(double ux, double uy, double &ox,
                                           double &oy) {
  const double n2 = ux * ux + uy * uy;
  if (n2 <= 1e-24) {
    ox = 1.0;
    oy = 0.0;
    return;
  }
  const double inv = 1.0 / sqrt(n2);
  ox = ux * inv;
  oy = uy * inv;
}

// StereoMatch::computeMinParallaxDeg + calTriangulateTol (:1082-1126)
__device__ inline double min_parallax_deg(const double (*u)[3], int n) {
  if (n < 2) return 180.0;
  double mn = 180.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double dot = u[i][0] * u[j][0] + u[i][1] * u[j][1] + u[i][2] * u[j][2];
      dot = fmax(-1.0, fmin(1.0, dot));
      const double deg = acos(dot) * 180.0 / M_PI;
      if (deg < mn) mn = deg;
    }
  return mn;
}

__device__ inline double tri_tol(double tol, int k, int m, double th_deg) {
  const int kk = k > 2 ? k : 2, mm = m > 2 ? m : 2;
  const double ad = sqrt((double)mm / (double)kk);
  const double th = fmax(th_deg, 1.0);
  const double num = sin(6.0 * M_PI / 180.0);
  const double den = fmax(1e-6, sin(th * M_PI / 180.0));
  const double aa = fmax(1.0, fmin(2.0, num / den));
  return tol * ad * aa;
}

// StereoMatch::selectSecondCameraByLineLength (:1307-1370)
__device__ inline int select_second(const CamT<double> *cams, const double pt[3],
                                    const double u[3], const int *rem, int nrem) {
  int best = -1;
  double best_len = INFINITY;
  for (int ri = 0; ri < nrem; ++ri) {
    const int cam = rem[ri];
    const Line2 L = make_line2d(cams[cam], pt, u);
    const double px = L.ax, py = L.ay, ux = L.ux, uy = L.uy;
    const double W = (double)cams[cam].n_col, H = (double)cams[cam].n_row;
    double qs[4][2];
    int np = 0;
    auto push_if = [&](double x, double y) {
      if (x >= 0.0 && x <= W && y >= 0.0 && y <= H) {
        qs[np][0] = x;
        qs[np][1] = y;
        ++np;
      }
    };
    if (fabs(ux) > 1e-12) {
      double t = (0.0 - px) / ux;
      push_if(0.0, py + t * uy);
      t = (W - 1 - px) / ux;
      push_if(W - 1, py + t * uy);
    }
    if (fabs(uy) > 1e-12) {
      double t = (0.0 - py) / uy;
      push_if(px + t * ux, 0.0);
      t = (H - 1 - py) / uy;
      push_if(px + t * ux, H - 1);
    }
    double uq[4][2];
    int nu = 0;
    for (int i = 0; i < np; ++i) {
      bool dup = false;
      for (int j = 0; j < nu; ++j)
        if (fabs(uq[j][0] - qs[i][0]) < 1e-9 && fabs(uq[j][1] - qs[i][1]) < 1e-9) dup = true;
      if (!dup) {
        uq[nu][0] = qs[i][0];
        uq[nu][1] = qs[i][1];
        ++nu;
      }
    }
    double len = 0.0;
    if (nu >= 2) {
      const double dx = uq[0][0] - uq[1][0], dy = uq[0][1] - uq[1][1];
      len = sqrt(dx * dx + dy * dy);
    }
    if (len < best_len - 1e-9) {
      best_len = len;
      best = cam;
    }
  }
  return best;
}

// StereoMatch::selectNextCameraByMaxPairAngle (:1375-1411)
__device__ inline int select_next(const CamT<double> *cams, const double (*pt)[3],
                                  const double (*u)[3], int n, const int *rem,
                                  int nrem) {
  int best = -1;
  double best_score = -1.0;
  for (int ri = 0; ri < nrem; ++ri) {
    const int cam = rem[ri];
    double lx[kMaxCam], ly[kMaxCam];
    for (int i = 0; i < n; ++i) {
      const Line2 L = make_line2d(cams[cam], pt[i], u[i]);
      lx[i] = L.ux;
      ly[i] = L.uy;
    }
    double score = 0.0;
    if (n >= 2) {
      for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
          const double s = fabs(lx[i] * ly[j] - ly[i] * lx[j]);
          if (s > score) score = s;
          if (score >= 1.0 - 1e-12) break;
        }
        if (score >= 1.0 - 1e-12) break;
      }
    }
    if (score > best_score) {
      best_score = score;
      best = cam;
    }
  }
  return best;
}

__device__ __forceinline__ bool in_limit(const double lim[6], const double X[3]) {
  return !(X[0] > lim[1] || X[0] < lim[0] || X[1] > lim[3] || X[1] < lim[2] ||
           X[2] > lim[5] || X[2] < lim[4]);
}

// --------------------------------------------------------------------------
// k_build: wave per reference observation
// --------------------------------------------------------------------------
constexpr int kRawCap = 1024;  // enumerated (pre back-projection) list
constexpr int kL1Cap = 1024;   // survivors on the 2nd build camera
constexpr int kLnCap = 256;    // survivors on deeper build cameras

template <int NL> struct BuildLds {
  int32_t raw[kRawCap];
  int32_t lvl_pid[NL > 0 ? kL1Cap + (NL - 1) * kLnCap : 1];
  int8_t lvl_next[NL > 0 ? kL1Cap + (NL - 1) * kLnCap : 1];
  // path (build order)
  int32_t pcam[kMaxCam];
  int32_t ppid[kMaxCam];
  double ppt[kMaxCam][2];   // 2D point
  double pu[kMaxCam][3];    // LOS unit direction
  // strip lines on the current target camera
  double la[kMaxCam][2], lU[kMaxCam][2], ln[kMaxCam][2], ld[kMaxCam];
  int32_t n_raw;
};

__device__ __forceinline__ int lvl_base(int lvl) {
  return lvl == 1 ? 0 : kL1Cap + (lvl - 2) * kLnCap;
}
__device__ __forceinline__ int lvl_cap(int lvl) { return lvl == 1 ? kL1Cap : kLnCap; }

// Enumerate points of camera t inside all k strips (the path's LOS
// projected into t); result appended to L.raw.
template <int NL>
__device__ void enumerate_strip(const StereoBuffers &b, const StereoParams &p,
                                BuildLds<NL> &L, int k, int t) {
  const int lane = lane_id();
  const CamT<double> *cams = b.cams;
  if (lane < k) {
    const int c = L.pcam[lane];
    const Line2 l2 = make_line2d(cams[t], cams[c].Tinv, L.pu[lane]);
    double Ux, Uy;
    normalized(l2.ux, l2.uy, Ux, Uy);
    L.la[lane][0] = l2.ax;
    L.la[lane][1] = l2.ay;
    L.lU[lane][0] = Ux;
    L.lU[lane][1] = Uy;
    const double n0 = -Uy, n1 = Ux;
    L.ln[lane][0] = n0;
    L.ln[lane][1] = n1;
    L.ld[lane] = -(n0 * l2.ax + n1 * l2.ay);
  }
  if (lane == 0) L.n_raw = 0;
  __syncthreads();
  const int rows = b.rows_cell, cols = b.cols_cell;
  const long ncell = (long)rows * cols;
  const int32_t *cstart = b.cell_start + t * ncell;
  const int32_t *ccnt = b.cell_cnt + t * ncell;
  const int32_t *cpid = b.cell_pid + (long)t * b.cap;
  const double *xy = b.xy + (long)t * b.cap * 2;
  const double cell = (double)b.cell_px;
  const double hw = 0.5 * b.cell_px, hh = 0.5 * b.cell_px;
  const double to_index = 1.0 / b.cell_px;
  const double tolg = p.tol2d + 1e-6;
  const double tol2 = p.tol2d * p.tol2d;
  for (int cy0 = 0; cy0 < rows; cy0 += kWave) {
    const int cy = cy0 + lane;
    int xa = 0, xb = cols - 1;
    bool valid = cy < rows;
    for (int l = 0; l < k && valid; ++l) {
      const double n0 = L.ln[l][0], n1 = L.ln[l][1];
      const double yc = (cy + 0.5) * cell;
      const double C = n1 * yc + L.ld[l];
      const double pe = fabs(n0) * hw + fabs(n1) * hh;
      const double T = tolg + pe;
      int a = 0, bb = cols - 1;
      if (fabs(n0) < 1e-12) {
        if (fabs(C) > T) valid = false;
      } else {
        double xl = (-T - C) / n0, xr = (T - C) / n0;
        if (xl > xr) {
          const double tt = xl;
          xl = xr;
          xr = tt;
        }
        a = (int)ceil(xl * to_index - 0.5 - 1e-9);
        bb = (int)floor(xr * to_index - 0.5 + 1e-9);
        a = a < 0 ? 0 : a;
        bb = bb > cols - 1 ? cols - 1 : bb;
        if (a > bb) valid = false;
      }
      if (valid) {
        xa = a > xa ? a : xa;
        xb = bb < xb ? bb : xb;
      }
    }
    // walk the lane's cells / points one point per iteration
    int cellid = valid ? cy * cols + xa : 0;
    const int cell_end = valid ? cy * cols + xb : -1;
    int kk = valid ? cstart[cellid] : 0, kend = valid ? kk + ccnt[cellid] : 0;
    while (valid && kk == kend && cellid < cell_end) {
      ++cellid;
      kk = cstart[cellid];
      kend = kk + ccnt[cellid];
    }
    bool has = valid && kk < kend;
    while (__any(has)) {
      bool acc = false;
      int pid = -1;
      if (has) {
        pid = cpid[kk];
        const double qx = xy[2 * pid], qy = xy[2 * pid + 1];
        acc = true;
        for (int l = 0; l < k; ++l) {
          const double dx = qx - L.la[l][0], dy = qy - L.la[l][1];
          const double z = dy * L.lU[l][0] - dx * L.lU[l][1];
          if (z * z > tol2) {
            acc = false;
            break;
          }
        }
        ++kk;
        while (kk == kend && cellid < cell_end) {
          ++cellid;
          kk = cstart[cellid];
          kend = kk + ccnt[cellid];
        }
        has = kk < kend;
      }
      const uint64_t m = ballot(acc);
      if (m) {
        const int base = L.n_raw;
        const int pos = base + mbcnt(m);
        if (acc) {
          if (pos < kRawCap)
            L.raw[pos] = pid;
          else
            atomicOr(&b.flags[1], 1);
        }
        __syncthreads();
        if (lane == 0) L.n_raw = base + (int)__popcll(m);
        __syncthreads();
      }
    }
  }
  __syncthreads();
  if (lane == 0 && L.n_raw > kRawCap) L.n_raw = kRawCap;
  __syncthreads();
}

// Lane-parallel: back-projection (checkBackProjection :1206-1227), quick
// triangulation (:397-407), then terminal checkMatch + emit, or next-camera
// selection and append to level list `lvl`.
template <int NL>
__device__ void filter_extend(const StereoBuffers &b, const StereoParams &p,
                              BuildLds<NL> &L, int k, int t, int lvl,
                              int32_t *n_lvl) {
  const int lane = lane_id();
  const CamT<double> *cams = b.cams;
  const double tol2 = p.tol2d * p.tol2d;
  const bool terminal = (k + 1 == p.build_count);
  if (lane == 0 && !terminal) n_lvl[lvl] = 0;
  __syncthreads();
  const int n_raw = L.n_raw;
  for (int i0 = 0; i0 < n_raw; i0 += kWave) {
    const int i = i0 + lane;
    bool keep = false;
    int pid = -1, next = -1;
    int32_t ids_out[kMaxCam];
    double e_out = 0.0;
    if (i < n_raw) {
      pid = L.raw[i];
      const double *q = b.xy + ((long)t * b.cap + pid) * 2;
      const double *ut = b.los + ((long)t * b.cap + pid) * 4;
      double u_t[3] = {ut[0], ut[1], ut[2]};
      keep = true;
      for (int j = 0; j < k && keep; ++j) {
        const Line2 l2 = make_line2d(cams[L.pcam[j]], cams[t].Tinv, u_t);
        const double d0 = L.ppt[j][0] - l2.ax, d1 = L.ppt[j][1] - l2.ay;
        const double z = d1 * l2.ux - d0 * l2.uy;
        if (z * z > tol2) keep = false;
      }
      if (keep) {
        double pts[kMaxCam][3], us[kMaxCam][3];
        for (int j = 0; j < k; ++j)
          for (int c = 0; c < 3; ++c) {
            pts[j][c] = cams[L.pcam[j]].Tinv[c];
            us[j][c] = L.pu[j][c];
          }
        for (int c = 0; c < 3; ++c) {
          pts[k][c] = cams[t].Tinv[c];
          us[k][c] = u_t[c];
        }
        const double th = min_parallax_deg(us, k + 1);
        const double tq = tri_tol(p.tol3d, k + 1, p.build_count, th);
        double X[3], e;
        if (!triangulate(pts, us, k + 1, X, e) || !(e <= tq)) keep = false;
        if (keep && terminal) {
          // checkMatch, all-active branch (:530-540): LOS in camera order
          double cp[kMaxCam][3], cu[kMaxCam][3];
          int n = 0;
          for (int c = 0; c < b.n_cam; ++c) ids_out[c] = -1;
          for (int j = 0; j < k; ++j) ids_out[L.pcam[j]] = L.ppid[j];
          ids_out[t] = pid;
          for (int c = 0; c < b.n_cam; ++c) {
            if (ids_out[c] < 0) continue;
            const double *uu = (c == t) ? u_t : nullptr;
            int jj = -1;
            for (int j = 0; j < k; ++j)
              if (L.pcam[j] == c) jj = j;
            for (int d = 0; d < 3; ++d) {
              cp[n][d] = cams[c].Tinv[d];
              cu[n][d] = uu ? uu[d] : L.pu[jj][d];
            }
            ++n;
          }
          double Xc[3];
          if (!triangulate(cp, cu, n, Xc, e_out)) {
            atomicOr(&b.flags[6], 1);
            keep = false;
          } else if (e_out > p.tol3d || !in_limit(p.limit, Xc)) {
            keep = false;
          }
        } else if (keep) {
          int rem[kMaxCam], nrem = 0;
          for (int c = 0; c < b.n_cam; ++c) {
            if (!p.active[c] || c == t) continue;
            bool used = false;
            for (int j = 0; j < k; ++j) used = used || (L.pcam[j] == c);
            if (!used) rem[nrem++] = c;
          }
          next = select_next(cams, pts, us, k + 1, rem, nrem);
          if (next < 0) keep = false;
        }
      }
    }
    const uint64_t m = ballot(keep);
    if (!m) continue;
    if (terminal) {
      int base = 0;
      if (lane == 0) base = atomicAdd(b.n_cand, (int)__popcll(m));
      base = __shfl(base, 0, kWave);
      if (keep) {
        const int slot = base + mbcnt(m);
        if (slot < b.cap_cand) {
          for (int c = 0; c < b.n_cam; ++c) b.cand_ids[(long)slot * b.n_cam + c] = ids_out[c];
          b.cand_e[slot] = e_out;
        } else {
          atomicOr(&b.flags[0], 1);
        }
      }
    } else {
      const int base = n_lvl[lvl];
      if (keep) {
        const int pos = base + mbcnt(m);
        if (pos < lvl_cap(lvl)) {
          L.lvl_pid[lvl_base(lvl) + pos] = pid;
          L.lvl_next[lvl_base(lvl) + pos] = (int8_t)next;
        } else {
          atomicOr(&b.flags[1], 1);
        }
      }
      __syncthreads();
      if (lane == 0) {
        const int nn = base + (int)__popcll(m);
        n_lvl[lvl] = nn < lvl_cap(lvl) ? nn : lvl_cap(lvl);
      }
      __syncthreads();
    }
  }
  __syncthreads();
}

template <int NL>
__global__ void __launch_bounds__(kWave) k_build(StereoBuffers b, StereoParams p) {
  __shared__ BuildLds<NL> L;
  __shared__ int32_t n_lvl[NL + 2];
  __shared__ int32_t cursor[NL + 2];
  __shared__ int32_t target[NL + 2];
  const int lane = lane_id();
  // reference camera: fewest observations among active cameras, lowest id
  // on ties (StereoMatch.cpp:70-85)
  int ref = -1;
  int best = 0x7fffffff;
  for (int c = 0; c < b.n_cam; ++c)
    if (p.active[c] && b.npts[c] < best) {
      best = b.npts[c];
      ref = c;
    }
  const int nref = best;
  const CamT<double> *cams = b.cams;
  while (true) {
    int r = 0;
    if (lane == 0) r = atomicAdd(b.work, 1);
    r = __shfl(r, 0, kWave);
    if (r >= nref) break;
    if (lane == 0) {
      L.pcam[0] = ref;
      L.ppid[0] = r;
      const double *q = b.xy + ((long)ref * b.cap + r) * 2;
      L.ppt[0][0] = q[0];
      L.ppt[0][1] = q[1];
      const double *u = b.los + ((long)ref * b.cap + r) * 4;
      L.pu[0][0] = u[0];
      L.pu[0][1] = u[1];
      L.pu[0][2] = u[2];
    }
    __syncthreads();
    int rem[kMaxCam], nrem = 0;
    for (int c = 0; c < b.n_cam; ++c)
      if (p.active[c] && c != ref) rem[nrem++] = c;
    if (nrem == 0) continue;
    const int t1 = select_second(cams, cams[ref].Tinv, L.pu[0], rem, nrem);
    if (t1 < 0) continue;
    enumerate_strip<NL>(b, p, L, 1, t1);
    filter_extend<NL>(b, p, L, 1, t1, 1, n_lvl);
    if (NL == 0) continue;
    // DFS over level lists
    int d = 1;
    if (lane == 0) {
      cursor[1] = 0;
      target[1] = t1;
    }
    __syncthreads();
    while (d >= 1) {
      if (cursor[d] >= n_lvl[d]) {
        --d;
        continue;
      }
      const int e = lvl_base(d) + cursor[d];
      const int pid = L.lvl_pid[e];
      const int nt = L.lvl_next[e];
      const int tc = target[d];
      __syncthreads();
      if (lane == 0) {
        cursor[d] = cursor[d] + 1;
        L.pcam[d] = tc;
        L.ppid[d] = pid;
        const double *q = b.xy + ((long)tc * b.cap + pid) * 2;
        L.ppt[d][0] = q[0];
        L.ppt[d][1] = q[1];
        const double *u = b.los + ((long)tc * b.cap + pid) * 4;
        L.pu[d][0] = u[0];
        L.pu[d][1] = u[1];
        L.pu[d][2] = u[2];
      }
      __syncthreads();
      const int k = d + 1;
      enumerate_strip<NL>(b, p, L, k, nt);
      filter_extend<NL>(b, p, L, k, nt, d + 1, n_lvl);
      if (k + 1 < p.build_count && d + 1 <= NL && n_lvl[d + 1] > 0) {
        if (lane == 0) {
          cursor[d + 1] = 0;
          target[d + 1] = nt;
        }
        __syncthreads();
        ++d;
      }
    }
  }
}

// --------------------------------------------------------------------------
// lexicographic sort of candidate id vectors
// --------------------------------------------------------------------------
__device__ __forceinline__ int primary_cam(const StereoParams &p, int n_cam) {
  for (int c = 0; c < n_cam; ++c)
    if (p.active[c]) return c;
  return 0;
}

__global__ void k_key_hist(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int pc = primary_cam(p, b.n_cam);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    const int key = b.cand_ids[(long)j * b.n_cam + pc];
    b.cand_rank[j] = atomicAdd(&b.key_cnt[key], 1);
  }
}

__global__ void k_key_scatter(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int pc = primary_cam(p, b.n_cam);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    const int key = b.cand_ids[(long)j * b.n_cam + pc];
    const int pos = b.key_start[key] + b.cand_rank[j];
    for (int c = 0; c < b.n_cam; ++c)
      b.cand_ids_s[(long)pos * b.n_cam + c] = b.cand_ids[(long)j * b.n_cam + c];
    b.cand_e_s[pos] = b.cand_e[j];
  }
}

__device__ __forceinline__ bool ids_less(const int32_t *a, const int32_t *c, int n) {
  for (int i = 0; i < n; ++i)
    if (a[i] != c[i]) return a[i] < c[i];
  return false;
}

__global__ void k_bucket_sort(StereoBuffers b, StereoParams p) {
  const int pc = primary_cam(p, b.n_cam);
  const int nkey = b.npts[pc];
  const int nc = b.n_cam;
  for (int key = blockIdx.x * blockDim.x + threadIdx.x; key < nkey;
       key += gridDim.x * blockDim.x) {
    const int s = b.key_start[key], n = b.key_cnt[key];
    if (n < 2) continue;
    int32_t *ids = b.cand_ids_s + (long)s * nc;
    double *e = b.cand_e_s + s;
    for (int i = 1; i < n; ++i) {
      int32_t cur[kMaxCam];
      for (int c = 0; c < nc; ++c) cur[c] = ids[(long)i * nc + c];
      const double ce = e[i];
      int j = i - 1;
      while (j >= 0 && ids_less(cur, ids + (long)j * nc, nc)) {
        for (int c = 0; c < nc; ++c) ids[(long)(j + 1) * nc + c] = ids[(long)j * nc + c];
        e[j + 1] = e[j];
        --j;
      }
      for (int c = 0; c < nc; ++c) ids[(long)(j + 1) * nc + c] = cur[c];
      e[j + 1] = ce;
    }
  }
}

// --------------------------------------------------------------------------
// pruneMatch
// --------------------------------------------------------------------------
__device__ __forceinline__ bool eq_eps(double a, double c) {
  const double s = fmax(fabs(a), fabs(c));
  return fabs(a - c) <= (s * 1e-9 + 1e-12);
}

// order of StereoMatch.cpp:820-851: pct (eq_eps), e_check (eq_eps), index
__device__ __forceinline__ bool less_order(const double *pct, const double *e,
                                           int aj, int bj) {
  const double pa = pct[aj], pb = pct[bj];
  const bool af = isfinite(pa), bf = isfinite(pb);
  if (af && bf) {
    if (!eq_eps(pa, pb)) return pa < pb;
  } else if (af) {
    return true;
  } else if (bf) {
    return false;
  }
  const double ea = e[aj], eb = e[bj];
  const bool aef = isfinite(ea), bef = isfinite(eb);
  if (aef && bef) {
    if (!eq_eps(ea, eb)) return ea < eb;
  } else if (aef) {
    return true;
  } else if (bef) {
    return false;
  }
  return aj < bj;
}

__global__ void k_flat(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    int k = 0;
    for (int c = 0; c < nc; ++c) {
      if (!p.active[c]) continue;
      const int pid = b.cand_ids_s[(long)j * nc + c];
      if (pid < 0) continue;
      const long f = (long)c * b.cap + pid;
      b.flat_rank[(long)j * nc + k] = atomicAdd(&b.pt_count[f], 1);
      ++k;
    }
  }
}

__global__ void k_pt_fill(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    int k = 0;
    for (int c = 0; c < nc; ++c) {
      if (!p.active[c]) continue;
      const int pid = b.cand_ids_s[(long)j * nc + c];
      if (pid < 0) continue;
      const long f = (long)c * b.cap + pid;
      b.pt_list[b.pt_start[f] + b.flat_rank[(long)j * nc + k]] = j;
      ++k;
    }
  }
}

// slot of camera c among the candidate's used cameras (camera order)
__device__ __forceinline__ int flat_slot(const StereoBuffers &b,
                                         const StereoParams &p, int j, int cam) {
  int k = 0;
  for (int c = 0; c < cam; ++c)
    if (p.active[c] && b.cand_ids_s[(long)j * b.n_cam + c] >= 0) ++k;
  return k;
}

// Percentile of each candidate on each used 2D point (StereoMatch.cpp:
// 727-800): stable sort of the point's candidates by e_check (eq_eps ties by
// index), average rank for eq_eps blocks, rank/(n-1).
__global__ void k_pct(StereoBuffers b, StereoParams p) {
  const long total = (long)b.n_cam * b.cap;
  for (long f = blockIdx.x * (long)blockDim.x + threadIdx.x; f < total;
       f += (long)gridDim.x * blockDim.x) {
    const int n = b.pt_count[f];
    if (n == 0) continue;
    const int cam = (int)(f / b.cap);
    int32_t *lst = b.pt_list + b.pt_start[f];
    if (n == 1) {
      const int j = lst[0];
      b.pv[(long)j * b.n_cam + flat_slot(b, p, j, cam)] = 0.0;
      continue;
    }
    // insertion sort by (e eq_eps, index)
    for (int i = 1; i < n; ++i) {
      const int cur = lst[i];
      const double ce = b.cand_e_s[cur];
      int k = i - 1;
      while (k >= 0) {
        const int o = lst[k];
        const double oe = b.cand_e_s[o];
        const bool less = !eq_eps(ce, oe) ? (ce < oe) : (cur < o);
        if (!less) break;
        lst[k + 1] = o;
        --k;
      }
      lst[k + 1] = cur;
    }
    int k = 0;
    while (k < n) {
      int k0 = k, k1 = k + 1;
      const double e0 = b.cand_e_s[lst[k0]];
      while (k1 < n && eq_eps(b.cand_e_s[lst[k1]], e0)) ++k1;
      const double avg_rank = 0.5 * (double)(k0 + (k1 - 1));
      const double denom = (double)(n - 1);
      const double pv = (denom > 0.0) ? (avg_rank / denom) : 0.0;
      for (int t = k0; t < k1; ++t) {
        const int j = lst[t];
        b.pv[(long)j * b.n_cam + flat_slot(b, p, j, cam)] = pv;
      }
      k = k1;
    }
  }
}

__global__ void k_pct_final(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    int nf = 0;
    double s = 0.0;
    for (int c = 0; c < nc; ++c)
      if (p.active[c] && b.cand_ids_s[(long)j * nc + c] >= 0) {
        s += b.pv[(long)j * nc + nf];
        ++nf;
      }
    b.pct[j] = nf > 0 ? s / (double)nf : INFINITY;
    b.state[j] = 0;
  }
}

// One round of the priority greedy.  In-place updates are safe: a decided
// state never changes, and an undecided neighbour only delays a decision.
__global__ void k_mis_round(StereoBuffers b, StereoParams p, int round) {
  if (round > 0 && b.undecided[round - 1] == 0) return;
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  int und = 0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    if (b.state[j] != 0) continue;
    bool out = false, blocked = false;
    for (int c = 0; c < nc && !out; ++c) {
      if (!p.active[c]) continue;
      const int pid = b.cand_ids_s[(long)j * nc + c];
      if (pid < 0) continue;
      const long f = (long)c * b.cap + pid;
      const int s0 = b.pt_start[f], n = b.pt_count[f];
      for (int t = 0; t < n; ++t) {
        const int m = b.pt_list[s0 + t];
        if (m == j) continue;
        if (!less_order(b.pct, b.cand_e_s, m, j)) continue;
        const int sm = __atomic_load_n(&b.state[m], __ATOMIC_RELAXED);
        if (sm == 1) {
          out = true;
          break;
        }
        if (sm == 0) blocked = true;
      }
    }
    if (out)
      b.state[j] = 2;
    else if (!blocked)
      b.state[j] = 1;
    else
      ++und;
  }
  und = wave_sum(und);
  if (lane_id() == 0 && und) atomicAdd(&b.undecided[round], und);
}

__global__ void k_rounds_check(StereoBuffers b, int max_rounds) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && b.undecided[max_rounds - 1] > 0)
    atomicOr(&b.flags[5], 1);
}

// Replacement branch of the reference (StereoMatch.cpp:906-968), evaluated
// for each rejected candidate against the owners it met at its turn.
__global__ void k_replace_check(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    if (b.state[j] != 2) continue;
    int S[kMaxCam], ns = 0;
    for (int c = 0; c < nc; ++c) {
      if (!p.active[c]) continue;
      const int pid = b.cand_ids_s[(long)j * nc + c];
      if (pid < 0) continue;
      const long f = (long)c * b.cap + pid;
      const int s0 = b.pt_start[f], n = b.pt_count[f];
      for (int t = 0; t < n; ++t) {
        const int m = b.pt_list[s0 + t];
        if (m == j || b.state[m] != 1 || !less_order(b.pct, b.cand_e_s, m, j)) continue;
        bool seen = false;
        for (int q = 0; q < ns; ++q) seen = seen || S[q] == m;
        if (!seen && ns < kMaxCam) S[ns++] = m;
      }
    }
    double ap = 0.0, ae = 0.0;
    int cnt = 0;
    for (int q = 0; q < ns; ++q) {
      const double po = b.pct[S[q]], eo = b.cand_e_s[S[q]];
      if (!isfinite(po) || !isfinite(eo)) continue;
      ap += po;
      ae += eo;
      ++cnt;
    }
    if (cnt == 0) continue;
    ap /= (double)cnt;
    ae /= (double)cnt;
    const double pj = b.pct[j], ej = b.cand_e_s[j];
    const bool better_pct = (!eq_eps(pj, ap)) && ((pj + 1e-12) < ap);
    const bool tie_better = eq_eps(pj, ap) && (!eq_eps(ej, ae)) && ((ej + 1e-12) < ae);
    if (better_pct || tie_better) atomicAdd(&b.flags[3], 1);
  }
}

// triangulateMatch: camera-order LOS, FP64, gate on tol_3d
__global__ void k_tri(StereoBuffers b, StereoParams p) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  int nsel = 0;
  // entries past M are cleared: the keep flags are scanned over cap_cand
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < b.cap_cand;
       j += gridDim.x * blockDim.x) {
    int keep = 0;
    if (j < M && b.state[j] == 1) {
      ++nsel;
      double pt[kMaxCam][3], u[kMaxCam][3];
      int n = 0;
      for (int c = 0; c < nc; ++c) {
        const int pid = b.cand_ids_s[(long)j * nc + c];
        if (pid < 0) continue;
        const double *uu = b.los + ((long)c * b.cap + pid) * 4;
        for (int d = 0; d < 3; ++d) {
          pt[n][d] = b.cams[c].Tinv[d];
          u[n][d] = uu[d];
        }
        ++n;
      }
      double X[3], e;
      if (n >= 2 && triangulate(pt, u, n, X, e) && !(e > p.tol3d)) {
        keep = 1;
        b.tri_xyz[(long)j * 3 + 0] = X[0];
        b.tri_xyz[(long)j * 3 + 1] = X[1];
        b.tri_xyz[(long)j * 3 + 2] = X[2];
      }
    }
    b.keep[j] = keep;
  }
  nsel = wave_sum(nsel);
  if (lane_id() == 0 && nsel) atomicAdd(b.n_sel, nsel);
}

__global__ void k_obj_write(StereoBuffers b) {
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    if (!b.keep[j]) continue;
    const int o = b.keep_off[j];
    if (o >= b.cap_obj) {
      atomicOr(&b.flags[4], 1);
      continue;
    }
    for (int d = 0; d < 3; ++d) b.obj_xyz[(long)o * 3 + d] = b.tri_xyz[(long)j * 3 + d];
    for (int c = 0; c < nc; ++c) b.obj_ids[(long)o * nc + c] = b.cand_ids_s[(long)j * nc + c];
  }
}

__global__ void k_obj_count(StereoBuffers b) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int t = b.totals[0];
    *b.n_obj = t < b.cap_obj ? t : b.cap_obj;
  }
}

// --------------------------------------------------------------------------
// launcher
// --------------------------------------------------------------------------
static int grid_for(long n, int threads, int max_blocks) {
  long g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

void launch_stereo(const StereoBuffers &b, const StereoParams &p,
                   int max_rounds, hipStream_t st) {
  const long ncell = (long)b.rows_cell * b.cols_cell;
  const int T = 256;
  // counters
  zero_i32(b.n_cand, 1, st);
  zero_i32(b.n_sel, 1, st);
  zero_i32(b.work, 1, st);
  zero_i32(b.undecided, max_rounds, st);
  zero_i32(b.cell_cnt, ncell * b.n_cam, st);
  // LOS + cell index
  dim3 gc(grid_for(b.cap, T, 1024), b.n_cam);
  k_los<<<gc, T, 0, st>>>(b, p);
  k_cell_count<<<gc, T, 0, st>>>(b, p);
  scan_exclusive(b.cell_cnt, b.cell_start, ncell, b.n_cam, ncell, nullptr,
                 b.scan_scratch, st);
  k_cell_fill<<<gc, T, 0, st>>>(b, p);
  // build
  const int NL = p.build_count - 2;
  const int waves = 256 * 12;
  stage_begin(ST_BUILD, st);
  switch (NL < 0 ? 0 : NL) {
  case 0: k_build<0><<<waves, kWave, 0, st>>>(b, p); break;
  case 1: k_build<1><<<waves, kWave, 0, st>>>(b, p); break;
  case 2: k_build<2><<<waves, kWave, 0, st>>>(b, p); break;
  case 3: k_build<3><<<waves, kWave, 0, st>>>(b, p); break;
  case 4: k_build<4><<<waves, kWave, 0, st>>>(b, p); break;
  case 5: k_build<5><<<waves, kWave, 0, st>>>(b, p); break;
  default: k_build<6><<<waves, kWave, 0, st>>>(b, p); break;
  }
  stage_end(ST_BUILD, st);
  stage_begin(ST_PRUNE, st);
  // sort
  const int gM = grid_for(b.cap_cand, T, 2048);
  zero_i32(b.key_cnt, b.cap, st);
  k_key_hist<<<gM, T, 0, st>>>(b, p);
  scan_exclusive(b.key_cnt, b.key_start, b.cap, 1, b.cap, nullptr, b.scan_scratch, st);
  k_key_scatter<<<gM, T, 0, st>>>(b, p);
  k_bucket_sort<<<grid_for(b.cap, T, 1024), T, 0, st>>>(b, p);
  // prune
  const long nflat = (long)b.n_cam * b.cap;
  zero_i32(b.pt_count, nflat, st);
  k_flat<<<gM, T, 0, st>>>(b, p);
  scan_exclusive(b.pt_count, b.pt_start, nflat, 1, nflat, nullptr, b.scan_scratch, st);
  k_pt_fill<<<gM, T, 0, st>>>(b, p);
  k_pct<<<grid_for(nflat, T, 2048), T, 0, st>>>(b, p);
  k_pct_final<<<gM, T, 0, st>>>(b, p);
  for (int r = 0; r < max_rounds; ++r) k_mis_round<<<gM, T, 0, st>>>(b, p, r);
  k_rounds_check<<<1, 1, 0, st>>>(b, max_rounds);
  k_replace_check<<<gM, T, 0, st>>>(b, p);
  // triangulate + compact in candidate order
  k_tri<<<gM, T, 0, st>>>(b, p);
  scan_exclusive(b.keep, b.keep_off, b.cap_cand, 1, b.cap_cand, b.totals,
                 b.scan_scratch, st);
  k_obj_write<<<gM, T, 0, st>>>(b);
  k_obj_count<<<1, 64, 0, st>>>(b);
  stage_end(ST_PRUNE, st);
}

} // namespace olpt
