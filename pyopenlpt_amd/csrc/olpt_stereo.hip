// olpt_stereo.hip - epipolar stereo matching (StereoMatch, StereoMatch.cpp).
//
// Stages (all counts device-resident, fixed grids):
//   k_los         one FP64 line of sight per 2D point (Camera::lineOfSight is
//                 deterministic, so it is computed once instead of per visit)
//   k_cell_*      IDMap as a CSR cell index per camera (StereoMatch.cpp:
//                 1413-1438); within-cell order is irrelevant to the result
//   k_build       one wave64 per reference observation, persistent waves with
//                 dynamic work fetch.  The first epipolar band is enumerated
//                 wave-cooperatively: lanes own cell rows of the strip
//                 (computeStripIntersection, :1475-1566) and append accepted
//                 points with ballot/mbcnt compaction into LDS; back-
//                 projection, triangulation and camera selection then run
//                 lane-parallel over that list.  The deeper build cameras
//                 (buildMatch's DFS, :184-444) are walked breadth-first with
//                 one survivor per lane: each lane enumerates only the cell
//                 rows inside the y-extent of its strip intersection (a
//                 superset of the reference's non-empty spans), so a deeper
//                 search costs a few cells instead of all 256 rows.  The
//                 emitted set does not depend on visiting order.
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
constexpr int kRawCap = 1024;   // level-1 band list (pre back-projection)
constexpr int kLvlCap = 1024;   // survivors per deeper level

// Survivors of level d are the 2D points chosen on the d-th build camera:
// their camera, the next camera to search, and the parent entry (level
// d-1) whose path they extend.  NLV = build_count - 2 non-terminal levels.
template <int NLV> struct BuildLds {
  int32_t raw[kRawCap];
  int32_t pid[NLV > 0 ? NLV : 1][kLvlCap];
  int16_t par[NLV > 0 ? NLV : 1][kLvlCap];
  int8_t cam[NLV > 0 ? NLV : 1][kLvlCap];
  int8_t nxt[NLV > 0 ? NLV : 1][kLvlCap];
  int32_t n[NLV + 1];
  int32_t n_raw;
  // reference observation and level-1 strip line
  double la[2], lU[2], ln[2], ld;
};

// Strip line of LOS (pt, u) on camera t, in the normal form the strip
// intersection uses (IDMap::computeStripIntersection, :1475-1566).
struct Strip {
  double ax, ay;   // point on the line
  double Ux, Uy;   // renormalised direction (IDMap::normalized)
  double n0, n1, d;
};

__device__ __forceinline__ Strip make_strip(const CamT<double> &ct, const double pt[3],
                                            const double u[3]) {
  const Line2 l2 = make_line2d(ct, pt, u);
  Strip s;
  s.ax = l2.ax;
  s.ay = l2.ay;
  normalized(l2.ux, l2.uy, s.Ux, s.Uy);
  s.n0 = -s.Uy;
  s.n1 = s.Ux;
  s.d = -(s.n0 * l2.ax + s.n1 * l2.ay);
  return s;
}

// Cell-row span of one row for one strip (same arithmetic as the
// reference); returns false if the row is empty for this strip.
__device__ __forceinline__ bool strip_span(const Strip &s, int cy, int cell_px, int cols,
                                           double tolg, int &a, int &bb) {
  const double cell = (double)cell_px;
  const double hw = 0.5 * cell_px, hh = 0.5 * cell_px;
  const double to_index = 1.0 / cell_px;
  const double yc = (cy + 0.5) * cell;
  const double C = s.n1 * yc + s.d;
  const double pe = fabs(s.n0) * hw + fabs(s.n1) * hh;
  const double T = tolg + pe;
  a = 0;
  bb = cols - 1;
  if (fabs(s.n0) < 1e-12) return !(fabs(C) > T);
  double xl = (-T - C) / s.n0, xr = (T - C) / s.n0;
  if (xl > xr) {
    const double tt = xl;
    xl = xr;
    xr = tt;
  }
  a = (int)ceil(xl * to_index - 0.5 - 1e-9);
  bb = (int)floor(xr * to_index - 0.5 + 1e-9);
  a = a < 0 ? 0 : a;
  bb = bb > cols - 1 ? cols - 1 : bb;
  return a <= bb;
}

// Rows whose cell centre can lie in both widened strips i, j: the
// intersection is a parallelogram around the crossing of the centre lines;
// its y-extent (+ one cell) bounds the non-empty rows.  Superset only.
__device__ __forceinline__ void pair_rows(const Strip &si, const Strip &sj, int cell_px,
                                          double tolg, int rows, int &r0, int &r1) {
  const double det = si.n0 * sj.n1 - si.n1 * sj.n0;
  if (!(fabs(det) >= 1e-6)) return;  // near-parallel: keep the current range
  const double hw = 0.5 * cell_px;
  const double Ti = tolg + (fabs(si.n0) + fabs(si.n1)) * hw;
  const double Tj = tolg + (fabs(sj.n0) + fabs(sj.n1)) * hw;
  // centre crossing: n_i.x = -d_i, n_j.x = -d_j
  const double py = (si.n0 * (-sj.d) - sj.n0 * (-si.d)) / det;
  // y-components of the unit offsets along each strip's normal band
  const double wiy = -sj.n0 / det;  // moves n_i.x by 1, keeps n_j.x
  const double wjy = si.n0 / det;
  const double ext = fabs(Ti * wiy) + fabs(Tj * wjy);
  // clamp in floating point before the integer conversion (far crossings)
  const double lo = fmax(floor((py - ext) / cell_px) - 1.0, -1.0);
  const double hi = fmin(floor((py + ext) / cell_px) + 1.0, (double)rows);
  if (!(lo <= hi)) return;  // NaN geometry: keep the current range
  if ((int)lo > r0) r0 = (int)lo;
  if ((int)hi < r1) r1 = (int)hi;
  if (r0 < 0) r0 = 0;
  if (r1 > rows - 1) r1 = rows - 1;
}

template <int MAXB> struct Path {
  int cam[MAXB];
  int pid[MAXB];
  double pt[MAXB][2];
  double P[MAXB][3];  // LOS origin (camera centre)
  double U[MAXB][3];  // LOS unit direction
};

template <int MAXB>
__device__ __forceinline__ void path_push(Path<MAXB> &P, int k, const StereoBuffers &b,
                                          int cam, int pid) {
  P.cam[k] = cam;
  P.pid[k] = pid;
  const double *q = b.xy + ((long)cam * b.cap + pid) * 2;
  P.pt[k][0] = q[0];
  P.pt[k][1] = q[1];
  const double *u = b.los + ((long)cam * b.cap + pid) * 4;
  for (int c = 0; c < 3; ++c) {
    P.P[k][c] = b.cams[cam].Tinv[c];
    P.U[k][c] = u[c];
  }
}

// Extend path (length k) by point pid on camera t: back-projection
// (checkBackProjection, :1206-1227), quick triangulation (:397-407), then
// either checkMatch + emission (terminal) or the next-camera choice.
// Returns 0 reject, 1 emitted, 2 survivor (next camera in *next).
template <int MAXB>
__device__ int extend(const StereoBuffers &b, const StereoParams &p, Path<MAXB> &P, int k,
                      int t, int pid, int *next) {
  const CamT<double> *cams = b.cams;
  const double tol2 = p.tol2d * p.tol2d;
  const double *ut = b.los + ((long)t * b.cap + pid) * 4;
  const double u_t[3] = {ut[0], ut[1], ut[2]};
  for (int j = 0; j < k; ++j) {
    const Line2 l2 = make_line2d(cams[P.cam[j]], cams[t].Tinv, u_t);
    const double d0 = P.pt[j][0] - l2.ax, d1 = P.pt[j][1] - l2.ay;
    const double z = d1 * l2.ux - d0 * l2.uy;
    if (z * z > tol2) return 0;
  }
  path_push<MAXB>(P, k, b, t, pid);
  const int kk = k + 1;
  {
    const double th = min_parallax_deg(P.U, kk);
    const double tq = tri_tol(p.tol3d, kk, p.build_count, th);
    double X[3], e;
    if (!triangulate(P.P, P.U, kk, X, e) || !(e <= tq)) return 0;
  }
  if (kk == p.build_count) {
    // checkMatch, all-active branch (:530-540): LOS in camera order
    double cp[MAXB][3], cu[MAXB][3];
    int32_t ids[kMaxCam];
    int n = 0;
    for (int c = 0; c < b.n_cam; ++c) {
      ids[c] = -1;
      for (int j = 0; j < kk; ++j)
        if (P.cam[j] == c) {
          ids[c] = P.pid[j];
          for (int d = 0; d < 3; ++d) {
            cp[n][d] = P.P[j][d];
            cu[n][d] = P.U[j][d];
          }
          ++n;
        }
    }
    double Xc[3], e;
    if (!triangulate(cp, cu, n, Xc, e)) {
      atomicOr(&b.flags[6], 1);
      return 0;
    }
    if (e > p.tol3d || !in_limit(p.limit, Xc)) return 0;
    const int slot = atomicAdd(b.n_cand, 1);
    if (slot < b.cap_cand) {
      for (int c = 0; c < b.n_cam; ++c) b.cand_ids[(long)slot * b.n_cam + c] = ids[c];
      b.cand_e[slot] = e;
    } else {
      atomicOr(&b.flags[0], 1);
    }
    return 1;
  }
  int rem[kMaxCam], nrem = 0;
  for (int c = 0; c < b.n_cam; ++c) {
    if (!p.active[c]) continue;
    bool used = false;
    for (int j = 0; j < kk; ++j) used = used || (P.cam[j] == c);
    if (!used) rem[nrem++] = c;
  }
  *next = select_next(cams, P.P, P.U, kk, rem, nrem);
  return *next < 0 ? 0 : 2;
}

template <int MAXB>
__global__ void __launch_bounds__(kWave) k_build(StereoBuffers b, StereoParams p) {
  constexpr int NLV = MAXB - 2;
  __shared__ BuildLds<NLV> L;
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
  const int rows = b.rows_cell, cols = b.cols_cell;
  const long ncell = (long)rows * cols;
  const double tolg = p.tol2d + 1e-6;
  const double tol2 = p.tol2d * p.tol2d;
  while (true) {
    int r = 0;
    if (lane == 0) r = atomicAdd(b.work, 1);
    r = __shfl(r, 0, kWave);
    if (r >= nref) break;
    Path<MAXB> P0;
    path_push<MAXB>(P0, 0, b, ref, r);
    int rem[kMaxCam], nrem = 0;
    for (int c = 0; c < b.n_cam; ++c)
      if (p.active[c] && c != ref) rem[nrem++] = c;
    if (nrem == 0) continue;
    const int t1 = select_second(cams, P0.P[0], P0.U[0], rem, nrem);
    if (t1 < 0) continue;
    // ---- level 1: one band, wave-cooperative enumeration over cell rows
    const Strip s1 = make_strip(cams[t1], P0.P[0], P0.U[0]);
    if (lane == 0) {
      L.n_raw = 0;
      for (int d = 0; d <= NLV; ++d) L.n[d] = 0;
    }
    __syncthreads();
    {
      const int32_t *cstart = b.cell_start + t1 * ncell;
      const int32_t *ccnt = b.cell_cnt + t1 * ncell;
      const int32_t *cpid = b.cell_pid + (long)t1 * b.cap;
      const double *xy = b.xy + (long)t1 * b.cap * 2;
      for (int cy0 = 0; cy0 < rows; cy0 += kWave) {
        const int cy = cy0 + lane;
        int xa = 0, xb = -1;
        bool valid = cy < rows && strip_span(s1, cy, b.cell_px, cols, tolg, xa, xb);
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
            const double dx = xy[2 * pid] - s1.ax, dy = xy[2 * pid + 1] - s1.ay;
            const double z = dy * s1.Ux - dx * s1.Uy;
            acc = !(z * z > tol2);
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
    }
    __syncthreads();
    const int n_raw = L.n_raw < kRawCap ? L.n_raw : kRawCap;
    // level-1 extension, lane-parallel over the band list
    for (int i = lane; i < n_raw; i += kWave) {
      int next = -1;
      const int res = extend<MAXB>(b, p, P0, 1, t1, L.raw[i], &next);
      if (res == 2 && NLV > 0) {
        const int pos = atomicAdd(&L.n[0], 1);
        if (pos < kLvlCap) {
          L.pid[0][pos] = L.raw[i];
          L.par[0][pos] = -1;
          L.cam[0][pos] = (int8_t)t1;
          L.nxt[0][pos] = (int8_t)next;
        } else {
          atomicOr(&b.flags[1], 1);
        }
      }
    }
    __syncthreads();
    // ---- deeper levels: one survivor per lane, lane-serial enumeration of
    // the (small) strip intersection on its next camera
    for (int lv = 1; lv <= NLV; ++lv) {
      const int n_in = L.n[lv - 1] < kLvlCap ? L.n[lv - 1] : kLvlCap;
      for (int e = lane; e < n_in; e += kWave) {
        // rebuild the path: ref + one entry per level
        Path<MAXB> P = P0;
        int chain[MAXB];
        int q = e;
        for (int d = lv - 1; d >= 0; --d) {
          chain[d] = q;
          q = L.par[d][q];
        }
        for (int d = 0; d < lv; ++d)
          path_push<MAXB>(P, d + 1, b, L.cam[d][chain[d]], L.pid[d][chain[d]]);
        const int k = lv + 1;              // path length
        const int t = L.nxt[lv - 1][e];    // camera to search
        Strip S[MAXB];
        for (int j = 0; j < k; ++j) S[j] = make_strip(cams[t], P.P[j], P.U[j]);
        int r0 = 0, r1 = rows - 1;
        for (int j = 1; j < k; ++j) pair_rows(S[0], S[j], b.cell_px, tolg, rows, r0, r1);
        const int32_t *cstart = b.cell_start + t * ncell;
        const int32_t *ccnt = b.cell_cnt + t * ncell;
        const int32_t *cpid = b.cell_pid + (long)t * b.cap;
        const double *xy = b.xy + (long)t * b.cap * 2;
        for (int cy = r0; cy <= r1; ++cy) {
          int xa = 0, xb = cols - 1;
          bool ok = true;
          for (int j = 0; j < k && ok; ++j) {
            int a, bb;
            ok = strip_span(S[j], cy, b.cell_px, cols, tolg, a, bb);
            if (ok) {
              xa = a > xa ? a : xa;
              xb = bb < xb ? bb : xb;
              ok = xa <= xb;
            }
          }
          if (!ok) continue;
          for (int cx = xa; cx <= xb; ++cx) {
            const int cell = cy * cols + cx;
            const int s0 = cstart[cell], s1n = s0 + ccnt[cell];
            for (int kk = s0; kk < s1n; ++kk) {
              const int pid = cpid[kk];
              const double qx = xy[2 * pid], qy = xy[2 * pid + 1];
              bool in = true;
              for (int j = 0; j < k && in; ++j) {
                const double dx = qx - S[j].ax, dy = qy - S[j].ay;
                const double z = dy * S[j].Ux - dx * S[j].Uy;
                in = !(z * z > tol2);
              }
              if (!in) continue;
              int next = -1;   // extend() only writes slot k of the path
              const int res = extend<MAXB>(b, p, P, k, t, pid, &next);
              if (res == 2 && lv < NLV) {
                const int pos = atomicAdd(&L.n[lv], 1);
                if (pos < kLvlCap) {
                  L.pid[lv][pos] = pid;
                  L.par[lv][pos] = (int16_t)e;
                  L.cam[lv][pos] = (int8_t)t;
                  L.nxt[lv][pos] = (int8_t)next;
                } else {
                  atomicOr(&b.flags[1], 1);
                }
              }
            }
          }
        }
      }
      __syncthreads();
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
  const int waves = 256 * 12;
  stage_begin(ST_BUILD, st);
  switch (p.build_count) {
  case 2: k_build<2><<<waves, kWave, 0, st>>>(b, p); break;
  case 3: k_build<3><<<waves, kWave, 0, st>>>(b, p); break;
  case 4: k_build<4><<<waves, kWave, 0, st>>>(b, p); break;
  case 5: k_build<5><<<waves, kWave, 0, st>>>(b, p); break;
  case 6: k_build<6><<<waves, kWave, 0, st>>>(b, p); break;
  case 7: k_build<7><<<waves, kWave, 0, st>>>(b, p); break;
  default: k_build<8><<<waves, kWave, 0, st>>>(b, p); break;
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
