// olpt_stereo.hip - epipolar stereo matching (StereoMatch, StereoMatch.cpp).
//
// Stages (all counts device-resident, fixed grids):
//   k_los         one FP64 line of sight per 2D point (Camera::lineOfSight is
//                 deterministic, so it is computed once instead of per visit)
//                 and its 2D line on every other camera (the line table)
//   k_cell_*      IDMap as a CSR cell index per camera (StereoMatch.cpp:
//                 1413-1438); within-cell order is irrelevant to the result
//   k_b_*         buildMatch's depth-first search (:184-444) run breadth-
//                 first over all reference observations at once, one kernel
//                 per step and one thread per work item: k_b_second picks
//                 each observation's second camera, k_b_band enumerates its
//                 epipolar band (wave per observation, lanes over cell rows
//                 and then over the rows' points, computeStripIntersection
//                 :1475-1566), k_b_extend
//                 runs back-projection, triangulation and camera selection
//                 on every (partial match, point) item, k_b_enum enumerates
//                 the strip intersection of every survivor on its next
//                 camera, but only the cell rows inside the y-extent of the
//                 intersection (a superset of the reference's non-empty
//                 spans).  Items and survivors carry their whole paths; lists
//                 grow by workgroup batches in LDS; the emitted set does not
//                 depend on visiting order.
//   k_key_* / k_bucket_rank  lexicographic sort of the id vectors
//                 (:153-156): counting sort on the first active camera's id,
//                 then each candidate's place inside its (small) bucket by
//                 counting the members that precede it.
//   k_grp_* / k_pct_* / k_prio_* / k_mis_round / k_replace_check / k_tri
//                 pruneMatch (:631-980) + triangulateMatch (:985-1050).
//                 Per-point orders (by e_check for the percentiles, by the
//                 greedy's priority for the packing) are rank counts, one
//                 thread per list entry: an eq_eps block's average rank
//                 follows from the entry's counts of lower and tied
//                 entries, and priority ranks that form a permutation are
//                 the unique order (the rare point where they do not is
//                 redone sequentially).  The greedy packing
//                 in (pct, e_check, index) order equals the greedy maximal
//                 independent set of the conflict graph under that priority,
//                 computed in rounds over the priority-sorted point lists: a
//                 candidate is selected once every entry ahead of it on its
//                 points is rejected and rejected once one of them is
//                 selected.  This equals the reference when its comparator
//                 is a strict total order, which k_hz_* verify (flags[2]
//                 otherwise), and when the replacement branch never fires,
//                 which k_replace_check verifies for every rejected
//                 candidate against the owners it met (flags[3] otherwise);
//                 either flag makes the caller report an error.
#include "olpt_kernels.hpp"
#include "olpt_scan.hpp"

namespace olpt {

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// --------------------------------------------------------------------------
// small kernels
// --------------------------------------------------------------------------
__global__ void k_zero_i32(int32_t *p, long n, long fstride) {
  p = fshift(p, (long)blockIdx.z * fstride);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0;
}

// zero n ints at p in every frame of b's batch (p: frame 0's buffer)
static void zero_i32(const StereoBuffers &b, int32_t *p, long n, hipStream_t st) {
  if (n <= 0) return;
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_zero_i32<<<dim3((unsigned)blocks, 1, b.nf), 256, 0, st>>>(p, n, b.fstride);
}

// Several arrays zeroed by one launch (blockIdx.y picks the array): the
// counters and count arrays of one pass are cleared together instead of by a
// launch each.
constexpr int kZeroMax = 12;
struct ZeroList {
  int32_t *p[kZeroMax];
  long n[kZeroMax];
  int count;
};

__global__ void k_zero_list(ZeroList z, long fstride) {
  const int a = blockIdx.y;
  int32_t *p = fshift(z.p[a], (long)blockIdx.z * fstride);
  const long n = z.n[a];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0;
}

static void zero_list(const StereoBuffers &b, const ZeroList &z, hipStream_t st) {
  long most = 0;
  for (int a = 0; a < z.count; ++a) most = std::max(most, z.n[a]);
  if (z.count <= 0 || most <= 0) return;
  const long blocks = std::min(4096L, (most + 255) / 256);
  k_zero_list<<<dim3((unsigned)blocks, z.count, b.nf), 256, 0, st>>>(z, b.fstride);
}

__global__ void k_cell_count(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int cam = blockIdx.y;
  if (!p.active[cam]) return;
  const int n = b.npts[cam];
  const long ncell = (long)b.rows_cell * b.cols_cell;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const double *q = b.xy + ((long)cam * b.cap + i) * 2;
    const int cx = (int)floor(q[0] / b.cell_px);
    const int cy = (int)floor(q[1] / b.cell_px);
    int cell = -1, rank;
    if (!(cx < 0 || cy < 0 || cx >= b.cols_cell || cy >= b.rows_cell)) {
      cell = cy * b.cols_cell + cx;
      rank = atomicAdd(&b.cell_cnt[cam * ncell + cell], 1);
    } else {
      rank = atomicAdd(&b.out_cnt[cam], 1);
    }
    b.pt_cell[(long)cam * b.cap + i] = cell;
    b.pt_rank[(long)cam * b.cap + i] = rank;
  }
}

// A point's position in its camera's cell order: its cell's start plus its
// rank in the cell; points outside the grid follow the last cell's.
__device__ __forceinline__ int point_pos(const StereoBuffers &b, int cam, int i) {
  const long ncell = (long)b.rows_cell * b.cols_cell;
  const int cell = b.pt_cell[(long)cam * b.cap + i];
  const int rank = b.pt_rank[(long)cam * b.cap + i];
  if (cell >= 0) return b.cell_start[cam * ncell + cell] + rank;
  const long last = cam * ncell + ncell - 1;
  return b.cell_start[last] + b.cell_cnt[last] + rank;
}

__global__ void k_cell_fill(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int cam = blockIdx.y;
  if (!p.active[cam]) return;
  const int n = b.npts[cam];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const int pos = point_pos(b, cam, i);
    b.cell_pid[(long)cam * b.cap + pos] = i;
    const double *q = b.xy + ((long)cam * b.cap + i) * 2;
    b.cell_xy[(long)cam * b.cap + pos] = make_double2(q[0], q[1]);
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

// One FP64 line of sight per point and its 2D line on every other camera, in
// the cell order of the point's camera (i is the position).
__global__ void k_los(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int cam = blockIdx.y;
  if (!p.active[cam]) return;
  const int n = b.npts[cam];
  const CamT<double> c = b.cams[cam];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const double2 q = b.cell_xy[(long)cam * b.cap + i];
    double u[3];
    line_of_sight(c, q.x, q.y, u);
    double *o = b.los + ((long)cam * b.cap + i) * 4;
    o[0] = u[0];
    o[1] = u[1];
    o[2] = u[2];
    o[3] = 0.0;
    for (int t = 0; t < b.n_cam; ++t) {
      if (t == cam || !p.active[t]) continue;
      const Line2 l = make_line2d(b.cams[t], c.Tinv, u);
      double *w = b.lines + (((long)cam * b.n_cam + t) * b.cap + i) * 4;
      w[0] = l.ax;
      w[1] = l.ay;
      w[2] = l.ux;
      w[3] = l.uy;
    }
  }
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

// Length of a 2D line's segment inside the image, as measured by
// StereoMatch::selectSecondCameraByLineLength (:1307-1370); l = (ax, ay, ux, uy).
__device__ inline double segment_in_image(const CamT<double> &cam, const double *l) {
  const double px = l[0], py = l[1], ux = l[2], uy = l[3];
  const double W = (double)cam.n_col, H = (double)cam.n_row;
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
  return len;
}

__device__ __forceinline__ bool in_limit(const double lim[6], const double X[3]) {
  return !(X[0] > lim[1] || X[0] < lim[0] || X[1] > lim[3] || X[1] < lim[2] ||
           X[2] > lim[5] || X[2] < lim[4]);
}

// --------------------------------------------------------------------------
// build: breadth-first over all reference observations
// --------------------------------------------------------------------------
// Line table: the 2D line of every observation's LOS on every other camera,
// lines[(c * n_cam + t) * cap + pid] = makeLine2DFromLOS3D(cam t, LOS(c, pid))
// (ax, ay, ux, uy).  The build's strips, back-projection tests and camera
// choices all read lines of this form; each is a pure function of (c, t, pid),
// so computing it once per pass gives the same values as recomputing it.
__device__ __forceinline__ const double *line_of(const StereoBuffers &b, int c, int t,
                                                 int pid) {
  return b.lines + (((long)c * b.n_cam + t) * b.cap + pid) * 4;
}

// Strip line of one LOS on the target camera in the normal form the strip
// intersection uses (IDMap::computeStripIntersection, :1475-1566).
struct Strip {
  double ax, ay;   // point on the line
  double Ux, Uy;   // renormalised direction (IDMap::normalized)
  double n0, n1, d;
  double inv_n0;   // 1 / n0 (the row spans' column bounds)
};

__device__ __forceinline__ Strip strip_of(const double *l) {
  Strip s;
  s.ax = l[0];
  s.ay = l[1];
  normalized(l[2], l[3], s.Ux, s.Uy);
  s.n0 = -s.Uy;
  s.n1 = s.Ux;
  s.d = -(s.n0 * s.ax + s.n1 * s.ay);
  s.inv_n0 = 1.0 / s.n0;
  return s;
}

__device__ __forceinline__ bool in_strip(const Strip &s, double qx, double qy, double tol2) {
  const double dx = qx - s.ax, dy = qy - s.ay;
  const double z = dy * s.Ux - dx * s.Uy;
  return !(z * z > tol2);
}

// Cell-column span of one cell row for one strip: the cells whose centre
// lies within tol + 1e-6 px plus the cell's own extent along the strip
// normal (the reference's IDMap::computeStripIntersection arithmetic).  The
// row-independent terms are taken once per strip (StripRows) and a row costs
// one FMA for the centre line's offset and a product with the normal's
// reciprocal.  A span only chooses which cells are visited: widened by the
// cell extent and 1e-6 px, it holds every cell with a point that passes the
// exact FP64 in_strip test, so its own last-bit rounding cannot change which
// points pass.
struct StripRows {
  double C0, dC, T, inv_n0, to_index;
  bool flat;  // |n0| < 1e-12: the strip runs along the rows
};

__device__ __forceinline__ StripRows strip_rows(const Strip &s, int cell_px, double tolg) {
  StripRows r;
  const double cell = (double)cell_px;
  const double hw = 0.5 * cell_px, hh = 0.5 * cell_px;
  r.to_index = 1.0 / cell_px;
  r.C0 = s.n1 * (0.5 * cell) + s.d;
  r.dC = s.n1 * cell;
  r.T = tolg + (fabs(s.n0) * hw + fabs(s.n1) * hh);
  r.flat = fabs(s.n0) < 1e-12;
  r.inv_n0 = s.inv_n0;
  return r;
}

// returns false if row cy holds no cell of the strip
__device__ __forceinline__ bool row_span(const StripRows &r, int cy, int cols, int &a, int &bb) {
  const double C = fma((double)cy, r.dC, r.C0);
  a = 0;
  bb = cols - 1;
  if (r.flat) return !(fabs(C) > r.T);
  double xl = (-r.T - C) * r.inv_n0, xr = (r.T - C) * r.inv_n0;
  if (xl > xr) {
    const double tt = xl;
    xl = xr;
    xr = tt;
  }
  a = (int)ceil(xl * r.to_index - 0.5 - 1e-9);
  bb = (int)floor(xr * r.to_index - 0.5 + 1e-9);
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
  // y-change per unit of each strip's signed distance
  const double wiy = -sj.n0 / det;
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

// Block queue: a workgroup collects its output records in LDS and claims
// room in the global list with one atomic per flush.  Global atomics on one
// address serialise across the XCDs' L2 slices, so a per-wave atomic on the
// list counter is what limits a list-building kernel; per-block batches
// also make the global writes contiguous.  All threads of the block call
// bq_push (at most R records each between flushes) and bq_flush<R> in the
// same order.
template <typename Rec, int CAP> struct BlockQueue {
  Rec buf[CAP];
  int n, base;
};

template <typename Rec, int CAP>
__device__ __forceinline__ void bq_init(BlockQueue<Rec, CAP> &q) {
  if (threadIdx.x == 0) q.n = 0;
  __syncthreads();
}

// Room for R records per thread is guaranteed by bq_flush<R>'s threshold.
// One LDS atomic per wave: the wave's pushing lanes take consecutive slots
// in lane order.
template <typename Rec, int CAP>
__device__ __forceinline__ void bq_push(BlockQueue<Rec, CAP> &q, bool want, const Rec &r) {
  const uint64_t m = ballot(want);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(&q.n, __popcll(m));
  base = __shfl(base, leader, kWave);
  if (want) q.buf[base + mbcnt(m)] = r;
}

// Flush when fewer than R * blockDim free slots remain (or any, with `all`).
template <int R = 1, typename Rec, int CAP, typename Out>
__device__ __forceinline__ void bq_flush(BlockQueue<Rec, CAP> &q, int32_t *counter, long cap,
                                         int32_t *overflow, bool all, Out out) {
  static_assert(R * 256 <= CAP / 2, "queue too small for R records per thread of 256");
  __syncthreads();
  const int n = q.n;
  if (all ? n > 0 : n > CAP - R * (int)blockDim.x) {
    if (threadIdx.x == 0) q.base = atomicAdd(counter, n);
    __syncthreads();
    const int base = q.base;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      if ((long)base + i < cap)
        out((long)base + i, q.buf[i]);
      else
        atomicOr(overflow, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
  }
}

// bq_push for lanes in divergent control flow (per-lane loops): the
// pushing lanes of a wave take slots with one LDS atomic; a record that
// finds the queue full goes straight to the global list.  Drained by
// bq_drain at the block's barrier points.
template <typename Rec, int CAP, typename Out>
__device__ __forceinline__ void bq_push_any(BlockQueue<Rec, CAP> &q, bool want, const Rec &r,
                                            int32_t *counter, long cap, int32_t *overflow,
                                            Out out) {
  const uint64_t m = ballot(want);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(&q.n, __popcll(m));
  base = __shfl(base, leader, kWave);
  if (want) {
    const int slot = base + mbcnt(m);
    if (slot < CAP) {
      q.buf[slot] = r;
    } else {
      const int g = atomicAdd(counter, 1);
      if (g < cap)
        out((long)g, r);
      else
        atomicOr(overflow, 1);
    }
  }
}

// Write the queued records out once half the queue is used (or any, with
// `all`); every thread of the block calls it at the same point.
template <typename Rec, int CAP, typename Out>
__device__ __forceinline__ void bq_drain(BlockQueue<Rec, CAP> &q, int32_t *counter, long cap,
                                         int32_t *overflow, bool all, Out out) {
  __syncthreads();
  const int n = min(q.n, CAP);
  if (all ? n > 0 : n >= CAP / 2) {
    if (threadIdx.x == 0) q.base = atomicAdd(counter, n);
    __syncthreads();
    const int base = q.base;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      if ((long)base + i < cap)
        out((long)base + i, q.buf[i]);
      else
        atomicOr(overflow, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
  }
}

struct CandRec {
  int32_t ids[kMaxCam];
  double e;
};

constexpr int kQueue = 2048;      // pair queue capacity (records)
constexpr int kPtBatch = 4;       // points tested per thread between flushes
constexpr int kRowsPass = 4;      // cell rows whose ranges are loaded together
// The rank kernels (bucket, percentile and priority ranks) compare one
// element with every other member of its list; they take kRankBatch members
// per step so that the members' dependent loads are in flight together.
// Counts do not depend on the order of the comparisons.
constexpr int kRankBatch = 4;
constexpr int kSurvQueue = 1024;
constexpr int kPathQueue = 1024;
constexpr int kCandQueue = 512;

// Path of a partial match in build order: slot 0 is the reference
// observation, slot k the point chosen on the k-th build camera.
struct Path {
  int cam[kMaxCam];
  int pid[kMaxCam];
  double pt[kMaxCam][2];
  double P[kMaxCam][3];  // LOS origin (camera centre)
  double U[kMaxCam][3];  // LOS unit direction
};

// pid: the point's position in its camera's cell order
__device__ __forceinline__ void path_set(Path &P, int k, const StereoBuffers &b, int cam,
                                         int pid) {
  P.cam[k] = cam;
  P.pid[k] = pid;
  const double2 q = b.cell_xy[(long)cam * b.cap + pid];
  P.pt[k][0] = q.x;
  P.pt[k][1] = q.y;
  const double *u = b.los + ((long)cam * b.cap + pid) * 4;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    P.P[k][c] = b.cams[cam].Tinv[c];
    P.U[k][c] = u[c];
  }
}

// reference camera: fewest observations among active cameras, lowest id on
// ties (StereoMatch.cpp:70-85)
__device__ __forceinline__ int ref_camera(const StereoBuffers &b, const StereoParams &p,
                                          int &nref) {
  int ref = -1;
  int best = 0x7fffffff;
  for (int c = 0; c < b.n_cam; ++c)
    if (p.active[c] && b.npts[c] < best) {
      best = b.npts[c];
      ref = c;
    }
  nref = best;
  return ref;
}

// cos(6.001 deg): a smallest pairwise parallax above 6.001 deg (and below
// 173.999 deg) makes calTriangulateTol's angle factor exactly 1.
constexpr double kCosParallax = 0.9945200708509531;

// Extend path P (length K) by point pid on camera t: back-projection
// (checkBackProjection, :1206-1227), quick triangulation (:397-407), then
// either checkMatch (terminal) or the next-camera choice
// (selectNextCameraByMaxPairAngle, :1375-1411).
// Returns 0 reject, 1 complete match (checkMatch error in *e_out),
// 2 survivor (next camera in *next).
// TERM: K + 1 is the build count (the host picks the instantiation), so
// each level compiles only its own branch.
template <int K, bool TERM>
__device__ int extend(const StereoBuffers &b, const StereoParams &p, Path &P, int t,
                      int pid, int *next, double *e_out) {
  constexpr int KK = K + 1;
  const int B = p.build_count;
  const double tol2 = p.tol2d * p.tol2d;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const double *l = line_of(b, t, P.cam[j], pid);
    const double d0 = P.pt[j][0] - l[0], d1 = P.pt[j][1] - l[1];
    const double z = d1 * l[2] - d0 * l[3];
    if (z * z > tol2) return 0;
  }
  path_set(P, K, b, t, pid);
  {
    // computeMinParallaxDeg: the smallest pairwise angle is the acos of the
    // largest clamped dot product (acos is monotone)
    double mx = -1.0;
#pragma unroll
    for (int i = 0; i < KK; ++i)
#pragma unroll
      for (int j = i + 1; j < KK; ++j) {
        double dot = P.U[i][0] * P.U[j][0] + P.U[i][1] * P.U[j][1] + P.U[i][2] * P.U[j][2];
        dot = fmax(-1.0, fmin(1.0, dot));
        mx = fmax(mx, dot);
      }
    double tq;
    if (mx < kCosParallax && mx > -kCosParallax) {
      const int kk = KK > 2 ? KK : 2, mm = B > 2 ? B : 2;
      tq = p.tol3d * sqrt((double)mm / (double)kk);
    } else {
      tq = tri_tol(p.tol3d, KK, B, acos(mx) * 180.0 / M_PI);
    }
    // only the decision counts: tri_decide settles it unless the error
    // lies within rounding of the tolerance
    auto quick = [&]() {
      int d = tri_decide<KK>(P.P, P.U, tq, mx);
      if (d < 0) {
        double X[3], e;
        d = triangulate_n<KK>(P.P, P.U, X, e) && e <= tq ? 1 : 0;
      }
      return d;
    };
    if constexpr (!TERM) {
      if (!quick()) return 0;
    } else {
      // checkMatch, all-active branch (:530-540): LOS in camera order
      int rank[KK];
#pragma unroll
      for (int j = 0; j < KK; ++j) {
        rank[j] = 0;
#pragma unroll
        for (int i = 0; i < KK; ++i) rank[j] += P.cam[i] < P.cam[j] ? 1 : 0;
      }
      double cp[KK][3], cu[KK][3];
#pragma unroll
      for (int r = 0; r < KK; ++r)
#pragma unroll
        for (int j = 0; j < KK; ++j)
          if (rank[j] == r) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              cp[r][d] = P.P[j][d];
              cu[r][d] = P.U[j][d];
            }
          }
      // checkMatch first: its error is needed anyway, and when it rejects
      // the quick test cannot change the outcome, so one exact
      // triangulation per item is left; flags[6] as when the quick test
      // runs first (only for items that pass it)
      double Xc[3], e;
      const bool okc = triangulate_n<KK>(cp, cu, Xc, e);
      if (okc && (e > p.tol3d || !in_limit(p.limit, Xc))) return 0;
      if (!quick()) return 0;
      if (!okc) {
        atomicOr(&b.flags[6], 1);
        return 0;
      }
      *e_out = e;
      return 1;
    }
  }
  unsigned rem = 0;
  for (int c = 0; c < b.n_cam; ++c) rem |= p.active[c] ? (1u << c) : 0u;
#pragma unroll
  for (int j = 0; j < KK; ++j) rem &= ~(1u << P.cam[j]);
  if (rem == 0) return 0;
  if ((rem & (rem - 1)) == 0) {
    // a single remaining camera: its score (>= 0) beats the initial -1
    *next = __ffs(rem) - 1;
    return 2;
  }
  int best = -1;
  double best_score = -1.0;
  for (int cam = 0; cam < b.n_cam; ++cam) {
    if (!(rem & (1u << cam))) continue;
    double lx[KK], ly[KK];
#pragma unroll
    for (int i = 0; i < KK; ++i) {
      const double *l = line_of(b, P.cam[i], cam, P.pid[i]);
      lx[i] = l[2];
      ly[i] = l[3];
    }
    double score = 0.0;
    bool stop = false;
#pragma unroll
    for (int i = 0; i < KK; ++i)
#pragma unroll
      for (int j = i + 1; j < KK; ++j) {
        if (!stop) {
          const double s = fabs(lx[i] * ly[j] - ly[i] * lx[j]);
          if (s > score) score = s;
          if (score >= 1.0 - 1e-12) stop = true;
        }
      }
    if (score > best_score) {
      best_score = score;
      best = cam;
    }
  }
  *next = best;
  return best < 0 ? 0 : 2;
}

// Second build camera of every reference observation
// (selectSecondCameraByLineLength, :1307-1370): shortest in-image segment.
__global__ void __launch_bounds__(256) k_b_second(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  int nref;
  const int ref = ref_camera(b, p, nref);
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nref;
       r += gridDim.x * blockDim.x) {
    int t1 = -1;
    double best_len = INFINITY;
    for (int cam = 0; cam < b.n_cam; ++cam) {
      if (!p.active[cam] || cam == ref) continue;
      const double len = segment_in_image(b.cams[cam], line_of(b, ref, cam, r));
      if (len < best_len - 1e-9) {
        best_len = len;
        t1 = cam;
      }
    }
    b.t1[r] = (int8_t)t1;
  }
}

// End of the CSR range of cells c0..c1 of a camera: the next cell's start
// (cell_start is an exclusive scan, so it usually shares c0's cache line
// and saves the count array's), the count for the camera's last cell.
__device__ __forceinline__ int cell_range_end(const int32_t *cstart, const int32_t *ccnt,
                                              long ncell, int c1) {
  return c1 + 1 < ncell ? cstart[c1 + 1] : cstart[c1] + ccnt[c1];
}

// The points of cell rows r0..r1 of camera t that lie in all K strips, each
// passed to push(accepted, pid).  kRowsPass rows at a time: their spans
// (FP64 arithmetic), then their CSR ranges loaded together (within a row,
// the cells of a span are consecutive in the CSR index), then the points in
// kPtBatch batches from the cell-ordered ids and coordinates - two
// dependent loads per pass rather than three per row.
template <int K, typename Push>
__device__ __forceinline__ void enum_rows(const StereoBuffers &b, int t, const Strip *S,
                                          const StripRows *SR, int r0, int r1, double tol2,
                                          Push push) {
  const int cols = b.cols_cell;
  const long ncell = (long)b.rows_cell * cols;
  const int32_t *cstart = b.cell_start + t * ncell;
  const int32_t *ccnt = b.cell_cnt + t * ncell;
  const double2 *cxy = b.cell_xy + (long)t * b.cap;
  for (int rb = r0; rb <= r1; rb += kRowsPass) {
    int c0[kRowsPass], c1[kRowsPass];
#pragma unroll
    for (int u = 0; u < kRowsPass; ++u) {
      const int cy = rb + u;
      c0[u] = -1;
      c1[u] = -1;
      if (cy <= r1) {
        int xa = 0, xb = cols - 1;
        bool ok = true;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          int a, bb;
          if (ok) {
            ok = row_span(SR[j], cy, cols, a, bb);
            xa = a > xa ? a : xa;
            xb = bb < xb ? bb : xb;
            ok = ok && xa <= xb;
          }
        }
        if (ok) {
          c0[u] = cy * cols + xa;
          c1[u] = cy * cols + xb;
        }
      }
    }
    int ka[kRowsPass], ke[kRowsPass];
#pragma unroll
    for (int u = 0; u < kRowsPass; ++u) {
      ka[u] = 0;
      ke[u] = 0;
      if (c0[u] >= 0) {
        ka[u] = cstart[c0[u]];
        ke[u] = cell_range_end(cstart, ccnt, ncell, c1[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kRowsPass; ++u) {
      for (int k = ka[u]; k < ke[u]; k += kPtBatch) {
        // coordinates first; the id only of the points that pass (most
        // visited points do not, and each id is a gather of its own)
        double2 q[kPtBatch];
#pragma unroll
        for (int v = 0; v < kPtBatch; ++v)
          q[v] = k + v < ke[u] ? cxy[k + v] : make_double2(0.0, 0.0);
#pragma unroll
        for (int v = 0; v < kPtBatch; ++v) {
          bool acc = k + v < ke[u];
#pragma unroll
          for (int j = 0; j < K; ++j) acc = acc && in_strip(S[j], q[v].x, q[v].y, tol2);
          push(acc, k + v);
        }
      }
    }
  }
}

// Level-1 band: one wave per reference observation (its strip on the second
// camera is wave-uniform).  Per pass of 64 cell rows each lane takes one row:
// its span (FP64, computeStripIntersection's arithmetic) and that span's CSR
// range; a wave scan lays the rows' points end to end, and the wave tests 64
// consecutive points per step, each lane the point at its position (its row
// found by a binary search over the lanes' prefix sums in LDS).  Consecutive
// lanes so read consecutive points of the cell-ordered arrays - coalesced
// loads, where one lane per row run would gather from its own cache lines -
// and the pairs leave in reference order, row by row along the epipolar line,
// which keeps the next stages' gathers coherent too.
__global__ void __launch_bounds__(256) k_b_band(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  int nref;
  const int ref = ref_camera(b, p, nref);
  const int rows = b.rows_cell, cols = b.cols_cell;
  const long ncell = (long)rows * cols;
  const double tolg = p.tol2d + 1e-6;
  const double tol2 = p.tol2d * p.tol2d;
  __shared__ BlockQueue<int2, kQueue> q;
  __shared__ int32_t wpre[256], wka[256];   // per lane: points before its row, its row's start
  bq_init(q);
  auto out = [&](long i, const int2 &v) { b.pairs[i] = v; };
  const int lane = threadIdx.x & (kWave - 1), w0 = threadIdx.x & ~(kWave - 1);
  const int nw = blockDim.x / kWave;
  // block-uniform loop over groups of nw reference observations
  for (int rbase = blockIdx.x * nw; rbase < nref; rbase += gridDim.x * nw) {
    const int r = rbase + (int)threadIdx.x / kWave;
    const int t1 = r < nref ? (int)b.t1[r] : -1;
    if (t1 >= 0) {
      const Strip S = strip_of(line_of(b, ref, t1, r));
      const StripRows SR = strip_rows(S, b.cell_px, tolg);
      const int32_t *cstart = b.cell_start + t1 * ncell;
      const int32_t *ccnt = b.cell_cnt + t1 * ncell;
      const double2 *cxy = b.cell_xy + (long)t1 * b.cap;
      for (int rb = 0; rb < rows; rb += kWave) {
        const int cy = rb + lane;
        int ka = 0, cnt = 0;
        if (cy < rows) {
          int a, bq;
          if (row_span(SR, cy, cols, a, bq)) {
            ka = cstart[cy * cols + a];
            cnt = cell_range_end(cstart, ccnt, ncell, cy * cols + bq) - ka;
          }
        }
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
          const int o = __shfl_up(incl, off, kWave);
          if (lane >= off) incl += o;
        }
        const int total = __shfl(incl, kWave - 1, kWave);
        wpre[threadIdx.x] = incl - cnt;
        wka[threadIdx.x] = ka;
        __builtin_amdgcn_wave_barrier();
        for (int t0 = 0; t0 < total; t0 += kWave) {
          const int jf = t0 + lane;
          // the row: the last lane whose prefix is <= jf
          int o = 0;
#pragma unroll
          for (int step = kWave / 2; step > 0; step >>= 1)
            if (wpre[w0 + o + step] <= jf) o += step;
          bool acc = jf < total;
          int pid = -1;
          if (acc) {
            const int k = wka[w0 + o] + jf - wpre[w0 + o];
            const double2 xy = cxy[k];
            acc = in_strip(S, xy.x, xy.y, tol2);
            pid = k;
          }
          bq_push_any(q, acc, make_int2(r, pid), b.n_pairs, b.cap_pairs, &b.flags[1], out);
        }
        __builtin_amdgcn_wave_barrier();   // the next pass rewrites this wave's LDS words
      }
    }
    bq_drain(q, b.n_pairs, b.cap_pairs, &b.flags[1], false, out);
  }
  bq_drain(q, b.n_pairs, b.cap_pairs, &b.flags[1], true, out);
}

// Deeper bands: one thread per survivor of level K-1 - a path of K points
// and its next camera, read whole from one record - enumerating the cell
// rows of its K-strip intersection on that camera; every point found
// becomes a level-K item (the path with the point in slot K).
template <int K>
__global__ void __launch_bounds__(256) k_b_enum(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int rows = b.rows_cell;
  const double tolg = p.tol2d + 1e-6;
  const double tol2 = p.tol2d * p.tol2d;
  const int n_in = min(b.n_surv[K - 1], (int)b.cap_pairs);
  const int stride = gridDim.x * blockDim.x;
  const PathRec<K> *surv = reinterpret_cast<const PathRec<K> *>(b.path[1]);
  PathRec<K> *items = reinterpret_cast<PathRec<K> *>(b.path[0]);
  __shared__ BlockQueue<PathRec<K>, kPathQueue> q;
  bq_init(q);
  auto out = [&](long i, const PathRec<K> &v) { items[i] = v; };
  for (int base = blockIdx.x * blockDim.x; base < n_in; base += stride) {
    const int s = base + threadIdx.x;
    if (s < n_in) {
      const PathRec<K> rec = surv[s];
      const int t = rec.cam[K];
      Strip S[K];
      StripRows SR[K];
#pragma unroll
      for (int j = 0; j < K; ++j) {
        S[j] = strip_of(line_of(b, rec.cam[j], t, rec.pid[j]));
        SR[j] = strip_rows(S[j], b.cell_px, tolg);
      }
      int r0 = 0, r1 = rows - 1;
#pragma unroll
      for (int j = 1; j < K; ++j) pair_rows(S[0], S[j], b.cell_px, tolg, rows, r0, r1);
      enum_rows<K>(b, t, S, SR, r0, r1, tol2, [&](bool acc, int pid_t) {
        PathRec<K> r = rec;
        r.pid[K] = pid_t;
        bq_push_any(q, acc, r, &b.n_path[K], b.cap_pairs, &b.flags[1], out);
      });
    }
    bq_drain(q, &b.n_path[K], b.cap_pairs, &b.flags[1], false, out);
  }
  bq_drain(q, &b.n_path[K], b.cap_pairs, &b.flags[1], true, out);
}

// Extension of every level-K item (a path of K points and a point under test
// on its next camera): a survivor goes to level K's survivor list as a whole
// path with its next camera, a complete match (K + 1 == build_count) to the
// candidate list.  Level 1's items are the band's (observation, point)
// pairs.  Levels 1 and 2 are held to 168 registers (3 waves per SIMD:
// measured faster than the unconstrained 2 waves, while 4 waves spill);
// deeper levels keep 2 waves.
template <int K> constexpr int extend_waves() { return K <= 2 ? 3 : 2; }

template <int K> union ExtendQueues {
  BlockQueue<CandRec, kCandQueue> c;           // terminal level
  BlockQueue<PathRec<K + 1>, kSurvQueue> s;    // the others
};

template <int K, bool TERM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(extend_waves<K>())))
k_b_extend(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  int nref;
  const int ref = ref_camera(b, p, nref);
  const int n = min(K == 1 ? *b.n_pairs : b.n_path[K], (int)b.cap_pairs);
  constexpr bool terminal = TERM;
  const int stride = gridDim.x * blockDim.x;
  __shared__ ExtendQueues<K> eq;
  if (threadIdx.x == 0) {
    if constexpr (terminal)
      eq.c.n = 0;
    else
      eq.s.n = 0;
  }
  __syncthreads();
  const int nc = b.n_cam;
  auto out_c = [&](long i, const CandRec &v) {
    for (int c = 0; c < nc; ++c) b.cand_ids[i * nc + c] = v.ids[c];
    b.cand_e[i] = v.e;
  };
  const PathRec<K> *items = reinterpret_cast<const PathRec<K> *>(b.path[0]);
  PathRec<K + 1> *surv_out = reinterpret_cast<PathRec<K + 1> *>(b.path[1]);
  auto out_s = [&](long i, const PathRec<K + 1> &v) { surv_out[i] = v; };
  for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const int q = base + threadIdx.x;
    int res = 0, next = -1;
    double e = 0.0;
    Path P;
    if (q < n) {
      int cam[K + 1], pid[K + 1];
      if constexpr (K == 1) {
        const int2 pr = b.pairs[q];
        cam[0] = ref;
        pid[0] = pr.x;
        cam[1] = b.t1[pr.x];
        pid[1] = pr.y;
      } else {
        const PathRec<K> r = items[q];
#pragma unroll
        for (int j = 0; j <= K; ++j) {
          cam[j] = r.cam[j];
          pid[j] = r.pid[j];
        }
      }
#pragma unroll
      for (int j = 0; j < K; ++j) path_set(P, j, b, cam[j], pid[j]);
      res = extend<K, TERM>(b, p, P, cam[K], pid[K], &next, &e);
    }
    if constexpr (terminal) {
      CandRec cr;
      if (res == 1) {
        // compile-time camera index: the record stays in registers
#pragma unroll
        for (int c = 0; c < kMaxCam; ++c) {
          int id = -1;
#pragma unroll
          for (int j = 0; j <= K; ++j) id = P.cam[j] == c ? P.pid[j] : id;
          // the candidate carries the points' original indices
          cr.ids[c] = c < nc && id >= 0 ? b.cell_pid[(long)c * b.cap + id] : -1;
        }
        cr.e = e;
      }
      bq_push(eq.c, res == 1, cr);
      bq_flush(eq.c, b.n_cand, b.cap_cand, &b.flags[0], false, out_c);
    } else {
      PathRec<K + 1> sr;
      if (res == 2) {
#pragma unroll
        for (int j = 0; j <= K; ++j) {
          sr.pid[j] = P.pid[j];
          sr.cam[j] = (int8_t)P.cam[j];
        }
        sr.pid[K + 1] = -1;
        sr.cam[K + 1] = (int8_t)next;
      }
      bq_push(eq.s, res == 2, sr);
      bq_flush(eq.s, &b.n_surv[K], b.cap_pairs, &b.flags[1], false, out_s);
    }
  }
  if constexpr (terminal)
    bq_flush(eq.c, b.n_cand, b.cap_cand, &b.flags[0], true, out_c);
  else
    bq_flush(eq.s, &b.n_surv[K], b.cap_pairs, &b.flags[1], true, out_s);
}

// --------------------------------------------------------------------------
// lexicographic sort of candidate id vectors
// --------------------------------------------------------------------------
__device__ __forceinline__ int primary_cam(const StereoParams &p, int n_cam) {
  for (int c = 0; c < n_cam; ++c)
    if (p.active[c]) return c;
  return 0;
}

__global__ void k_key_hist(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  const int pc = primary_cam(p, b.n_cam);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    const int key = b.cand_ids[(long)j * b.n_cam + pc];
    b.cand_rank[j] = atomicAdd(&b.key_cnt[key], 1);
  }
}

__global__ void k_key_scatter(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  const int pc = primary_cam(p, b.n_cam);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    const int key = b.cand_ids[(long)j * nc + pc];
    const int q = b.key_start[key] + b.cand_rank[j];
    b.bucket_idx[q] = j;
    if (nc == 4)  // one 16-byte row load and store
      reinterpret_cast<int4 *>(b.bucket_ids)[q] = reinterpret_cast<const int4 *>(b.cand_ids)[j];
    else
      for (int c = 0; c < nc; ++c) b.bucket_ids[(long)q * nc + c] = b.cand_ids[(long)j * nc + c];
  }
}

__device__ __forceinline__ bool ids_less(const int32_t *a, const int32_t *c, int n) {
  for (int i = 0; i < n; ++i)
    if (a[i] != c[i]) return a[i] < c[i];
  return false;
}

// Final position of each candidate = bucket start + number of bucket members
// that precede it lexicographically (equal vectors, which the build never
// emits, by emission index so positions stay a permutation).  One thread per
// bucket slot: the members' id vectors sit contiguously in bucket_ids, so
// the threads of one bucket read the same lines.
__global__ void k_bucket_rank(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  const int pc = primary_cam(p, b.n_cam);
  const int nc = b.n_cam;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < M; q += gridDim.x * blockDim.x) {
    const int j = b.bucket_idx[q];
    int32_t idj[kMaxCam];
    if (nc == 4) {
      const int4 v = reinterpret_cast<const int4 *>(b.bucket_ids)[q];
      idj[0] = v.x;
      idj[1] = v.y;
      idj[2] = v.z;
      idj[3] = v.w;
    } else {
      for (int c = 0; c < nc; ++c) idj[c] = b.bucket_ids[(long)q * nc + c];
    }
    const int key = idj[pc];
    const int s = b.key_start[key], n = b.key_cnt[key];
    int r = 0;
    if (nc == 4) {
      // four cameras: the members' id vectors are loaded whole, a batch at
      // a time, and compared without branches
      const int4 *ids4 = reinterpret_cast<const int4 *>(b.bucket_ids);
      for (int t0 = 0; t0 < n; t0 += kRankBatch) {
        int m[kRankBatch];
        int4 im[kRankBatch];
#pragma unroll
        for (int u = 0; u < kRankBatch; ++u) {
          const int t = t0 + u < n ? s + t0 + u : q;
          m[u] = t == q ? j : b.bucket_idx[t];
          im[u] = ids4[t];
        }
#pragma unroll
        for (int u = 0; u < kRankBatch; ++u) {
          const int32_t a[4] = {im[u].x, im[u].y, im[u].z, im[u].w};
          int cmp = 0;
#pragma unroll
          for (int c = 0; c < 4; ++c)
            cmp = cmp != 0 ? cmp : (a[c] < idj[c] ? -1 : (a[c] > idj[c] ? 1 : 0));
          r += m[u] != j && (cmp < 0 || (cmp == 0 && m[u] < j)) ? 1 : 0;
        }
      }
    } else {
      for (int t = s; t < s + n; ++t) {
        if (t == q) continue;
        const int m = b.bucket_idx[t];
        const int32_t *idm = b.bucket_ids + (long)t * nc;
        int cmp = 0;
        for (int c = 0; c < nc && cmp == 0; ++c)
          cmp = idm[c] < idj[c] ? -1 : (idm[c] > idj[c] ? 1 : 0);
        r += (cmp < 0 || (cmp == 0 && m < j)) ? 1 : 0;
      }
    }
    const int pos = s + r;
    if (nc == 4)
      reinterpret_cast<int4 *>(b.cand_ids_s)[pos] = make_int4(idj[0], idj[1], idj[2], idj[3]);
    else
      for (int c = 0; c < nc; ++c) b.cand_ids_s[(long)pos * nc + c] = idj[c];
    b.cand_e_s[pos] = b.cand_e[j];
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
__device__ __forceinline__ bool less_order_v(double pa, double ea, int aj, double pb, double eb,
                                             int bj) {
  const bool af = isfinite(pa), bf = isfinite(pb);
  if (af && bf) {
    if (!eq_eps(pa, pb)) return pa < pb;
  } else if (af) {
    return true;
  } else if (bf) {
    return false;
  }
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

__device__ __forceinline__ bool less_order(const double *pct, const double *e,
                                           int aj, int bj) {
  return less_order_v(pct[aj], e[aj], aj, pct[bj], e[bj], bj);
}

// Point lists: entry = (candidate << 3) | slot, slot = rank of the camera
// among the candidate's used cameras (camera order).
__device__ __forceinline__ int ent_cand(int32_t v) { return v >> 3; }
__device__ __forceinline__ int ent_slot(int32_t v) { return v & 7; }

// The flat point (camera * cap + pid) of each used camera of candidate j in
// camera order, -1 past the used ones; returns their number.  Unrolled over
// the camera bound so that the id loads (and what depends on them) issue
// together.
__device__ __forceinline__ int cand_points(const StereoBuffers &b, const StereoParams &p, int j,
                                           long f[kMaxCam]) {
  const int nc = b.n_cam;
  int pid[kMaxCam];
#pragma unroll
  for (int c = 0; c < kMaxCam; ++c)
    pid[c] = c < nc && p.active[c] ? b.cand_ids_s[(long)j * nc + c] : -1;
  int k = 0;
#pragma unroll
  for (int c = 0; c < kMaxCam; ++c) f[c] = -1;
#pragma unroll
  for (int c = 0; c < kMaxCam; ++c) {
    if (pid[c] >= 0) {
#pragma unroll
      for (int u = 0; u <= c; ++u)
        if (u == k) f[u] = (long)c * b.cap + pid[c];
      ++k;
    }
  }
  return k;
}

// Point lists by a two-level grouping instead of one counting atomic per
// entry: the entries (flat point f, candidate << 3 | slot) go into buckets
// of 2^low consecutive flat points - each workgroup histograms its own
// contiguous run of candidates in LDS, a scan over (bucket, workgroup) gives
// every workgroup its slot ranges, no global atomics - then one workgroup
// per bucket counting-sorts the bucket's entries in LDS, which also gives
// every point's list start and length.  The order inside a list is
// arbitrary: the ranks and the greedy do not depend on it.
struct GrpShape {
  int low, nb, g1;   // bucket bits, buckets, counting workgroups
};

// pass 1: per (bucket, workgroup) counts of the workgroup's run
__global__ void __launch_bounds__(256) k_grp_count(StereoBuffers bb, StereoParams p, GrpShape g) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  __shared__ int h[kGrpMaxBuckets];
  for (int i = threadIdx.x; i < g.nb; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const long per = ((long)M + g.g1 - 1) / g.g1;
  const int lo = (int)min((long)M, blockIdx.x * per), hi = (int)min((long)M, lo + per);
  for (int j = lo + threadIdx.x; j < hi; j += blockDim.x) {
    long f[kMaxCam];
    cand_points(b, p, j, f);
#pragma unroll
    for (int k = 0; k < kMaxCam; ++k)
      if (f[k] >= 0) atomicAdd(&h[f[k] >> g.low], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < g.nb; i += blockDim.x) b.grp_cnt[(long)i * g.g1 + blockIdx.x] = h[i];
}

// pass 1b: the same runs, each entry to its slot (staging in pt_sorted /
// pt_occ, free until the priority ranks)
__global__ void __launch_bounds__(256) k_grp_place(StereoBuffers bb, StereoParams p, GrpShape g) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  __shared__ int cur[kGrpMaxBuckets];
  for (int i = threadIdx.x; i < g.nb; i += blockDim.x)
    cur[i] = b.grp_off[(long)i * g.g1 + blockIdx.x];
  __syncthreads();
  const long per = ((long)M + g.g1 - 1) / g.g1;
  const int lo = (int)min((long)M, blockIdx.x * per), hi = (int)min((long)M, lo + per);
  for (int j = lo + threadIdx.x; j < hi; j += blockDim.x) {
    long f[kMaxCam];
    cand_points(b, p, j, f);
#pragma unroll
    for (int k = 0; k < kMaxCam; ++k) {
      if (f[k] < 0) continue;
      const int q = atomicAdd(&cur[f[k] >> g.low], 1);
      b.pt_sorted[q] = (int32_t)f[k];
      b.pt_occ[q] = (j << 3) | k;
    }
  }
}

// pass 2: one workgroup per bucket
__global__ void __launch_bounds__(256) k_grp_sort(StereoBuffers bb, GrpShape g) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int x = blockIdx.x;
  const long nflat = (long)b.n_cam * b.cap;
  const int s0 = b.grp_off[(long)x * g.g1];
  const int s1 = x + 1 < g.nb ? b.grp_off[(long)(x + 1) * g.g1] : *b.n_entries;
  const int P = 1 << g.low;
  __shared__ int cnt[1 << kGrpMaxLow], cur[1 << kGrpMaxLow];
  __shared__ int wsum[256 / kWave];
  for (int i = threadIdx.x; i < P; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  for (int q = s0 + threadIdx.x; q < s1; q += blockDim.x)
    atomicAdd(&cnt[b.pt_sorted[q] & (P - 1)], 1);
  __syncthreads();
  // exclusive scan of cnt: each thread a contiguous run, then the runs' sums
  const int per = P / (int)blockDim.x > 0 ? P / (int)blockDim.x : 1;
  const int i0 = threadIdx.x * per;
  int run = 0;
  for (int i = i0; i < i0 + per && i < P; ++i) run += cnt[i];
  int incl = run;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int o = __shfl_up(incl, off, kWave);
    if (lane_id() >= off) incl += o;
  }
  if (lane_id() == kWave - 1) wsum[threadIdx.x / kWave] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < (int)threadIdx.x / kWave; ++w) wbase += wsum[w];
  int acc = s0 + wbase + incl - run;
  for (int i = i0; i < i0 + per && i < P; ++i) {
    const long f = (long)x * P + i;
    cur[i] = acc;
    if (f < nflat) {
      b.pt_start[f] = acc;
      b.pt_count[f] = cnt[i];
      b.head[f] = acc;
    }
    acc += cnt[i];
  }
  __syncthreads();
  for (int q = s0 + threadIdx.x; q < s1; q += blockDim.x) {
    const int f = b.pt_sorted[q];
    const int32_t v = b.pt_occ[q];
    const int d = atomicAdd(&cur[f & (P - 1)], 1);
    b.pt_list[d] = v;
    b.pt_of[d] = f;
    b.e_ent[d] = b.cand_e_s[ent_cand(v)];
  }
}

// each list entry's percentile score, in list order (after k_pct_final)
__global__ void k_pct_ent(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int ne = *b.n_entries;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ne; q += gridDim.x * blockDim.x)
    b.p_ent[q] = b.pct[ent_cand(b.pt_list[q])];
}

__global__ void k_zero_entries(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int n = *b.n_entries;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    b.pt_occ[q] = 0;
    b.st_ent[q] = 0;
  }
}

// The rank kernels below take the list entries in tiles of kRankTile
// consecutive entries per workgroup.  A tile's entries belong to a run of
// whole point lists [A, B) (lists are contiguous and in order); when the run
// fits in kRankLds entries its keys are staged in LDS once and every entry
// counts against its list there - lanes of one list read the same LDS words
// (broadcast) - instead of re-reading the list's keys from the caches per
// entry.  A run that does not fit (a tile inside a very long list) is read
// from global memory as before.  Counts do not depend on where the keys are
// read from.
constexpr int kRankTile = 512;
// k_prio_rank stages three arrays of this many entries (20 KB): eight
// workgroups fit a CU's LDS, its thread limit (2048 entries took 40 KB and
// halved the resident workgroups: -26% time at 1024)
constexpr int kRankLds = 1024;

// the run of lists [A, B) holding entries [a, a1) (a < a1)
__device__ __forceinline__ void tile_lists(const StereoBuffers &b, int a, int a1, int &A,
                                           int &B) {
  const int f0 = b.pt_of[a], f1 = b.pt_of[a1 - 1];
  A = b.pt_start[f0];
  B = b.pt_start[f1] + b.pt_count[f1];
}

// Percentile of each candidate on each used 2D point (StereoMatch.cpp:
// 727-800): the point's candidates sorted by e_check (eq_eps ties by
// index), eq_eps blocks take their average rank, pv = rank / (n - 1).
// One thread per list entry counts the entries below it (r) and its eq_eps
// partners (t).  Where eq_eps is an equivalence on the frame's values (the
// condition k_hz_* verify, flags[2] otherwise) the blocks are its classes:
// the entry's block holds it and its t partners and starts after the r
// entries of lower blocks, so its average rank is (r + (r + t)) / 2.
__global__ void __launch_bounds__(256) k_pct_rank(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int ne = *b.n_entries;
  __shared__ double se[kRankLds];
  for (int a = blockIdx.x * kRankTile; a < ne; a += gridDim.x * kRankTile) {
    const int a1 = min(a + kRankTile, ne);
    int A, B;
    tile_lists(b, a, a1, A, B);
    const bool lds = B - A <= kRankLds;
    if (lds)
      for (int i = A + threadIdx.x; i < B; i += blockDim.x) se[i - A] = b.e_ent[i];
    __syncthreads();
    for (int q = a + threadIdx.x; q < a1; q += blockDim.x) {
      const int f = b.pt_of[q];
      const int s0 = b.pt_start[f], n = b.pt_count[f];
      const double ej = b.e_ent[q];
      int r = 0, nt = 0;
      // the point's entries are e_ent[s0, s0 + n); a candidate holds at
      // most one entry per point, so the entry itself is skipped by position
      if (lds) {
        const double *le = se + (s0 - A);
        for (int t = 0; t < n; ++t) {
          if (s0 + t == q) continue;
          const double em = le[t];
          if (eq_eps(em, ej))
            ++nt;
          else
            r += em < ej ? 1 : 0;
        }
      } else {
        for (int t0 = 0; t0 < n; t0 += kRankBatch) {
          double em[kRankBatch];
#pragma unroll
          for (int u = 0; u < kRankBatch; ++u) em[u] = b.e_ent[t0 + u < n ? s0 + t0 + u : q];
#pragma unroll
          for (int u = 0; u < kRankBatch; ++u) {
            if (t0 + u >= n || s0 + t0 + u == q) continue;
            if (eq_eps(em[u], ej))
              ++nt;
            else
              r += em[u] < ej ? 1 : 0;
          }
        }
      }
      const int32_t v = b.pt_list[q];
      const double denom = (double)(n - 1);
      const double avg_rank = 0.5 * (double)(r + (r + nt));
      b.pv[(long)ent_cand(v) * b.n_cam + ent_slot(v)] = (denom > 0.0) ? (avg_rank / denom) : 0.0;
    }
    __syncthreads();
  }
}

__global__ void k_pct_final(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
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
    const double pj = nf > 0 ? s / (double)nf : INFINITY;
    b.pct[j] = pj;
    b.state[j] = 0;
  }
}

// Greedy state of a candidate that k_prio_rank found below the first entry
// of one of its lists; k_mis_top turns it back into undecided (0).
constexpr int8_t kNotFirst = 3;

// Each point's candidates in the greedy's priority order (the global sort
// of StereoMatch.cpp:820-851 restricted to the point): position = number of
// entries that precede the entry.  A point whose positions collide (the
// pairwise order is not transitive there) is flagged for k_prio_fallback.
__global__ void __launch_bounds__(256) k_prio_rank(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int ne = *b.n_entries;
  __shared__ double sp[kRankLds], se[kRankLds];
  __shared__ int32_t sm[kRankLds];
  for (int a = blockIdx.x * kRankTile; a < ne; a += gridDim.x * kRankTile) {
    const int a1 = min(a + kRankTile, ne);
    int A, B;
    tile_lists(b, a, a1, A, B);
    const bool lds = B - A <= kRankLds;
    if (lds)
      for (int i = A + threadIdx.x; i < B; i += blockDim.x) {
        sp[i - A] = b.p_ent[i];
        se[i - A] = b.e_ent[i];
        sm[i - A] = ent_cand(b.pt_list[i]);
      }
    __syncthreads();
    for (int q = a + threadIdx.x; q < a1; q += blockDim.x) {
      const int f = b.pt_of[q];
      const int s0 = b.pt_start[f], n = b.pt_count[f];
      const int32_t v = b.pt_list[q];
      const int j = ent_cand(v);
      const double pj = b.p_ent[q], ej = b.e_ent[q];
      int r = 0;
      // the point's keys are the contiguous ranges [s0, s0 + n) of pt_list,
      // p_ent and e_ent
      if (lds) {
        // percentiles lie in [0, 1], where eq_eps's tolerance is at most
        // 1e-9 + 1e-12: a larger difference decides the order by itself
        const int o = s0 - A;
        for (int t = 0; t < n; ++t) {
          const int m = sm[o + t];
          const double pm = sp[o + t];
          const double d = pm - pj;
          bool lt;
          if (fabs(d) > 1.001e-9)
            lt = d < 0.0;
          else
            lt = m != j && less_order_v(pm, se[o + t], m, pj, ej, j);
          r += lt ? 1 : 0;
        }
      } else {
        for (int t0 = 0; t0 < n; t0 += kRankBatch) {
          int m[kRankBatch];
          double pm[kRankBatch], em[kRankBatch];
#pragma unroll
          for (int u = 0; u < kRankBatch; ++u) {
            const int t = t0 + u < n ? s0 + t0 + u : q;
            m[u] = ent_cand(b.pt_list[t]);
            pm[u] = b.p_ent[t];
            em[u] = b.e_ent[t];
          }
#pragma unroll
          for (int u = 0; u < kRankBatch; ++u)
            r += m[u] != j && less_order_v(pm[u], em[u], m[u], pj, ej, j) ? 1 : 0;
        }
      }
      b.ppos[(long)j * b.n_cam + ent_slot(v)] = s0 + r;
      b.pt_sorted[s0 + r] = v;
      if (atomicAdd(&b.pt_occ[s0 + r], 1) != 0) atomicOr(&b.pt_flag[f], 2);
      // not first on this point: the greedy's first round can skip it
      if (r != 0) b.state[j] = kNotFirst;
    }
    __syncthreads();
  }
}

__global__ void k_prio_fallback(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const long total = (long)b.n_cam * b.cap;
  for (long f = blockIdx.x * (long)blockDim.x + threadIdx.x; f < total;
       f += (long)gridDim.x * blockDim.x) {
    if (!(b.pt_flag[f] & 2)) continue;
    const int n = b.pt_count[f];
    const int s0 = b.pt_start[f];
    int32_t *lst = b.pt_sorted + s0;
    for (int i = 0; i < n; ++i) lst[i] = b.pt_list[s0 + i];
    for (int i = 1; i < n; ++i) {
      const int32_t cur = lst[i];
      int k = i - 1;
      while (k >= 0 && less_order(b.pct, b.cand_e_s, ent_cand(cur), ent_cand(lst[k]))) {
        lst[k + 1] = lst[k];
        --k;
      }
      lst[k + 1] = cur;
    }
    for (int i = 0; i < n; ++i)
      b.ppos[(long)ent_cand(lst[i]) * b.n_cam + ent_slot(lst[i])] = s0 + i;
  }
}

// One round of the priority greedy over the priority-sorted point lists.
// head[f] only moves past rejected entries, so every entry before it is
// rejected: a candidate is selected when on each of its points it is the
// head, and rejected when a selected entry precedes it on any point.
// In-place updates are safe: decided states never change, and an
// undecided entry ahead only delays a decision.
// One candidate's turn: selected (1) when every entry ahead of it on each of
// its points is rejected, rejected (2) when one is selected, else still
// undecided (0).  A decision is written to state[j] and to st_ent at the
// candidate's position in each of its points' sorted lists, where the scans
// of later candidates read it; states and heads are read with
// device-coherent loads.
__device__ __forceinline__ int mis_decide(const StereoBuffers &b, const StereoParams &p, int j) {
  const int nc = b.n_cam;
  long f[kMaxCam];
  cand_points(b, p, j, f);
  // every point's position and head (both indices into the sorted lists)
  // loaded together
  int pos[kMaxCam], h0[kMaxCam];
#pragma unroll
  for (int k = 0; k < kMaxCam; ++k) {
    pos[k] = 0;
    h0[k] = 0;
    if (f[k] >= 0) {
      pos[k] = b.ppos[(long)j * nc + k];
      h0[k] = __atomic_load_n(&b.head[f[k]], __ATOMIC_RELAXED);
    }
  }
  bool out = false, blocked = false;
#pragma unroll
  for (int k = 0; k < kMaxCam; ++k) {
    if (f[k] < 0 || out) continue;
    // first entry ahead of pos that is not rejected, kRankBatch entries per
    // step (their state loads in flight together).  States only leave 0 for
    // a final 1 or 2, so a batch's snapshot at worst defers a decision to a
    // later round.
    int h = h0[k], st = 2;
    while (h < pos[k]) {
      int sv[kRankBatch];
#pragma unroll
      for (int u = 0; u < kRankBatch; ++u)
        sv[u] = h + u < pos[k] ? __atomic_load_n(&b.st_ent[h + u], __ATOMIC_RELAXED) : 2;
      int adv = kRankBatch;
#pragma unroll
      for (int u = kRankBatch - 1; u >= 0; --u)
        if (sv[u] != 2) adv = u;
      h = min(h + adv, pos[k]);
      if (adv < kRankBatch) {
        st = sv[adv];
        break;
      }
    }
    if (h > h0[k]) atomicMax(&b.head[f[k]], h);
    if (h < pos[k]) {
      if (st == 1)
        out = true;
      else
        blocked = true;
    }
  }
  const int d = out ? 2 : (blocked ? 0 : 1);
  if (d != 0) {
    b.state[j] = (int8_t)d;
#pragma unroll
    for (int k = 0; k < kMaxCam; ++k)
      if (f[k] >= 0) {
        __atomic_store_n(&b.st_ent[pos[k]], d, __ATOMIC_RELAXED);
        // a point has at most one selected candidate
        if (d == 1) b.own[f[k]] = ((unsigned long long)(pos[k] + 1) << 32) | (unsigned)j;
      }
  }
  return d;
}

// After the first round most candidates are decided: a wave sweeps 64
// states at a time, queues the undecided in LDS and takes them 64 at a time,
// every lane on a candidate, instead of stepping over decided ones.
__global__ void __launch_bounds__(256) k_mis_round(StereoBuffers bb, StereoParams p, int round) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  if (round > 0 && b.undecided[round - 1] == 0) return;
  const int M = min(*b.n_cand, b.cap_cand);
  __shared__ int queue[256 / kWave][2 * kWave];
  int *qw = queue[threadIdx.x / kWave];
  const int lane = lane_id();
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const long nwave = (long)gridDim.x * blockDim.x / kWave;
  int cnt = 0, und = 0;
  for (long c0 = wave * kWave; c0 < M; c0 += nwave * kWave) {
    const int j = (int)c0 + lane;
    const bool pend = j < M && b.state[j] == 0;
    const uint64_t m = ballot(pend);
    if (pend) qw[cnt + mbcnt(m)] = j;
    cnt += __popcll(m);
    if (cnt >= kWave) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (mis_decide(b, p, qw[lane]) == 0) ++und;
      const int rest = cnt - kWave;
      const int v = lane < rest ? qw[kWave + lane] : 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane < rest) qw[lane] = v;
      cnt = rest;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < cnt && mis_decide(b, p, qw[lane]) == 0) ++und;
  und = wave_sum(und);
  if (lane == 0 && und) atomicAdd(&b.undecided[round], und);
}

// The greedy's first round: only a candidate that is first on every one of
// its lists can be selected before anything else is decided.  Those that
// k_prio_rank marked kNotFirst become undecided without a look at their
// lists; the others take their turn (mis_decide, which also covers points
// whose order k_prio_fallback rewrote: it reads the final positions).
__global__ void k_mis_top(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    const int8_t s = b.state[j];
    if (s == kNotFirst)
      b.state[j] = 0;
    else if (s == 0)
      (void)mis_decide(b, p, j);
  }
}

// After kMisFull rounds over all candidates the few still undecided are
// listed (in keep_off, free until the final compaction; count in
// undecided[kMisFull]) and one workgroup runs the remaining rounds over the
// list, a barrier between rounds.  The decisions are the same fixed point
// (each depends only on entries ahead of it), so the result is unchanged.
constexpr int kMisFull = 10;
constexpr int kMisFinishThreads = 1024;
constexpr int kMisFinishRounds = 1 << 14;

__global__ void k_mis_compact(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  if (b.undecided[kMisFull - 1] == 0) return;
  const int M = min(*b.n_cand, b.cap_cand);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x)
    if (b.state[j] == 0) b.keep_off[atomicAdd(&b.undecided[kMisFull], 1)] = j;
}

__global__ void __launch_bounds__(kMisFinishThreads)
k_mis_finish(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  __shared__ int s_und[kMisFinishThreads / 64];
  const int n = b.undecided[kMisFull];
  if (n == 0) return;
  for (int r = 0; r < kMisFinishRounds; ++r) {
    int und = 0;
    for (int i = threadIdx.x; i < n; i += kMisFinishThreads) {
      const int j = b.keep_off[i];
      if (__atomic_load_n(&b.state[j], __ATOMIC_RELAXED) != 0) continue;
      if (mis_decide(b, p, j) == 0) ++und;
    }
    und = wave_sum(und);
    if (lane_id() == 0) s_und[threadIdx.x / 64] = und;
    __syncthreads();
    int tot = 0;
    for (int w = 0; w < kMisFinishThreads / 64; ++w) tot += s_und[w];
    __syncthreads();
    if (tot == 0) return;
  }
  if (threadIdx.x == 0) atomicOr(&b.flags[5], 1);  // rounds exhausted
}

// Replacement branch of the reference (StereoMatch.cpp:906-968), evaluated
// for each rejected candidate against the owners it met at its turn: the
// selected candidates ahead of it on its points (at most one per point, the
// point's own[] record).
__global__ void k_replace_check(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  const int nc = b.n_cam;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x) {
    if (b.state[j] != 2) continue;
    long f[kMaxCam];
    cand_points(b, p, j, f);
    // every point's position and owner record loaded together
    int pos[kMaxCam];
    unsigned long long o[kMaxCam];
#pragma unroll
    for (int k = 0; k < kMaxCam; ++k) {
      pos[k] = f[k] >= 0 ? b.ppos[(long)j * nc + k] : 0;
      o[k] = f[k] >= 0 ? b.own[f[k]] : 0ULL;
    }
    int S[kMaxCam] = {}, ns = 0;
#pragma unroll
    for (int k = 0; k < kMaxCam; ++k) {
      if (o[k] == 0ULL || (int)(o[k] >> 32) - 1 >= pos[k]) continue;
      const int m = (int)(o[k] & 0xffffffffULL);
      bool seen = false;
#pragma unroll
      for (int q = 0; q < k; ++q) seen = seen || (q < ns && S[q] == m);
      if (seen) continue;
#pragma unroll
      for (int q = 0; q <= k; ++q)
        if (q == ns) S[q] = m;
      ++ns;
    }
    double ap = 0.0, ae = 0.0;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < kMaxCam; ++q) {
      if (q >= ns) continue;
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

// --------------------------------------------------------------------------
// eq_eps class check.  pruneMatch orders candidates with a comparator that
// treats values within rtol 1e-9 / atol 1e-12 as equal (StereoMatch.cpp:
// 820-851, the percentile blocks of :742-790 likewise).  That relation is an
// equivalence on a set of non-negative values exactly when, in sorted order,
// every run of eq_eps-adjacent values has eq_eps end points (a run whose
// ends are within the tolerance is a clique, and values of different runs
// are never within it).  Then the comparator is the strict total order
// (pct class, e class, index), every sort yields it, and the per-point
// ranks above are its restriction.  Otherwise std::stable_sort's result
// depends on the comparisons it happens to make, which this path does not
// replay: flags[2] is set and the caller reports an error.
//
// Without a global sort: g(v) = log(v + 1e-3) grows by at most ~1e-9
// between eq_eps values (v - u <= 1e-9 (v + 1e-3)), so buckets of width 2e-9
// in g put every eq_eps pair in the same or adjacent buckets and never
// relate values two buckets apart.  Each distinct value is stored once in a
// linear-probing table (k_hz_insert) whose home slot for bucket k is k mod
// the table size, unhashed, so that buckets k, k + 1, k + 2 are neighbouring
// slots and every value of them lies between home(k) and the first empty
// slot at or after home(k + 2) (a value sits after its home with no empty
// slot between).  The thread that stored a value scans that range, sorts the
// values of the three buckets in registers and checks every run in that
// window (a run reaching beyond it spans more than two buckets, so its
// visible part already fails).  More than kHzWindow values in a window, or a
// scan or probe longer than its bound, set the flag without a claim either
// way.  The values checked are the finite percentile scores (type 0) and
// their candidates' e_check (type 1): the candidates the greedy considers;
// the others sort after them whatever their order.
// --------------------------------------------------------------------------
constexpr double kHzWidth = 2e-9;
constexpr unsigned long long kHzEmpty = ~0ULL;   // no tagged finite value has these bits
constexpr unsigned long long kHzType = 1ULL << 63;
constexpr int kHzMaxProbe = 256;   // slots an insert may probe
constexpr int kHzMaxScan = 64;     // slots a window scan may visit

__device__ __forceinline__ long long hz_key(int type, double v) {
  const double g = log(v + 1e-3);
  const long long k = (long long)floor(g / kHzWidth);
  return ((long long)type << 42) + k + (1LL << 41);
}

// a value >= +0 with its type in the (free) sign bit
__device__ __forceinline__ unsigned long long hz_tag(int type, double v) {
  return (unsigned long long)__double_as_longlong(v) | (type ? kHzType : 0ULL);
}
__device__ __forceinline__ int hz_tag_type(unsigned long long w) { return (int)(w >> 63); }
__device__ __forceinline__ double hz_tag_val(unsigned long long w) {
  return __longlong_as_double((long long)(w & ~kHzType));
}

// slot mask of this call's part of the table (its candidate count's size)
__device__ __forceinline__ long hz_mask(const StereoBuffers &b, int M) {
  const long n = stereo_hz_size(M);
  return (n < b.hz_size ? n : b.hz_size) - 1;
}

// One thread per value (i = 2 j + type): the two values of a candidate
// insert independently.  The thread whose store makes a value present owns
// it: it alone checks and clears it (repeated values - most percentile
// scores - find their bits already there).
__global__ void k_hz_insert(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  const long mask = hz_mask(b, M);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * M; i += gridDim.x * blockDim.x) {
    const int j = i >> 1, t = i & 1;
    b.hz_slot[i] = -1;
    const double pj = b.pct[j];
    const double pv = t == 0 ? pj : b.cand_e_s[j];
    if (!isfinite(pj) || !isfinite(pv)) continue;
    const double v = pv + 0.0;  // -0 -> +0
    if (v < 0.0) {
      atomicOr(&b.flags[2], 1);
      continue;
    }
    const unsigned long long tag = hz_tag(t, v);
    long s = (long)((unsigned long long)hz_key(t, v) & (unsigned long long)mask);
    bool placed = false, owner = false;
    for (int n = 0; n < kHzMaxProbe && !placed; ++n) {
      // a plain read first: repeated values meet their bits without an atomic
      unsigned long long o = __atomic_load_n(b.hz_tab + s, __ATOMIC_RELAXED);
      if (o == kHzEmpty) o = atomicCAS(b.hz_tab + s, kHzEmpty, tag);
      owner = o == kHzEmpty;
      placed = owner || o == tag;
      if (!placed) s = (s + 1) & mask;
    }
    if (!placed) atomicOr(&b.flags[2], 1);   // probe bound
    if (owner) b.hz_slot[i] = (int32_t)s;
  }
}

constexpr int kHzChunk = 4;   // slots a window scan loads per step

__global__ void k_hz_check(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  const long mask = hz_mask(b, M);
  constexpr int NW = kHzWindow;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * M; i += gridDim.x * blockDim.x) {
    const long s = b.hz_slot[i];
    if (s < 0) continue;
    const unsigned long long tag = b.hz_tab[s];
    const int type = hz_tag_type(tag);
    const long long key = hz_key(type, hz_tag_val(tag));
    const long home = (long)((unsigned long long)key & (unsigned long long)mask);
    // the distinct values of buckets key, key + 1, key + 2 (absent ones as
    // +inf), gathered from home(key) to the first empty slot at or after
    // home(key + 2)
    double w[NW];
#pragma unroll
    for (int u = 0; u < NW; ++u) w[u] = INFINITY;
    int n = 0;
    bool bad = false, done = false;
    for (int d0 = 0; d0 < kHzMaxScan && !done; d0 += kHzChunk) {
      unsigned long long o[kHzChunk];
#pragma unroll
      for (int c = 0; c < kHzChunk; ++c) o[c] = b.hz_tab[(home + d0 + c) & mask];
#pragma unroll
      for (int c = 0; c < kHzChunk; ++c) {
        const int d = d0 + c;
        if (done) continue;
        if (o[c] == kHzEmpty) {
          done = d >= 2;
          continue;
        }
        if (hz_tag_type(o[c]) != type) continue;
        const double u = hz_tag_val(o[c]);
        const long long ku = hz_key(type, u);
        if (ku < key || ku > key + 2) continue;
        bad = bad || n == NW;     // window full
#pragma unroll
        for (int q = 0; q < NW; ++q)
          if (q == n) w[q] = u;
        n += n < NW;
      }
    }
    bad = bad || !done;           // scan bound
    // sorted by a transposition network in registers
#pragma unroll
    for (int r = 0; r < NW; ++r)
#pragma unroll
      for (int k = r & 1; k + 1 < NW; k += 2) {
        const double lo = fmin(w[k], w[k + 1]), hi = fmax(w[k], w[k + 1]);
        w[k] = lo;
        w[k + 1] = hi;
      }
    // every run of eq_eps-adjacent values must have eq_eps end points
    double r0 = w[0];
#pragma unroll
    for (int k = 1; k <= NW; ++k) {
      const bool end = k == NW || !isfinite(w[k]) || !eq_eps(w[k - 1], w[k]);
      if (isfinite(w[k - 1]) && end) {
        bad = bad || !eq_eps(r0, w[k - 1]);
        if (k < NW) r0 = w[k];
      }
    }
    if (bad) atomicOr(&b.flags[2], 1);
  }
}

// empty the slots used by this call (every used slot holds a value whose
// owner recorded it)
__global__ void k_hz_clear(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * M; i += gridDim.x * blockDim.x) {
    const long s = b.hz_slot[i];
    if (s >= 0) b.hz_tab[s] = kHzEmpty;
  }
}

__global__ void k_hz_init(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < b.hz_size;
       s += (long)gridDim.x * blockDim.x)
    b.hz_tab[s] = kHzEmpty;
}

void stereo_init_tables(const StereoBuffers &b, hipStream_t st) {
  k_hz_init<<<dim3(1024, 1, b.nf), 256, 0, st>>>(b);
}

// triangulateMatch: camera-order LOS, FP64, gate on tol_3d.  A candidate's
// cameras are gathered into slots 0..n-1 in camera order by selects over
// compile-time slot and camera indices (MC >= n_cam), so the LOS arrays stay
// in registers.
// One selected candidate: its LOS in camera order, triangulated; keep[j]
// and tri_xyz[j] written.
template <int MC>
__device__ __forceinline__ void tri_one(const StereoBuffers &b, const StereoParams &p, int j) {
  const int nc = b.n_cam;
  int keep = 0;
  int pidc[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) pidc[c] = c < nc ? b.cand_ids_s[(long)j * nc + c] : -1;
  double pt[MC][3], u[MC][3];
#pragma unroll
  for (int s = 0; s < MC; ++s)
#pragma unroll
    for (int d = 0; d < 3; ++d) pt[s][d] = u[s][d] = 0.0;
  int n = 0;
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    if (pidc[c] < 0) continue;
    const double *uu = b.los + ((long)c * b.cap + point_pos(b, c, pidc[c])) * 4;
    double tc[3], uc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      tc[d] = b.cams[c].Tinv[d];
      uc[d] = uu[d];
    }
#pragma unroll
    for (int s = 0; s <= c; ++s)
      if (s == n) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          pt[s][d] = tc[d];
          u[s][d] = uc[d];
        }
      }
    ++n;
  }
  double X[3], e;
  if (n >= 2 && triangulate_upto<MC>(pt, u, n, X, e) && !(e > p.tol3d)) {
    keep = 1;
    b.tri_xyz[(long)j * 3 + 0] = X[0];
    b.tri_xyz[(long)j * 3 + 1] = X[1];
    b.tri_xyz[(long)j * 3 + 2] = X[2];
  }
  b.keep[j] = keep;
}

// A wave sweeps 64 candidates at a time: the unselected get keep = 0 (the
// keep flags are scanned over the first n_cand entries) and the selected
// (about one in twelve) join the wave's LDS queue, which is triangulated 64
// at a time with every lane busy instead of a few lanes of each sweep step.
template <int MC>
__global__ void __launch_bounds__(256) k_tri(StereoBuffers bb, StereoParams p) {
  const StereoBuffers b = bb.frame(blockIdx.z);
  const int M = min(*b.n_cand, b.cap_cand);
  __shared__ int queue[256 / kWave][2 * kWave];
  int *qw = queue[threadIdx.x / kWave];
  const int lane = lane_id();
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const long nwave = (long)gridDim.x * blockDim.x / kWave;
  int cnt = 0, nsel = 0;
  for (long c0 = wave * kWave; c0 < M; c0 += nwave * kWave) {
    const int j = (int)c0 + lane;
    const bool sel = j < M && b.state[j] == 1;
    if (j < M && !sel) b.keep[j] = 0;
    const uint64_t m = ballot(sel);
    if (sel) qw[cnt + mbcnt(m)] = j;
    cnt += __popcll(m);
    nsel += sel ? 1 : 0;
    if (cnt >= kWave) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      tri_one<MC>(b, p, qw[lane]);
      const int rest = cnt - kWave;
      const int v = lane < rest ? qw[kWave + lane] : 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane < rest) qw[lane] = v;
      cnt = rest;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < cnt) tri_one<MC>(b, p, qw[lane]);
  nsel = wave_sum(nsel);
  if (lane == 0 && nsel) atomicAdd(b.n_sel, nsel);
}

__global__ void k_obj_write(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
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

__global__ void k_obj_count(StereoBuffers bb) {
  const StereoBuffers b = bb.frame(blockIdx.z);
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

template <int K>
static void build_level(const StereoBuffers &b, const StereoParams &p, hipStream_t st) {
  if (K == 2) stage_begin(ST_ENUM2, st);
  k_b_enum<K><<<dim3(launch_blocks(b.grid_build, b.nf), 1, b.nf), 256, 0, st>>>(b, p);
  if (K == 2) stage_end(ST_ENUM2, st);
  if (K + 1 == p.build_count)
    k_b_extend<K, true><<<dim3(launch_blocks(b.grid_build, b.nf), 1, b.nf), 256, 0, st>>>(b, p);
  else
    k_b_extend<K, false><<<dim3(launch_blocks(b.grid_build, b.nf), 1, b.nf), 256, 0, st>>>(b, p);
}

GrpShape grp_shape(const StereoBuffers &b) {
  const long nflat = (long)b.n_cam * b.cap;
  GrpShape g;
  g.low = grp_low(nflat);
  g.nb = (int)grp_buckets(nflat);
  g.g1 = std::max(kGrpMinBlocks, kGrpBlocks / std::max(1, b.nf));
  return g;
}

// pruneMatch on the sorted candidates (cand_ids_s, cand_e_s, n_cand):
// per-point percentile and priority orders, the eq_eps check, the greedy
// rounds and the replacement check; state[j] = 1 marks the selection.
static void prune_section(const StereoBuffers &b, const StereoParams &p, int gM,
                          hipStream_t st) {
  const int T = 256;
  const long nflat = (long)b.n_cam * b.cap;
  zero_list(b, ZeroList{{b.pt_flag, (int32_t *)b.own}, {nflat, 2 * nflat}, 2}, st);
  const GrpShape g = grp_shape(b);
  k_grp_count<<<dim3(g.g1, 1, b.nf), T, 0, st>>>(b, p, g);
  scan_exclusive(b.grp_cnt, b.grp_off, (long)g.nb * g.g1, 1, (long)g.nb * g.g1, b.n_entries,
                 b.scan_scratch, st, b.nf, b.fstride);
  k_grp_place<<<dim3(g.g1, 1, b.nf), T, 0, st>>>(b, p, g);
  k_grp_sort<<<dim3(g.nb, 1, b.nf), T, 0, st>>>(b, g);
  const int gE = grid_for((long)b.cap_cand * 4, T, 4096);
  k_zero_entries<<<dim3(gE, 1, b.nf), T, 0, st>>>(b);
  const int gR = grid_for((long)b.cap_cand * 4, kRankTile, 4096);
  k_pct_rank<<<dim3(gR, 1, b.nf), T, 0, st>>>(b);
  k_pct_final<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  k_pct_ent<<<dim3(gE, 1, b.nf), T, 0, st>>>(b);
  k_hz_insert<<<dim3(gM, 1, b.nf), T, 0, st>>>(b);
  k_hz_check<<<dim3(gM, 1, b.nf), T, 0, st>>>(b);
  k_hz_clear<<<dim3(gM, 1, b.nf), T, 0, st>>>(b);
  k_prio_rank<<<dim3(gR, 1, b.nf), T, 0, st>>>(b);
  k_prio_fallback<<<dim3(grid_for(nflat, T, 2048), 1, b.nf), T, 0, st>>>(b);
  // the first rounds decide almost every candidate; one workgroup finishes
  // the few left
  k_mis_top<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  for (int r = 0; r < kMisFull; ++r) k_mis_round<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p, r);
  k_mis_compact<<<dim3(gM, 1, b.nf), T, 0, st>>>(b);
  k_mis_finish<<<dim3(1, 1, b.nf), kMisFinishThreads, 0, st>>>(b, p);
  k_replace_check<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
}

void launch_stereo(const StereoBuffers &b, const StereoParams &p,
                   int max_rounds, hipStream_t st) {
  const long ncell = (long)b.rows_cell * b.cols_cell;
  const int T = 256;
  // the pass's counters and count arrays, all in one launch (nothing before
  // their first use reads them)
  zero_list(b,
            ZeroList{{b.n_cand, b.n_sel, b.work, b.undecided, b.cell_cnt, b.out_cnt, b.n_path,
                      b.n_surv, b.n_pairs, b.key_cnt},
                     {1, 1, 1, max_rounds, ncell * b.n_cam, b.n_cam, kMaxCam, kMaxCam, 1, b.cap},
                     10},
            st);
  // LOS + cell index
  dim3 gc(grid_for(b.cap, T, 1024), b.n_cam, b.nf);
  k_cell_count<<<gc, T, 0, st>>>(b, p);
  scan_exclusive(b.cell_cnt, b.cell_start, ncell, b.n_cam, ncell, nullptr,
                 b.scan_scratch, st, b.nf, b.fstride);
  k_cell_fill<<<gc, T, 0, st>>>(b, p);
  k_los<<<gc, T, 0, st>>>(b, p);
  // build: level 1 from every reference observation, then one enumerate +
  // extend pair of kernels per deeper build camera
  const int G = launch_blocks(b.grid_build, b.nf);
  stage_begin(ST_BUILD, st);
  k_b_second<<<dim3(grid_for(b.cap, T, 1024), 1, b.nf), T, 0, st>>>(b, p);
  k_b_band<<<dim3(G, 1, b.nf), T, 0, st>>>(b, p);
  if (p.build_count == 2)
    k_b_extend<1, true><<<dim3(G, 1, b.nf), T, 0, st>>>(b, p);
  else
    k_b_extend<1, false><<<dim3(G, 1, b.nf), T, 0, st>>>(b, p);
  if (p.build_count > 2) build_level<2>(b, p, st);
  if (p.build_count > 3) build_level<3>(b, p, st);
  if (p.build_count > 4) build_level<4>(b, p, st);
  if (p.build_count > 5) build_level<5>(b, p, st);
  if (p.build_count > 6) build_level<6>(b, p, st);
  if (p.build_count > 7) build_level<7>(b, p, st);
  stage_end(ST_BUILD, st);
  stage_begin(ST_PRUNE, st);
  // sort
  const int gM = grid_for(b.cap_cand, T, launch_blocks(b.grid_prune, b.nf));
  k_key_hist<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  scan_exclusive(b.key_cnt, b.key_start, b.cap, 1, b.cap, nullptr, b.scan_scratch, st, b.nf, b.fstride);
  k_key_scatter<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  k_bucket_rank<<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  prune_section(b, p, gM, st);
  // triangulate + compact in candidate order
  if (b.n_cam <= 4)
    k_tri<4><<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  else
    k_tri<kMaxCam><<<dim3(gM, 1, b.nf), T, 0, st>>>(b, p);
  scan_exclusive(b.keep, b.keep_off, b.cap_cand, 1, b.cap_cand, b.totals,
                 b.scan_scratch, st, b.nf, b.fstride, b.n_cand);
  k_obj_write<<<dim3(gM, 1, b.nf), T, 0, st>>>(b);
  k_obj_count<<<dim3(1, 1, b.nf), 64, 0, st>>>(b);
  stage_end(ST_PRUNE, st);
}

void launch_prune(const StereoBuffers &b, const StereoParams &p, int max_rounds,
                  hipStream_t st) {
  zero_i32(b, b.undecided, max_rounds, st);
  prune_section(b, p, grid_for(b.cap_cand, 256, launch_blocks(b.grid_prune, b.nf)), st);
}

} // namespace olpt
