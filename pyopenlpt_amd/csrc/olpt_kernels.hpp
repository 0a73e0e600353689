// olpt_kernels.hpp - host-side launchers for the IPR kernels.  All counts
// live in device memory; every launcher takes fixed grids sized from the
// context's capacities, so a whole frame is a fixed kernel sequence that a
// hipGraph can capture.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "olpt_c.h"
#include "olpt_device.hpp"

namespace olpt {

// Optional stage timing with HIP events on the launch stream (eager runs
// only).  Slots: see olpt_c.h (OLPT_STAGE_*).
constexpr int kStageEvents = 2048;  // timed occurrences per stage
struct StageTimer {
  bool on = false;
  hipEvent_t ev[2][8][kStageEvents];
  int n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  void begin(int slot, hipStream_t st) {
    if (on && n[slot] < kStageEvents) (void)hipEventRecord(ev[0][slot][n[slot]], st);
  }
  void end(int slot, hipStream_t st) {
    if (on && n[slot] < kStageEvents) {
      (void)hipEventRecord(ev[1][slot][n[slot]], st);
      ++n[slot];
    }
  }
};
extern thread_local StageTimer *g_timer;
inline void stage_begin(int slot, hipStream_t st) {
  if (g_timer) g_timer->begin(slot, st);
}
inline void stage_end(int slot, hipStream_t st) {
  if (g_timer) g_timer->end(slot, st);
}
enum { ST_PEAKS = 0, ST_BUILD = 1, ST_PRUNE = 2, ST_SHAKE = 3, ST_RESID = 4, ST_ENUM2 = 5 };

// Frames of one batch share a workspace layout: frame f's buffers lie f *
// fstride bytes after frame 0's (one arena per workspace), so a kernel
// launched over a batch takes its frame from blockIdx.z and shifts the
// workspace pointers it was given for frame 0.  Null stays null.
template <typename P> __host__ __device__ __forceinline__ P *fshift(P *p, long off) {
  return p ? (P *)((char *)p + off) : p;
}

// Which cameras a stage works on (host-side list, passed by value).
struct CamList {
  int32_t n;
  int32_t id[kMaxCam];
};

// ---------------- peaks (ObjectFinder2D::findTracer2D) --------------------
// img: [n_cam][H][W] (element stride cam_stride); output xy as (x,y) double
// pairs at xy[(cam*cap + i)*2]; npts[cam] = min(count, cap); overflow[0] is
// set to 1 if any camera had more than cap peaks.  Over nf frames every
// pointer (img included) is frame 0's and moves fstride bytes per frame.  Work buffers:
// mask/cnt/off hold (H-2)*ceil((W-2)/64) entries per listed camera.
template <typename TIn>
void launch_find_peaks(const TIn *img, long cam_stride, int H, int W,
                       double min_int, CamList cams, uint64_t *mask,
                       int32_t *cnt, int32_t *off, int32_t *scan_scratch,
                       int32_t *totals, double *xy, int cap, int32_t *npts,
                       int32_t *overflow, hipStream_t st, int nf = 1,
                       long fstride = 0);

// limit2DObjectNumber (IPR.cpp:19-39) on the peak lists of `cams` (xy and
// npts as launch_find_peaks writes them; a no-op unless 0 < cap < cap2d).
// Work buffers: perm [n_cam][cap2d] int32, lim [n_cam], tmp [n_cam][cap2d][2].
void launch_limit2d(double *xy, int32_t *npts, CamList cams, int cap, int cap2d,
                    int32_t *perm, int32_t *lim, double *tmp, hipStream_t st, int nf = 1,
                    long fstride = 0);

inline long peak_chunks(int H, int W) {
  return (long)(H > 2 ? H - 2 : 0) * ((W - 2 + 63) / 64);
}

// ---------------- stereo matching (StereoMatch) ---------------------------
// A partial match of the stereo build with the point it is being extended
// by: the points and cameras of path slots 0..K (slot 0 the reference
// observation, slot K the point under test).
// Workgroups per frame of a launch covering nf frames: the tuned count, but
// at least min_total in all (kMinLaunchBlocks 256-thread workgroups, four
// times as many one-wave ones), so that launches over few frames (one frame
// per call) still fill the chip.  A launch shape never changes what a
// kernel computes.
constexpr int kMinLaunchBlocks = 1024;
__host__ __device__ inline int launch_blocks(int tuned, int nf,
                                             int min_total = kMinLaunchBlocks) {
  const int need = (min_total + (nf > 0 ? nf : 1) - 1) / (nf > 0 ? nf : 1);
  const int g = tuned > need ? tuned : need;
  return g < 65536 ? g : 65536;
}

template <int K> struct PathRec {
  int32_t pid[K + 1];
  int8_t cam[K + 1];
};

// bytes of the largest path record a build with n_cam cameras lists
inline long path_rec_bytes(int n_cam) {
  const long k1 = n_cam > 1 ? n_cam : 2;   // K + 1 <= n_cam
  return (k1 * 5 + 3) / 4 * 4;
}

// point-list grouping (prune): buckets of 2^low consecutive flat points,
// at most kGrpMaxBuckets of them; kGrpBlocks workgroups in the first pass
// over a whole launch (kGrpBlocks / nf per frame, at least kGrpMinBlocks):
// the (bucket, workgroup) counts it scans grow with the workgroups, so 16
// frames at 64 each beat 1,024 each (-1.2% of the step)
constexpr int kGrpBlocks = 1024;
constexpr int kGrpMinBlocks = 64;
constexpr int kGrpMaxBuckets = 4096;
// max_2d <= 2^20 and at most 8 cameras give nflat <= 2^23, so low <= 12:
// k_grp_sort's LDS counters are sized for that bound
constexpr int kGrpMaxLow = 12;
inline int grp_low(long nflat) {
  int low = 9;
  while ((nflat >> low) >= kGrpMaxBuckets) ++low;
  return low;
}
inline long grp_buckets(long nflat) {
  const int low = grp_low(nflat);
  return (nflat + (1L << low) - 1) >> low;
}
constexpr int kHzWindow = 12;  // distinct values an eq_eps check window may hold

struct StereoBuffers {
  int n_cam;
  int cap;               // per-camera 2D capacity
  const CamT<double> *cams;
  const double *xy;      // [n_cam][cap][2]
  const int32_t *npts;   // [n_cam]
  // the build numbers each camera's points by their position in cell order
  // (cell_xy / cell_pid), the points outside the cell grid last; los and
  // lines are indexed by that position
  double *los;           // [n_cam][cap][4]  unit direction (w unused)
  double *lines;         // [n_cam][n_cam][cap][4] LOS (c, pos) on camera t
  // build
  int8_t *t1;            // [cap] second build camera per reference observation
  long cap_pairs;
  int2 *pairs;           // [cap_pairs] (partial match, point) work list
  int32_t *n_pairs;      // device
  // the build's work lists as whole paths: path[0] the (partial match,
  // point) items of level K >= 2 (PathRec<K>), path[1] the survivors of
  // level K (PathRec<K + 1>: K + 1 points and the next camera in slot K + 1)
  void *path[2];         // [cap_pairs] records each
  int32_t *n_path;       // [kMaxCam] device: items of each level
  int32_t *n_surv;       // [kMaxCam] device: survivors of each level
  int cell_px, rows_cell, cols_cell;
  int32_t *cell_cnt;     // [n_cam][ncell]
  int32_t *cell_start;   // [n_cam][ncell]
  int32_t *cell_pid;     // [n_cam][cap]
  double2 *cell_xy;      // [n_cam][cap] the points' coordinates in cell_pid order
  int32_t *pt_cell;      // [n_cam][cap]
  int32_t *pt_rank;      // [n_cam][cap]
  int32_t *out_cnt;      // [n_cam] points outside the cell grid
  // candidates
  int cap_cand;
  int32_t *cand_ids;     // [cap_cand][n_cam] raw (emission order)
  double *cand_e;
  int32_t *n_cand;       // device counter
  int32_t *cand_ids_s;   // sorted
  double *cand_e_s;
  int32_t *key_cnt;      // [cap] histogram on the primary camera
  int32_t *key_start;    // [cap]
  int32_t *cand_rank;    // [cap_cand]
  int32_t *bucket_idx;   // [cap_cand] raw candidate index, grouped by key
  int32_t *bucket_ids;   // [cap_cand][n_cam] id vectors in bucket_idx order
  // prune
  int32_t *pt_count;     // [n_cam*cap] candidates per flat point
  int32_t *pt_start;     // [n_cam*cap]
  int32_t *pt_list;      // [cap_cand*n_cam] (candidate << 3 | slot)
  int32_t *pt_of;        // [cap_cand*n_cam] flat point of each list entry
  int32_t *pt_sorted;    // [cap_cand*n_cam] lists in priority order
  int32_t *pt_occ;       // [cap_cand*n_cam] position occupancy (rank check)
  int32_t *ppos;         // [cap_cand][n_cam] index of the entry in pt_sorted (its point's
                         // priority-sorted run)
  int32_t *pt_flag;      // [n_cam*cap] bit1: priority order redone sequentially
  int32_t *head;         // [n_cam*cap] first entry not known to be rejected (pt_sorted index)
  // per flat point its selected candidate, if any: (pt_sorted index + 1) << 32 |
  // candidate, 0 = none (written by the greedy as it selects)
  unsigned long long *own; // [n_cam*cap]
  int32_t *n_entries;    // device: total list entries
  int32_t *grp_cnt;      // [grp buckets * kGrpBlocks] point-list grouping counts
  int32_t *grp_off;      // [grp buckets * kGrpBlocks] their exclusive scan
  // per list entry, in pt_list order: the entry's candidate's e_check and
  // percentile score, so that a point's rank loops read its own contiguous
  // range
  double *e_ent;         // [cap_cand*n_cam]
  double *p_ent;         // [cap_cand*n_cam]
  // state of the candidate at each pt_sorted position (the greedy's scans
  // read a point's run of states contiguously)
  int32_t *st_ent;       // [cap_cand*n_cam]
  double *pv;            // [cap_cand][n_cam] percentile per used point
  double *pct;           // [cap_cand]
  int8_t *state;         // [cap_cand] 0 undecided, 1 selected, 2 rejected
  int32_t *undecided;    // [max_rounds] per-round undecided counts
  int32_t *keep;         // [cap_cand] 1 if selected and triangulated
  int32_t *keep_off;     // [cap_cand]
  double *tri_xyz;       // [cap_cand][3]
  // objects
  int cap_obj;
  double *obj_xyz;       // [cap_obj][3]
  int32_t *obj_ids;      // [cap_obj][n_cam]
  int32_t *n_obj;        // device
  int32_t *n_sel;        // device
  // eq_eps class check of the prune keys (k_hz_*): a linear-probing table
  // of distinct values, one 64-bit word each (the value's bits with its
  // type in bit 63; all ones = empty), home slot = bucket key mod size
  unsigned long long *hz_tab; // [hz_size]
  int32_t *hz_slot;           // [2 * cap_cand] slot of each inserted value
  long hz_size;               // power of two
  int32_t *flags;        // [8]: 0 cand overflow, 1 list overflow, 2 eq_eps
                         //      ties not transitive, 3 replacement,
                         //      4 obj overflow, 5 rounds exhausted,
                         //      6 triangulation throw
  int32_t *work;         // dynamic work counter
  int32_t *scan_scratch;
  int32_t *totals;       // small scratch for scan totals
  int grid_build;        // workgroups of the build kernels (per frame)
  int grid_prune;        // cap on workgroups of the sort/prune kernels (per frame)
  int nf;                // frames of this batch (grid z of every launch)
  long fstride;          // bytes between consecutive frames' buffers
  // the buffers of frame f (cameras are shared)
  __host__ __device__ StereoBuffers frame(long f) const {
    StereoBuffers o = *this;
    const long d = f * fstride;
    o.xy = fshift(xy, d);
    o.npts = fshift(npts, d);
    o.los = fshift(los, d);
    o.lines = fshift(lines, d);
    o.t1 = fshift(t1, d);
    o.pairs = fshift(pairs, d);
    o.n_pairs = fshift(n_pairs, d);
    o.path[0] = fshift(path[0], d);
    o.path[1] = fshift(path[1], d);
    o.n_path = fshift(n_path, d);
    o.n_surv = fshift(n_surv, d);
    o.cell_cnt = fshift(cell_cnt, d);
    o.cell_start = fshift(cell_start, d);
    o.cell_pid = fshift(cell_pid, d);
    o.cell_xy = fshift(cell_xy, d);
    o.pt_cell = fshift(pt_cell, d);
    o.pt_rank = fshift(pt_rank, d);
    o.out_cnt = fshift(out_cnt, d);
    o.cand_ids = fshift(cand_ids, d);
    o.cand_e = fshift(cand_e, d);
    o.n_cand = fshift(n_cand, d);
    o.cand_ids_s = fshift(cand_ids_s, d);
    o.cand_e_s = fshift(cand_e_s, d);
    o.key_cnt = fshift(key_cnt, d);
    o.key_start = fshift(key_start, d);
    o.cand_rank = fshift(cand_rank, d);
    o.bucket_idx = fshift(bucket_idx, d);
    o.bucket_ids = fshift(bucket_ids, d);
    o.pt_count = fshift(pt_count, d);
    o.pt_start = fshift(pt_start, d);
    o.pt_list = fshift(pt_list, d);
    o.pt_of = fshift(pt_of, d);
    o.pt_sorted = fshift(pt_sorted, d);
    o.pt_occ = fshift(pt_occ, d);
    o.ppos = fshift(ppos, d);
    o.pt_flag = fshift(pt_flag, d);
    o.head = fshift(head, d);
    o.own = fshift(own, d);
    o.n_entries = fshift(n_entries, d);
    o.grp_cnt = fshift(grp_cnt, d);
    o.grp_off = fshift(grp_off, d);
    o.e_ent = fshift(e_ent, d);
    o.p_ent = fshift(p_ent, d);
    o.st_ent = fshift(st_ent, d);
    o.pv = fshift(pv, d);
    o.pct = fshift(pct, d);
    o.state = fshift(state, d);
    o.undecided = fshift(undecided, d);
    o.keep = fshift(keep, d);
    o.keep_off = fshift(keep_off, d);
    o.tri_xyz = fshift(tri_xyz, d);
    o.obj_xyz = fshift(obj_xyz, d);
    o.obj_ids = fshift(obj_ids, d);
    o.n_obj = fshift(n_obj, d);
    o.n_sel = fshift(n_sel, d);
    o.hz_tab = fshift(hz_tab, d);
    o.hz_slot = fshift(hz_slot, d);
    o.flags = fshift(flags, d);
    o.work = fshift(work, d);
    o.scan_scratch = fshift(scan_scratch, d);
    o.totals = fshift(totals, d);
    return o;
  }
};

struct StereoParams {
  int32_t active[kMaxCam];
  int n_active;
  int build_count;
  double tol2d, tol3d;
  double limit[6];
};

void launch_stereo(const StereoBuffers &b, const StereoParams &p,
                   int max_rounds, hipStream_t st);
// pruneMatch alone on candidates already in cand_ids_s / cand_e_s (count
// in n_cand): state[j] == 1 marks the selection, flags as launch_stereo.
void launch_prune(const StereoBuffers &b, const StereoParams &p, int max_rounds,
                  hipStream_t st);
// Empty the eq_eps check table (once, after allocation; every stereo call
// leaves it empty again).
void stereo_init_tables(const StereoBuffers &b, hipStream_t st);
// Number of slots of the eq_eps check table for n candidates: two values
// each, at most a quarter of the slots in use.  The table is allocated for
// the candidate capacity; each call uses the part its candidate count needs.
__host__ __device__ inline long stereo_hz_size(long n_cand) {
  long n = 1024;
  while (n < 8 * n_cand) n <<= 1;
  return n;
}

// ---------------- shake (Shake::runShake) ---------------------------------
template <typename T> struct ShakeBuffers {
  int n_cam, H, W;
  const CamT<T> *cams;
  OtfT<T> otf;
  T *img;             // [n_cam][H][W] loop input (orig for shake)
  T *res;             // [n_cam][H][W] residual (work)
  T *img_next;        // [n_cam][H][W] IPR-mode residual (next loop input)
  int cap_obj;
  double *xyz;        // [cap_obj][3]   in/out positions
  T *uv;              // [cap_obj][n_cam][2]
  double *score;      // [cap_obj]
  uint8_t *flag;      // [cap_obj]
  const int32_t *n_obj;
  // binning of 4x4 projection boxes (residual gather)
  int bin_px, bin_rows, bin_cols;
  int32_t *bin_cnt, *bin_start, *obj_bin, *obj_rank; // per cam
  void *bin_rec;      // [n_cam * cap_obj] the projection records in bin order
  // repeated-object grid
  int grid_n;         // cells per axis
  double grid_lo[3], grid_cell;
  int32_t *g_cnt, *g_start, *g_obj, *g_rank, *g_cellof;
  int32_t *has_nb;
  // order in which the per-object shake kernels visit the objects (null:
  // index order); frame 0's array, shifted like the others
  const int32_t *order;
  // per (object, camera) projection records, reduction partials
  void *rec;          // [cap_obj][n_cam] records (sizeof depends on T)
  double *partial;    // [2*256]
  int32_t *keep_off;  // [cap_obj]
  int32_t *n_keep;
  int32_t *scan_scratch;
  int32_t *totals;
  int32_t *work;
  // learned proposal (optional)
  void *patches;      // [cap_obj][512] bf16 network input
  float *prop;        // [cap_obj][4] network output
  int grid_shake;     // one-wave workgroups of the shake/proposal kernels (per frame)
  int nf;             // frames of this batch (grid z of every launch)
  long fstride;       // bytes between consecutive frames' workspace buffers
  long lstride;       // bytes between consecutive frames' patches
  long pstride;       // bytes between consecutive frames' network outputs
  // the buffers of frame f (cameras and OTF are shared)
  __host__ __device__ ShakeBuffers frame(long f) const {
    ShakeBuffers o = *this;
    const long d = f * fstride;
    o.img = fshift(img, d);
    o.res = fshift(res, d);
    o.img_next = fshift(img_next, d);
    o.xyz = fshift(xyz, d);
    o.uv = fshift(uv, d);
    o.score = fshift(score, d);
    o.flag = fshift(flag, d);
    o.n_obj = fshift(n_obj, d);
    o.bin_cnt = fshift(bin_cnt, d);
    o.bin_start = fshift(bin_start, d);
    o.bin_rec = fshift(bin_rec, d);
    o.obj_bin = fshift(obj_bin, d);
    o.obj_rank = fshift(obj_rank, d);
    o.g_cnt = fshift(g_cnt, d);
    o.g_start = fshift(g_start, d);
    o.g_obj = fshift(g_obj, d);
    o.g_rank = fshift(g_rank, d);
    o.g_cellof = fshift(g_cellof, d);
    o.has_nb = fshift(has_nb, d);
    o.order = fshift(order, d);
    o.rec = fshift(rec, d);
    o.partial = fshift(partial, d);
    o.keep_off = fshift(keep_off, d);
    o.n_keep = fshift(n_keep, d);
    o.scan_scratch = fshift(scan_scratch, d);
    o.totals = fshift(totals, d);
    o.work = fshift(work, d);
    o.patches = fshift(patches, f * lstride);
    o.prop = fshift(prop, f * pstride);
    return o;
  }
};

struct ShakeParams {
  int32_t active[kMaxCam];
  double shake_width, shakewidth_min, thred_ghost, tol3d, r_px;
  const void *mlp;            // learned proposal network (olpt_mlp), or null
  double voxel_mm, max_step_vox;
  double ghost_logit;         // learned proposal: ghost logit above this flags a ghost
  int n_shake_loop, ratio_augimg;
  // calROIBound box edges (box_width): the object box (r_px) and the shake
  // ROI (r_px * ratio_augimg), fixed on the host in double arithmetic
  int box_w, roi_w;
};

template <typename T>
void launch_shake(const ShakeBuffers<T> &b, const ShakeParams &p,
                  hipStream_t st);

// IPR-mode residual of the kept objects: img_next = max(0, img - max proj)
// over all cameras (Shake::calResidualImage, Shake.h:40-48).
template <typename T>
void launch_residual_ipr(const ShakeBuffers<T> &b, const int32_t *n_keep,
                         const double *xyz_keep, T *uv_keep, double r_px,
                         hipStream_t st);

// order-preserving compaction of the objects with flag == 0 into
// (xyz_keep, uv_keep); count to *n_keep
template <typename T>
void launch_filter(const ShakeBuffers<T> &b, double *xyz_keep, T *uv_keep,
                   int32_t *n_keep, hipStream_t st);

// size in bytes of one (object, camera) record
size_t shake_rec_bytes(bool f64);

// Learned shake proposal: network inputs of the objects (residual from their
// positions, then the 4 x 11 x 11 patches, bf16 [cap_obj][512]).
template <typename T>
void launch_patches(const ShakeBuffers<T> &b, const ShakeParams &p, void *X, hipStream_t st);
// Render objects (max-fusion of their projections; amp, if given, replaces
// the OTF amplitude per object) into out [n_cam][H][W].
template <typename T>
void launch_render(const ShakeBuffers<T> &b, const double *xyz, const double *amp,
                   const int32_t *n_ptr, double r_px, T *out, hipStream_t st);
// MLP forward on a device-resident row count (olpt_mlp.hip), over nf frames
// whose inputs, counts and outputs lie xs, ns, ys bytes apart.
void mlp_infer_device(const void *mlp, const void *X, const int32_t *d_n, int cap, float *Y,
                      hipStream_t st, int nf = 1, long xs = 0, long ns = 0, long ys = 0);

template <typename T>
void launch_convert_image(const void *src, int dtype, long n, T *dst,
                          hipStream_t st);

// What modules built on a context (tracking, OTF estimate) read from it.
struct CtxInfo {
  hipStream_t st;
  int prec, n_cam, H, W, cap_obj;
  const olpt_camera *cams;     // host copies
  const CamT<double> *d_cam64; // device copies (FP64)
};
// False when the context has no workspace or cameras yet.
bool ctx_info(const olpt_ctx *c, CtxInfo *out);
// Set the thread-local message read by olpt_last_error and return s.
int32_t set_error(int32_t s, const char *msg);

} // namespace olpt
