// olpt_api.hip - C-ABI, device workspace and IPR orchestration.
//
// IPR::runIPR (IPR.cpp:195-306) becomes a fixed sequence of kernel launches
// per loop: peaks -> stereo -> shake -> filter -> IPR-mode residual ->
// append.  Every count stays in device memory, so the same sequence is what
// olpt_ipr_graph_capture records into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "olpt_c.h"
#include "olpt_kernels.hpp"
#include "olpt_scan.hpp"

using namespace olpt;

namespace olpt {
thread_local StageTimer *g_timer = nullptr;
}

namespace {

thread_local std::string g_err;

olpt_status fail(olpt_status s, const std::string &msg) {
  g_err = msg;
  return s;
}

#define HIP_TRY(x)                                                             \
  do {                                                                         \
    hipError_t _e = (x);                                                       \
    if (_e != hipSuccess)                                                      \
      return fail(OLPT_E_GPU, std::string(#x ": ") + hipGetErrorString(_e));   \
  } while (0)

constexpr int kMaxRounds = 64;

olpt_status read_flags(olpt_ctx *c, int nw, hipStream_t st, int32_t out[8]);

std::string prune_ties_msg(const int32_t *fl) {
  return std::string("pruneMatch: ") +
         (fl[2] ? "values within the eq_eps tolerance (rtol 1e-9) are not transitive here, so "
                  "the reference's order is the one its std::stable_sort happens to produce"
                : "the replacement branch (StereoMatch.cpp:906-968) would fire") +
         "; this selection is not reproduced";
}

} // namespace

// Per-stream workspace: everything one frame's IPR writes.  Frames of one
// olpt_ipr_run call are spread over the workspaces, each on its own stream,
// so the many small kernels of one frame overlap with another frame's work.
struct Workspace {
  hipStream_t st = nullptr;     // own stream (workspace 0: the context stream)
  hipEvent_t done = nullptr;
  // one arena of nf_cap frame slices, fstride bytes each; every pointer
  // below is frame 0's (frame f's buffer lies f * fstride bytes further)
  void *arena = nullptr;
  int nf_cap = 0;
  long fstride = 0;
  long lstride = 0, pstride = 0;  // learned-proposal buffers (own allocations)
  // images (sized for FP64)
  void *img_a = nullptr, *img_b = nullptr, *img_res = nullptr;
  // peaks
  uint64_t *pk_mask = nullptr;
  int32_t *pk_cnt = nullptr, *pk_off = nullptr;
  // stereo
  StereoBuffers sb{};
  double *xy = nullptr;
  int32_t *npts = nullptr;
  // limit2DObjectNumber work buffers
  int32_t *lim_perm = nullptr, *lim_flag = nullptr;
  double *lim_xy = nullptr;
  // shake
  void *sh_uv = nullptr, *keep_uv = nullptr;
  double *sh_xyz = nullptr, *keep_xyz = nullptr, *sh_score = nullptr;
  uint8_t *sh_flag = nullptr;
  int32_t *sh_n = nullptr, *n_keep = nullptr;
  int32_t *bin_cnt = nullptr, *bin_start = nullptr, *obj_bin = nullptr,
          *obj_rank = nullptr;
  int32_t *g_cnt = nullptr, *g_start = nullptr, *g_obj = nullptr, *g_rank = nullptr,
          *g_cellof = nullptr, *has_nb = nullptr, *sh_keep_off = nullptr;
  void *rec = nullptr, *bin_rec = nullptr;
  double *partial = nullptr;
  void *patches = nullptr;   // learned proposal: bf16 [cap_obj][512]
  float *prop = nullptr;     // [cap_obj][4]
  // misc
  int32_t *flags = nullptr;  // [16]
  int32_t *scan_scratch = nullptr;
  int32_t *totals = nullptr; // [64]
};

constexpr int kMaxWorkspaces = 8;
constexpr int kMaxBatch = 1024;

struct olpt_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  int prec = OLPT_PREC_F32;
  int n_cam = 0, H = 0, W = 0, cap2d = 0, cap_cand = 0, cap_obj = 0;
  int cams_set = 0, otf_set = 0;
  olpt_camera cams[OLPT_MAX_CAM];
  std::vector<void *> allocs;
  // cameras / OTF on device
  CamT<double> *d_cam64 = nullptr;
  CamT<float> *d_cam32 = nullptr;
  int otf_nx = 2, otf_ny = 2, otf_nz = 2;
  double otf_bnd[6] = {0, 0, 0, 0, 0, 0};
  double *d_otf64 = nullptr; // a | b | c | alpha, each n_cam*ng
  float *d_otf32 = nullptr;
  int otf_ng = 0;
  // an OTF with the same (a, b, c, alpha) at every grid node of each camera:
  // per camera (a, b, c, cos alpha, sin alpha) in FP32 for the FP32 lookups
  int otf_uniform = 0;
  float *d_otf_uni32 = nullptr;
  int bin_rows = 0, bin_cols = 0;
  int grid_n = 80;
  // learned shake proposal (not owned)
  const olpt_mlp *mlp = nullptr;
  double mlp_voxel_mm = 0.0, mlp_max_step_vox = 0.5;
  double mlp_ghost_logit = INFINITY;  // olpt_set_ghost_logit (off by default)
  // workspaces
  Workspace ws[kMaxWorkspaces];
  int n_ws = 0;          // allocated
  int n_ws_want = 1;     // workspaces to keep allocated
  int n_ws_active = 1;   // workspaces olpt_ipr_run spreads frames over
  hipEvent_t ev_fork = nullptr;
  int32_t *loop_tot = nullptr; // [n_frames*cap_loop] device
  int loop_tot_cap = 0;
  // stage timing and algorithmic-byte accounting
  StageTimer *timer = nullptr;
  bool stats = false;
  double st_bytes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t st_launch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // graph
  // one captured graph per workspace (per stream)
  hipGraph_t graph[kMaxWorkspaces] = {};
  hipGraphExec_t gexec[kMaxWorkspaces] = {};
  int n_graph = 0;
  bool capturing = false;
  // host time spent issuing olpt_ipr_run's work (not waiting for it)
  double host_ms = 0.0;
  int32_t host_calls = 0;
  // launch shapes (olpt_set_tuning)
  olpt_tuning tune = {1, 4096, 1024, 512, 1};
  int batch = 1;  // frames per workspace launch (arena slices)
};

namespace {

olpt_status dalloc(olpt_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(OLPT_E_ALLOC, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  c->allocs.push_back(*p);
  return OLPT_OK;
}

template <typename P> olpt_status dalloc_t(olpt_ctx *c, P **p, size_t count) {
  return dalloc(c, reinterpret_cast<void **>(p), count * sizeof(P));
}

void timer_free(olpt_ctx *c) {
  if (!c->timer) return;
  for (int a = 0; a < 2; ++a)
    for (int s = 0; s < 8; ++s)
      for (int i = 0; i < kStageEvents; ++i) (void)hipEventDestroy(c->timer->ev[a][s][i]);
  delete c->timer;
  c->timer = nullptr;
}

// RAII: route stage events of this thread to the context's timer
struct TimerScope {
  TimerScope(olpt_ctx *c, bool capturing) {
    g_timer = (!capturing && c->timer && c->timer->on) ? c->timer : nullptr;
  }
  ~TimerScope() { g_timer = nullptr; }
};

void free_streams(olpt_ctx *c);

void free_graphs(olpt_ctx *c) {
  for (int k = 0; k < kMaxWorkspaces; ++k) {
    if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
    if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
    c->gexec[k] = nullptr;
    c->graph[k] = nullptr;
  }
  c->n_graph = 0;
}

void free_workspace_memory(olpt_ctx *c) {
  for (int k = 0; k < kMaxWorkspaces; ++k) {
    Workspace &w = c->ws[k];
    if (w.arena) (void)hipFree(w.arena);
    if (w.patches) (void)hipFree(w.patches);
    if (w.prop) (void)hipFree(w.prop);
    w.arena = nullptr;
    w.patches = nullptr;
    w.prop = nullptr;
  }
}

void free_all(olpt_ctx *c) {
  free_graphs(c);
  free_workspace_memory(c);
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
}

template <typename TIn, typename T>
__global__ void k_convert(const TIn *__restrict__ src, long n, T *__restrict__ dst, long sstride,
                          long dstride) {
  src = fshift(src, (long)blockIdx.z * sstride);
  dst = fshift(dst, (long)blockIdx.z * dstride);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    dst[i] = (T)src[i];
}

// n pixels of each of nf frames: source frames sstride bytes apart,
// destination frames dstride bytes apart
template <typename T>
olpt_status convert_image(const void *src, int dtype, long n, T *dst, hipStream_t st, int nf = 1,
                          long sstride = 0, long dstride = 0) {
  const int threads = 256;
  long blocks = (n + threads - 1) / threads;
  if (blocks > 8192) blocks = 8192;
  const dim3 g((unsigned)blocks, 1, (unsigned)nf);
  switch (dtype) {
  case OLPT_U8: k_convert<uint8_t, T><<<g, threads, 0, st>>>((const uint8_t *)src, n, dst, sstride, dstride); break;
  case OLPT_U16: k_convert<uint16_t, T><<<g, threads, 0, st>>>((const uint16_t *)src, n, dst, sstride, dstride); break;
  case OLPT_F32: k_convert<float, T><<<g, threads, 0, st>>>((const float *)src, n, dst, sstride, dstride); break;
  case OLPT_F64: k_convert<double, T><<<g, threads, 0, st>>>((const double *)src, n, dst, sstride, dstride); break;
  default: return fail(OLPT_E_UNSUPPORTED, "unsupported image dtype");
  }
  return OLPT_OK;
}

__global__ void k_fill_i32(int32_t *p, long n, int32_t v, long fstride) {
  p = fshift(p, (long)blockIdx.z * fstride);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}
// n ints at p set to v, in each of nf frames fstride bytes apart
void fill_i32(int32_t *p, long n, int32_t v, hipStream_t st, int nf = 1, long fstride = 0) {
  long blocks = (n + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  k_fill_i32<<<dim3((unsigned)blocks, 1, (unsigned)nf), 256, 0, st>>>(p, n, v, fstride);
}

// append kept objects of one loop to the frame output (frame blockIdx.z:
// workspace buffers fstride bytes apart, outputs cap objects apart)
template <typename T>
__global__ void k_append(const double *__restrict__ kx, const T *__restrict__ kuv,
                         const int32_t *__restrict__ n_keep, int n_cam,
                         double *__restrict__ ox, double *__restrict__ ouv, int cap,
                         const int32_t *__restrict__ n_out, long fstride) {
  const long f = blockIdx.z;
  kx = fshift(kx, f * fstride);
  kuv = fshift(kuv, f * fstride);
  n_keep = fshift(n_keep, f * fstride);
  ox += f * (long)cap * 3;
  if (ouv) ouv += f * (long)cap * n_cam * 2;
  n_out += f;
  const int base = *n_out;
  const int n = *n_keep;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int o = base + i;
    if (o >= cap) break;
    for (int d = 0; d < 3; ++d) ox[3L * o + d] = kx[3L * i + d];
    if (ouv)
      for (int k = 0; k < 2 * n_cam; ++k) ouv[((long)o * n_cam) * 2 + k] = (double)kuv[(long)i * n_cam * 2 + k];
  }
}

__global__ void k_append_count(const int32_t *n_keep, int cap, int32_t *n_out,
                               int32_t *loop_tot, int loop_stride, int32_t *flags,
                               long fstride) {
  if (threadIdx.x || blockIdx.x) return;
  const long f = blockIdx.z;
  n_keep = fshift(n_keep, f * fstride);
  flags = fshift(flags, f * fstride);
  n_out += f;
  if (loop_tot) loop_tot += f * loop_stride;
  int t = *n_out + *n_keep;
  if (t > cap) {
    t = cap;
    atomicOr(&flags[4], 1);
  }
  *n_out = t;
  if (loop_tot) *loop_tot = t;
}

__global__ void k_sel_flags(const int8_t *state, const int32_t *n_cand, int cap,
                            uint8_t *sel) {
  const int M = min(*n_cand, cap);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < M; j += gridDim.x * blockDim.x)
    sel[j] = state[j] == 1 ? 1 : 0;
}

__global__ void k_peak_overflow(const int32_t *pk_flag, int32_t *flags, long fstride) {
  pk_flag = fshift(pk_flag, (long)blockIdx.z * fstride);
  flags = fshift(flags, (long)blockIdx.z * fstride);
  if (threadIdx.x == 0 && blockIdx.x == 0 && pk_flag[0]) atomicOr(&flags[7], 1);
}

template <typename T> const CamT<T> *dcams(olpt_ctx *c);
template <> const CamT<double> *dcams<double>(olpt_ctx *c) { return c->d_cam64; }
template <> const CamT<float> *dcams<float>(olpt_ctx *c) { return c->d_cam32; }

template <typename T> OtfT<T> dotf(olpt_ctx *c) {
  OtfT<T> o;
  o.nx = c->otf_nx;
  o.ny = c->otf_ny;
  o.nz = c->otf_nz;
  o.ng = c->otf_ng;
  for (int i = 0; i < 6; ++i) o.bnd[i] = (T)c->otf_bnd[i];
  const int n[3] = {c->otf_nx, c->otf_ny, c->otf_nz};
  for (int a = 0; a < 3; ++a) {
    // the same T operations as the device's (B[2a+1] - B[2a]) / (n - 1)
    o.step[a] = (o.bnd[2 * a + 1] - o.bnd[2 * a]) / (T)(n[a] - 1);
    o.inv_step[a] = (T)1 / o.step[a];
  }
  const long sz = (long)c->n_cam * c->otf_ng;
  const T *base;
  if constexpr (sizeof(T) == 8)
    base = (const T *)c->d_otf64;
  else
    base = (const T *)c->d_otf32;
  if constexpr (sizeof(T) == 4) {
    o.uniform = c->otf_uniform;
    o.uni = (const T *)c->d_otf_uni32;
  } else {
    o.uniform = 0;
    o.uni = nullptr;
  }
  o.a = base;
  o.b = base + sz;
  o.c = base + 2 * sz;
  o.alpha = base + 3 * sz;
  return o;
}

template <typename T>
ShakeBuffers<T> shake_buffers(olpt_ctx *c, Workspace &w, T *img, T *img_next, int nf = 1) {
  ShakeBuffers<T> b{};
  b.nf = nf;
  b.fstride = w.fstride;
  b.lstride = w.lstride;
  b.pstride = w.pstride;
  b.grid_shake = c->tune.shake_blocks;
  b.n_cam = c->n_cam;
  b.H = c->H;
  b.W = c->W;
  b.cams = dcams<T>(c);
  b.otf = dotf<T>(c);
  b.img = img;
  b.res = (T *)w.img_res;
  b.img_next = img_next;
  b.cap_obj = c->cap_obj;
  b.xyz = w.sh_xyz;
  b.uv = (T *)w.sh_uv;
  b.score = w.sh_score;
  b.flag = w.sh_flag;
  b.n_obj = w.sh_n;
  b.bin_px = 8;
  b.bin_rows = c->bin_rows;
  b.bin_cols = c->bin_cols;
  b.bin_cnt = w.bin_cnt;
  b.bin_start = w.bin_start;
  b.bin_rec = w.bin_rec;
  b.obj_bin = w.obj_bin;
  b.obj_rank = w.obj_rank;
  b.grid_n = c->grid_n;
  b.g_cnt = w.g_cnt;
  b.g_start = w.g_start;
  b.g_obj = w.g_obj;
  b.g_rank = w.g_rank;
  b.g_cellof = w.g_cellof;
  b.has_nb = w.has_nb;
  b.rec = w.rec;
  b.partial = w.partial;
  b.keep_off = w.sh_keep_off;
  b.n_keep = w.n_keep;
  b.scan_scratch = w.scan_scratch;
  b.totals = w.totals;
  b.work = w.sb.work;
  b.patches = c->mlp ? w.patches : nullptr;
  b.prop = c->mlp ? w.prop : nullptr;
  return b;
}

void set_grid(olpt_ctx *c, ShakeBuffers<float> *bf, ShakeBuffers<double> *bd,
              const olpt_ipr_params *p) {
  // repeated-object grid over the view volume (+ margin); cells >= tol_3d
  double lo[3], ext = 0;
  for (int d = 0; d < 3; ++d) {
    lo[d] = p->limit[2 * d] - 1.0;
    ext = std::max(ext, p->limit[2 * d + 1] - p->limit[2 * d] + 2.0);
  }
  double cell = std::max(p->tol_3d_mm, ext / c->grid_n);
  if (!(cell > 0)) cell = 1.0;
  auto fill = [&](auto *b) {
    if (!b) return;
    for (int d = 0; d < 3; ++d) b->grid_lo[d] = lo[d];
    b->grid_cell = cell;
  };
  fill(bf);
  fill(bd);
}

StereoParams stereo_params(olpt_ctx *c, const int32_t *active, const olpt_ipr_params *p,
                           double tol2d, olpt_status *st) {
  StereoParams sp{};
  sp.n_active = 0;
  for (int k = 0; k < OLPT_MAX_CAM; ++k) sp.active[k] = 0;
  for (int k = 0; k < c->n_cam; ++k) {
    sp.active[k] = active ? (active[k] != 0) : 1;
    sp.n_active += sp.active[k];
  }
  sp.tol2d = tol2d;
  sp.tol3d = p->tol_3d_mm;
  for (int i = 0; i < 6; ++i) sp.limit[i] = p->limit[i];
  *st = OLPT_OK;
  // StereoMatch ctor pre-checks (StereoMatch.cpp:27-47)
  if (sp.n_active < 2) {
    *st = fail(OLPT_E_INVALID_ARGUMENT, "preCheck: need at least 2 active cameras.");
    return sp;
  }
  if (p->match_cam_count < 2 || p->match_cam_count > sp.n_active) {
    *st = fail(OLPT_E_INVALID_ARGUMENT, "preCheck: match_cam_count must be within [2," +
                                            std::to_string(sp.n_active) + "].");
    return sp;
  }
  const int bc = std::min(std::max(2, p->match_cam_count), sp.n_active);
  sp.build_count = bc;
  if (bc != sp.n_active) {
    // with check cameras the reference's checkMatch leaves e_check NaN
    // (StereoMatch.cpp:522, :603-604) and match() raises InvalidArgument on
    // the first candidate that passes it (REQUIRE, :126-127)
    *st = fail(OLPT_E_INVALID_ARGUMENT,
               "StereoMatch::match: checkMatch returned NaN e_check. (match_cam_count below the "
               "number of active cameras)");
  }
  return sp;
}

ShakeParams shake_params(olpt_ctx *c, const int32_t *active, const olpt_ipr_params *p) {
  ShakeParams sp{};
  for (int k = 0; k < OLPT_MAX_CAM; ++k) sp.active[k] = 0;
  for (int k = 0; k < c->n_cam; ++k) sp.active[k] = active ? (active[k] != 0) : 1;
  sp.shake_width = p->shake_width;
  sp.shakewidth_min = p->shakewidth_min;
  sp.thred_ghost = p->thred_ghost;
  sp.tol3d = p->tol_3d_mm;
  sp.r_px = p->radius_obj;
  sp.n_shake_loop = p->n_shake_loop;
  sp.ratio_augimg = p->ratio_augimg;
  {
    const double d = std::max(0.0, p->radius_obj);   // calROISize, Shake.cpp:294-296
    sp.box_w = box_width(d);
    sp.roi_w = box_width(d * (double)p->ratio_augimg);
  }
  sp.mlp = c->mlp;
  sp.voxel_mm = c->mlp_voxel_mm;
  sp.max_step_vox = c->mlp_max_step_vox;
  sp.ghost_logit = c->mlp_ghost_logit;
  return sp;
}

olpt_status check_ready(olpt_ctx *c) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (!c->cams_set) return fail(OLPT_E_INVALID_ARGUMENT, "cameras not set");
  if (!c->n_ws || !c->ws[0].xy) return fail(OLPT_E_INVALID_ARGUMENT, "workspace not reserved");
  return OLPT_OK;
}

olpt_status check_params(olpt_ctx *c, const olpt_ipr_params *p) {
  if (!p) return fail(OLPT_E_INVALID_ARGUMENT, "null params");
  if (p->radius_obj <= 0 || p->ratio_augimg <= 0)
    return fail(OLPT_E_INVALID_ARGUMENT, "radius_obj and ratio_augimg must be positive");
  const int box = box_width(p->radius_obj);
  const int roi = box_width(p->radius_obj * (double)p->ratio_augimg);
  if (roi * roi > 4 * 64 || box > 24)
    return fail(OLPT_E_UNSUPPORTED, "ROI larger than 16x16 pixels is not supported");
  if (c->mlp && roi > 8)
    return fail(OLPT_E_UNSUPPORTED, "the learned shake proposal supports ROIs up to 8x8 pixels");
  if (p->idmap_cell_px < 1) return fail(OLPT_E_INVALID_ARGUMENT, "idmap_cell_px must be >= 1");
  if (p->idmap_cell_px != c->ws[0].sb.cell_px)
    return fail(OLPT_E_UNSUPPORTED, "idmap_cell_px differs from the reserved workspace (4)");
  return OLPT_OK;
}

// One IPR loop on the active subset for the nf frames of a workspace batch:
// img (loop input, T) -> img_next.  Outputs of frame f: out_xyz + f * cap_out
// objects, d_n_out[f], d_loop_tot[f * loop_stride].
template <typename T>
olpt_status ipr_loop(olpt_ctx *c, Workspace &w, int nf, T *img, T *img_next,
                     const int32_t *active, const olpt_ipr_params *p, double tol2d,
                     double *out_xyz, double *out_uv, int cap_out, int32_t *d_n_out,
                     int32_t *d_loop_tot, int loop_stride) {
  hipStream_t st = w.st;
  const long fs = w.fstride;
  olpt_status s;
  StereoParams sp = stereo_params(c, active, p, tol2d, &s);
  if (s) return s;
  CamList cl{0, {0}};
  for (int k = 0; k < c->n_cam; ++k)
    if (sp.active[k]) cl.id[cl.n++] = k;
  const long HW = (long)c->H * c->W;
  // 1) ObjectFinder2D on the active cameras (inactive slots stay empty)
  fill_i32(w.npts, c->n_cam, 0, st, nf, fs);
  launch_find_peaks<T>(img, HW, c->H, c->W, (double)p->min_obj_int, cl, w.pk_mask, w.pk_cnt,
                       w.pk_off, w.scan_scratch, w.totals, w.xy, c->cap2d, w.npts,
                       w.flags + 8, st, nf, fs);
  k_peak_overflow<<<dim3(1, 1, nf), 1, 0, st>>>(w.flags + 8, w.flags, fs);
  // 1.1) limit2DObjectNumber (IPR.cpp:105-107)
  if (p->n_obj2d_process_max > 0 && p->n_obj2d_process_max < c->cap2d)
    launch_limit2d(w.xy, w.npts, cl, p->n_obj2d_process_max, c->cap2d, w.lim_perm, w.lim_flag,
                   w.lim_xy, st, nf, fs);
  // 2) StereoMatch
  StereoBuffers sb = w.sb;
  sb.nf = nf;
  launch_stereo(sb, sp, kMaxRounds, st);
  // 3) Shake the matched objects in place
  ShakeBuffers<T> b = shake_buffers<T>(c, w, img, img_next, nf);
  b.xyz = w.sb.obj_xyz;
  b.n_obj = w.sb.n_obj;
  if constexpr (sizeof(T) == 8)
    set_grid(c, nullptr, &b, p);
  else
    set_grid(c, &b, nullptr, p);
  ShakeParams shp = shake_params(c, active, p);
  launch_shake<T>(b, shp, st);
  // 4) filterOutInvalid + calResidualImage (IPR mode)
  launch_filter<T>(b, w.keep_xyz, (T *)w.keep_uv, w.n_keep, st);
  launch_residual_ipr<T>(b, w.n_keep, w.keep_xyz, (T *)w.keep_uv, p->radius_obj, st);
  // 5) append
  if (out_xyz) {
    k_append<T><<<dim3(256, 1, nf), 256, 0, st>>>(w.keep_xyz, (const T *)w.keep_uv, w.n_keep,
                                                   c->n_cam, out_xyz, out_uv, cap_out, d_n_out,
                                                   fs);
    k_append_count<<<dim3(1, 1, nf), 1, 0, st>>>(w.n_keep, cap_out, d_n_out, d_loop_tot,
                                                  loop_stride, w.flags, fs);
  }
  HIP_TRY(hipGetLastError());
  if (c->stats) {
    // Algorithmic bytes per launch (what each stage must move at minimum),
    // summed over the batch's frames:
    //  peaks:    read the loop image of the active cameras + write the peaks
    //  build:    read 2D points + LOS of the active cameras + the cell CSR,
    //            write the checked candidates
    //  prune:    read/write the candidates twice + point lists + objects
    //  shake:    per iteration, the N x C x P^2 patch gathers of the
    //            reference ROI (P = 2 round(r ratio)) + positions
    //  residual: read the loop image and write the residual per camera
    //  enum2:    the enumeration of the level-1 survivors' next camera:
    //            each survivor's path record read, each level-2 item written,
    //            and the line table once (each 2D point's lines on the other
    //            cameras, 32 B each)
    const double s_t = sizeof(T);
    const double na = cl.n, nc = c->n_cam;
    const double ncell = (double)w.sb.rows_cell * w.sb.cols_cell;
    const double P = std::round(2.0 * p->radius_obj * p->ratio_augimg);
    for (int f = 0; f < nf; ++f) {
      int32_t np[OLPT_MAX_CAM], cnt[3], lv[3] = {0, 0, 0};
      const long o = (long)f * fs;
      // level-1 survivors, level-2 items
      HIP_TRY(hipMemcpyAsync(&lv[0], fshift(w.sb.n_surv, o) + 1, 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(&lv[1], fshift(w.sb.n_path, o) + 2, 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(np, fshift(w.npts, o), sizeof(np), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(&cnt[0], fshift(w.sb.n_cand, o), 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(&cnt[1], fshift(w.sb.n_obj, o), 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(&cnt[2], fshift(w.n_keep, o), 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      double n2d = 0;
      for (int k = 0; k < c->n_cam; ++k) n2d += np[k];
      const double M = cnt[0], N = cnt[1];
      c->st_bytes[ST_PEAKS] += na * HW * s_t + n2d * 16.0;
      c->st_bytes[ST_BUILD] += n2d * 48.0 + na * ncell * 8.0 + M * (4.0 * nc + 8.0);
      c->st_bytes[ST_PRUNE] += 2.0 * M * (4.0 * nc + 8.0) + 2.0 * M * na * 4.0 + N * (24.0 + 4.0 * nc);
      c->st_bytes[ST_SHAKE] += p->n_shake_loop * (N * na * P * P * s_t + N * 48.0);
      c->st_bytes[ST_RESID] += p->n_shake_loop * (na * HW * 2.0 * s_t) + nc * HW * 2.0 * s_t;
      if (sp.build_count > 2)
        c->st_bytes[ST_ENUM2] += ((double)lv[0] + (double)lv[1]) * sizeof(PathRec<2>) +
                                 n2d * (na - 1.0) * 32.0;
    }
    c->st_launch[ST_PEAKS] += 1;
    c->st_launch[ST_BUILD] += 1;
    c->st_launch[ST_PRUNE] += 1;
    c->st_launch[ST_SHAKE] += p->n_shake_loop;
    c->st_launch[ST_RESID] += p->n_shake_loop + 1;
    if (sp.build_count > 2) c->st_launch[ST_ENUM2] += 1;
  }
  return OLPT_OK;
}

// IPR::runIPR on the nf frames of a workspace batch, already converted into
// img_a (T); frame f's outputs at out_xyz + f * cap_out objects, d_n_out[f]
// and d_loop_tot[f * cap_loop + loop].
template <typename T>
olpt_status ipr_frame(olpt_ctx *c, Workspace &w, int nf, const olpt_ipr_params *p,
                      double *out_xyz, double *out_uv, int cap_out, int32_t *d_n_out,
                      int32_t *d_loop_tot, int cap_loop, int *n_loop_out) {
  const int n_cam = c->n_cam;
  fill_i32(d_n_out, 1, 0, w.st, nf, (long)sizeof(int32_t));
  int n_loop = 0;
  if (n_cam < 2) { // IPR.cpp:200-204
    *n_loop_out = 0;
    return OLPT_OK;
  }
  T *img = (T *)w.img_a, *nxt = (T *)w.img_b;
  int max_drop = std::min(p->n_cam_reduced, std::max(0, n_cam - 2));
  olpt_ipr_params q = *p;
  const int orig_mcc = p->match_cam_count;
  for (int drop = 0; drop <= max_drop; ++drop) {
    const int K = n_cam - drop;
    const int loops = drop == 0 ? p->n_loop_ipr : p->n_loop_ipr_reduced;
    q.match_cam_count = std::min(orig_mcc, K);
    // lexicographic K-subsets (myMATH::generateCombinations, myMATH.cpp:39-71)
    std::vector<int> comb(K);
    for (int i = 0; i < K; ++i) comb[i] = i;
    while (true) {
      int32_t act[OLPT_MAX_CAM] = {0};
      for (int i = 0; i < K; ++i) act[comb[i]] = 1;
      double tol = q.tol_2d_px;
      for (int loop = 0; loop < loops; ++loop) {
        int32_t *lt = (d_loop_tot && n_loop < cap_loop) ? d_loop_tot + n_loop : nullptr;
        olpt_status s = ipr_loop<T>(c, w, nf, img, nxt, act, &q, tol, out_xyz, out_uv, cap_out,
                                    d_n_out, lt, cap_loop);
        if (s) return s;
        std::swap(img, nxt);
        ++n_loop;
        tol = tol * 1.5; // IPR.cpp:291-292
      }
      int i = K - 1;
      while (i >= 0 && comb[i] == n_cam - K + i) --i;
      if (i < 0) break;
      comb[i]++;
      for (int j = i + 1; j < K; ++j) comb[j] = comb[j - 1] + 1;
    }
  }
  *n_loop_out = n_loop;
  return OLPT_OK;
}

// OTF::getOTFParam at n points (FP64)
__global__ void k_otf_param(OtfT<double> o, int cam, const double *__restrict__ xyz, int n,
                            double *__restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double X[3] = {xyz[3L * i], xyz[3L * i + 1], xyz[3L * i + 2]};
    double P[5];
    otf_param(o, cam, X, P);
    for (int q = 0; q < 5; ++q) out[5L * i + q] = P[q];
  }
}

size_t tsize(olpt_ctx *c) { return c->prec == OLPT_PREC_F64 ? 8 : 4; }

} // namespace

namespace olpt {
bool ctx_info(const olpt_ctx *c, CtxInfo *out) {
  if (!c || !c->cams_set || !c->n_ws || !c->d_cam64) return false;
  out->st = c->st;
  out->prec = c->prec;
  out->n_cam = c->n_cam;
  out->H = c->H;
  out->W = c->W;
  out->cap_obj = c->cap_obj;
  out->cams = c->cams;
  out->d_cam64 = c->d_cam64;
  return true;
}
int32_t set_error(int32_t s, const char *msg) { return fail(s, msg); }
} // namespace olpt

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C" {

const char *olpt_last_error(void) { return g_err.c_str(); }

olpt_status olpt_ctx_create(int device, void *stream, olpt_ctx **out) {
  if (!out) return fail(OLPT_E_INVALID_ARGUMENT, "null out");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(OLPT_E_INVALID_ARGUMENT, "bad device index");
  HIP_TRY(hipSetDevice(device));
  olpt_ctx *c = new olpt_ctx();
  c->device = device;
  c->st = (hipStream_t)stream;
  *out = c;
  return OLPT_OK;
}

olpt_status olpt_ctx_destroy(olpt_ctx *c) {
  if (!c) return OLPT_OK;
  (void)hipSetDevice(c->device);
  timer_free(c);
  if (c->st)
    (void)hipStreamSynchronize(c->st);
  else
    (void)hipDeviceSynchronize();
  free_all(c);
  free_streams(c);
  delete c;
  return OLPT_OK;
}

void *olpt_stream(olpt_ctx *c) { return c ? (void *)c->st : nullptr; }

olpt_status olpt_sync(olpt_ctx *c) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  HIP_TRY(hipStreamSynchronize(c->st));
  return OLPT_OK;
}

olpt_status olpt_set_precision(olpt_ctx *c, int prec) {
  if (!c || (prec != OLPT_PREC_F32 && prec != OLPT_PREC_F64))
    return fail(OLPT_E_INVALID_ARGUMENT, "precision must be OLPT_PREC_F32 or OLPT_PREC_F64");
  c->prec = prec;
  return OLPT_OK;
}

} // extern "C"

namespace {

// Buffer offsets inside one frame slice of a workspace arena (256-byte
// aligned); resolved to pointers once the arena is allocated.
struct SliceLayout {
  size_t off = 0;
  std::vector<std::pair<void **, size_t>> at;
  void add(void **p, size_t bytes) {
    at.emplace_back(p, off);
    off += (std::max<size_t>(bytes, 16) + 255) / 256 * 256;
  }
};

// All buffers of one workspace, sized from the context's reserve, for
// c->batch frames: one arena of identical frame slices.
olpt_status alloc_workspace(olpt_ctx *c, Workspace &w) {
  const int H = c->H, W = c->W;
  const long HW = (long)H * W;
  const int nc = c->n_cam, cap = c->cap2d, M = c->cap_cand, O = c->cap_obj;
  const int F = c->batch;
  SliceLayout L;
#define A(ptr, T, n) L.add(reinterpret_cast<void **>(&(ptr)), sizeof(T) * (size_t)(n))
  L.add(&w.img_a, 8 * (size_t)nc * HW);
  L.add(&w.img_b, 8 * (size_t)nc * HW);
  L.add(&w.img_res, 8 * (size_t)nc * HW);
  const long nchunk = peak_chunks(H, W);
  A(w.pk_mask, uint64_t, nchunk * nc);
  A(w.pk_cnt, int32_t, nchunk * nc);
  A(w.pk_off, int32_t, nchunk * nc);
  A(w.xy, double, (long)nc * cap * 2);
  A(w.npts, int32_t, OLPT_MAX_CAM);
  A(w.lim_perm, int32_t, (long)nc * cap);
  A(w.lim_flag, int32_t, OLPT_MAX_CAM);
  A(w.lim_xy, double, (long)nc * cap * 2);
  StereoBuffers &b = w.sb;
  b = StereoBuffers{};
  b.n_cam = nc;
  b.cap = cap;
  b.cams = c->d_cam64;
  b.xy = w.xy;
  b.npts = w.npts;
  A(b.los, double, (long)nc * cap * 4);
  A(b.lines, double, (long)nc * nc * cap * 4);
  A(b.t1, int8_t, cap);
  b.cap_pairs = std::max<long>(1L << 22, 16L * M);
  A(b.pairs, int2, b.cap_pairs);
  A(b.n_pairs, int32_t, 1);
  L.add(&b.path[0], path_rec_bytes(nc) * (size_t)b.cap_pairs);
  L.add(&b.path[1], path_rec_bytes(nc) * (size_t)b.cap_pairs);
  A(b.n_path, int32_t, kMaxCam);
  A(b.n_surv, int32_t, kMaxCam);
  b.cell_px = 4;
  b.rows_cell = std::max(1, H / 4 + (H % 4 ? 1 : 0));
  b.cols_cell = std::max(1, W / 4 + (W % 4 ? 1 : 0));
  const long ncell = (long)b.rows_cell * b.cols_cell;
  A(b.cell_cnt, int32_t, ncell * nc);
  A(b.cell_start, int32_t, ncell * nc);
  A(b.cell_pid, int32_t, (long)nc * cap);
  A(b.out_cnt, int32_t, nc);
  A(b.cell_xy, double2, (long)nc * cap);
  A(b.pt_cell, int32_t, (long)nc * cap);
  A(b.pt_rank, int32_t, (long)nc * cap);
  b.cap_cand = M;
  A(b.cand_ids, int32_t, (long)M * nc);
  A(b.cand_e, double, M);
  A(b.n_cand, int32_t, 1);
  A(b.cand_ids_s, int32_t, (long)M * nc);
  A(b.cand_e_s, double, M);
  A(b.key_cnt, int32_t, cap);
  A(b.key_start, int32_t, cap);
  A(b.cand_rank, int32_t, M);
  A(b.bucket_idx, int32_t, M);
  A(b.bucket_ids, int32_t, (long)M * nc);
  A(b.pt_count, int32_t, (long)nc * cap);
  A(b.pt_start, int32_t, (long)nc * cap);
  A(b.pt_list, int32_t, (long)M * nc);
  A(b.pt_of, int32_t, (long)M * nc);
  A(b.pt_sorted, int32_t, (long)M * nc);
  A(b.pt_occ, int32_t, (long)M * nc);
  A(b.ppos, int32_t, (long)M * nc);
  A(b.pt_flag, int32_t, (long)nc * cap);
  A(b.head, int32_t, (long)nc * cap);
  A(b.own, unsigned long long, (long)nc * cap);
  A(b.n_entries, int32_t, 1);
  const long grp_n = grp_buckets((long)nc * cap) * (long)kGrpBlocks;
  A(b.grp_cnt, int32_t, grp_n);
  A(b.grp_off, int32_t, grp_n);
  A(b.e_ent, double, (long)M * nc);
  A(b.p_ent, double, (long)M * nc);
  A(b.st_ent, int32_t, (long)M * nc);
  A(b.pv, double, (long)M * nc);
  A(b.pct, double, M);
  A(b.state, int8_t, M);
  A(b.undecided, int32_t, kMaxRounds);
  A(b.keep, int32_t, M);
  A(b.keep_off, int32_t, M);
  A(b.tri_xyz, double, (long)M * 3);
  b.hz_size = stereo_hz_size(M);
  A(b.hz_tab, unsigned long long, b.hz_size);
  A(b.hz_slot, int32_t, 2L * M);
  b.cap_obj = O;
  A(b.obj_xyz, double, (long)O * 3);
  A(b.obj_ids, int32_t, (long)O * nc);
  A(b.n_obj, int32_t, 1);
  A(b.n_sel, int32_t, 1);
  A(w.flags, int32_t, 16);
  b.flags = w.flags;
  A(b.work, int32_t, 16);
  const long grid_cells = (long)c->grid_n * c->grid_n * c->grid_n;
  long scan_n = std::max<long>(std::max<long>(M, (long)nc * cap), std::max<long>(ncell, grid_cells));
  scan_n = std::max<long>(scan_n, nchunk);
  scan_n = std::max<long>(scan_n, grp_n);
  scan_n = std::max<long>(scan_n, (long)c->bin_rows * c->bin_cols * nc);
  A(w.scan_scratch, int32_t, scan_scratch_ints(scan_n, nc) + 64);
  b.scan_scratch = w.scan_scratch;
  A(w.totals, int32_t, 64);
  b.totals = w.totals;
  b.grid_build = c->tune.build_blocks;
  b.grid_prune = c->tune.prune_blocks;
  // shake
  L.add(&w.sh_uv, 8 * (size_t)O * nc * 2);
  L.add(&w.keep_uv, 8 * (size_t)O * nc * 2);
  A(w.sh_xyz, double, (long)O * 3);
  A(w.keep_xyz, double, (long)O * 3);
  A(w.sh_score, double, O);
  A(w.sh_flag, uint8_t, O);
  A(w.sh_n, int32_t, 1);
  A(w.n_keep, int32_t, 1);
  const long nbin = (long)c->bin_rows * c->bin_cols;
  A(w.bin_cnt, int32_t, nbin * nc);
  A(w.bin_start, int32_t, nbin * nc + 1);  // + the total: counts are start differences
  L.add(&w.bin_rec, shake_rec_bytes(true) * (size_t)O * nc);
  A(w.obj_bin, int32_t, (long)O * nc);
  A(w.obj_rank, int32_t, (long)O * nc);
  A(w.g_cnt, int32_t, grid_cells);
  A(w.g_start, int32_t, grid_cells + 1);
  A(w.g_obj, int32_t, O);
  A(w.g_rank, int32_t, O);
  A(w.g_cellof, int32_t, O);
  A(w.has_nb, int32_t, O);
  A(w.sh_keep_off, int32_t, O);
  L.add(&w.rec, shake_rec_bytes(true) * (size_t)O * nc);
  A(w.partial, double, 1024);
#undef A
  // the arena, then every buffer at its offset in frame 0's slice
  w.fstride = (long)L.off;
  w.nf_cap = F;
  {
    hipError_t e = hipMalloc(&w.arena, (size_t)F * L.off);
    if (e != hipSuccess) {
      w.arena = nullptr;
      return fail(OLPT_E_ALLOC, "hipMalloc(" + std::to_string((size_t)F * L.off) +
                                    ") for a workspace of " + std::to_string(F) +
                                    " frames: " + hipGetErrorString(e));
    }
  }
  for (auto &a : L.at) *a.first = (char *)w.arena + a.second;
  // stereo views of the arena (pointers set above through the layout)
  b.xy = w.xy;
  b.npts = w.npts;
  b.flags = w.flags;
  b.scan_scratch = w.scan_scratch;
  b.totals = w.totals;
  b.nf = 1;
  b.fstride = w.fstride;
  // device counters and flags start at zero (fresh allocations may hold
  // another context's old values), in every frame slice
  const long fs = w.fstride;
  fill_i32(w.flags, 16, 0, nullptr, F, fs);
  fill_i32(w.npts, OLPT_MAX_CAM, 0, nullptr, F, fs);
  StereoBuffers bf = b;
  bf.nf = F;
  stereo_init_tables(bf, nullptr);
  // count arrays of the count -> scan -> fill passes: zero here, and each
  // pass's scan leaves them zero again
  fill_i32(w.bin_cnt, nbin * nc, 0, nullptr, F, fs);
  fill_i32(w.g_cnt, grid_cells, 0, nullptr, F, fs);
  return OLPT_OK;
}

olpt_status alloc_learned(olpt_ctx *c, Workspace &w) {
  if (w.patches) return OLPT_OK;
  w.lstride = (long)c->cap_obj * 512 * 2;
  w.pstride = (long)c->cap_obj * 4 * 4;
  HIP_TRY(hipMalloc(&w.patches, (size_t)w.nf_cap * w.lstride));
  HIP_TRY(hipMalloc((void **)&w.prop, (size_t)w.nf_cap * w.pstride));
  return OLPT_OK;
}

void free_streams(olpt_ctx *c) {
  for (int i = 1; i < kMaxWorkspaces; ++i) {
    if (c->ws[i].st) (void)hipStreamDestroy(c->ws[i].st);
    c->ws[i].st = nullptr;
  }
  for (int i = 0; i < kMaxWorkspaces; ++i) {
    if (c->ws[i].done) (void)hipEventDestroy(c->ws[i].done);
    c->ws[i].done = nullptr;
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  c->ev_fork = nullptr;
}

// Allocate workspaces up to the requested concurrency; extra ones get a
// stream of their own (default priority, non-blocking) and a join event.
olpt_status ensure_workspaces(olpt_ctx *c) {
  if (!c->cap2d) return OLPT_OK;
  if (!c->ev_fork) HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  while (c->n_ws < c->n_ws_want) {
    Workspace &w = c->ws[c->n_ws];
    if (c->n_ws == 0)
      w.st = c->st;
    else if (!w.st)
      HIP_TRY(hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking));
    if (!w.done) HIP_TRY(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
    olpt_status s = alloc_workspace(c, w);
    if (s) return s;
    if (c->mlp && (s = alloc_learned(c, w))) return s;
    ++c->n_ws;
  }
  HIP_TRY(hipDeviceSynchronize());
  return OLPT_OK;
}

} // namespace

extern "C" {

olpt_status olpt_ctx_reserve(olpt_ctx *c, int n_cam, int H, int W, int max_2d, int max_cand,
                             int max_obj) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (n_cam < 1 || n_cam > OLPT_MAX_CAM || H < 3 || W < 3)
    return fail(OLPT_E_INVALID_ARGUMENT, "bad camera count or image size");
  if (max_cand >= (1 << 27))
    return fail(OLPT_E_INVALID_ARGUMENT, "max_cand must be below 2^27");
  if (max_2d > (1 << 20))
    return fail(OLPT_E_INVALID_ARGUMENT, "max_2d must be at most 2^20");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipDeviceSynchronize());
  free_all(c);
  c->n_cam = n_cam;
  c->H = H;
  c->W = W;
  // default: every possible local maximum of an image up to 724^2 pixels
  // (at most one in two pixels), and 262,144 above that - more than the
  // reference's default n_obj2d_process_max (200,000), so its
  // limit2DObjectNumber subsampling runs as the reference's does
  c->cap2d = max_2d > 0 ? max_2d : std::min(1 << 18, std::max(1024, H * W / 2));
  c->cap_cand = max_cand > 0 ? max_cand : (1 << 21);
  c->cap_obj = max_obj > 0 ? max_obj : (1 << 18);
  c->bin_rows = (H + 7) / 8;
  c->bin_cols = (W + 7) / 8;
  olpt_status s;
  if ((s = dalloc_t(c, &c->d_cam64, OLPT_MAX_CAM))) return s;
  if ((s = dalloc_t(c, &c->d_cam32, OLPT_MAX_CAM))) return s;
  for (int i = 0; i < kMaxWorkspaces; ++i) {
    hipStream_t keep_st = c->ws[i].st;
    hipEvent_t keep_ev = c->ws[i].done;
    c->ws[i] = Workspace{};
    c->ws[i].st = i == 0 ? c->st : keep_st;
    c->ws[i].done = keep_ev;
  }
  c->n_ws = 0;
  if ((s = ensure_workspaces(c))) return s;
  // re-upload cameras / OTF if they were set before
  if (c->cams_set) {
    CamT<double> h64[OLPT_MAX_CAM];
    CamT<float> h32[OLPT_MAX_CAM];
    for (int k = 0; k < c->n_cam; ++k) {
      h64[k] = make_cam<double>(c->cams[k]);
      h32[k] = make_cam<float>(c->cams[k]);
    }
    HIP_TRY(hipMemcpy(c->d_cam64, h64, sizeof(h64), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_cam32, h32, sizeof(h32), hipMemcpyHostToDevice));
  }
  c->otf_set = 0;
  c->d_otf64 = nullptr;
  c->d_otf32 = nullptr;
  c->otf_ng = 0;
  c->otf_uniform = 0;
  c->d_otf_uni32 = nullptr;
  c->loop_tot = nullptr;
  c->loop_tot_cap = 0;
  return OLPT_OK;
}

olpt_status olpt_set_concurrency(olpt_ctx *c, int n_streams) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (n_streams < 1 || n_streams > kMaxWorkspaces)
    return fail(OLPT_E_INVALID_ARGUMENT, "n_streams must be within [1, 8]");
  if (c->capturing) return fail(OLPT_E_INVALID_ARGUMENT, "cannot change concurrency while capturing");
  if (n_streams > c->n_ws_want) c->n_ws_want = n_streams;
  c->n_ws_active = n_streams;
  return ensure_workspaces(c);
}

olpt_status olpt_set_tuning(olpt_ctx *c, const olpt_tuning *t) {
  if (!c || !t) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (c->capturing) return fail(OLPT_E_INVALID_ARGUMENT, "cannot change tuning while capturing");
  const int32_t blk[3] = {t->shake_blocks, t->build_blocks, t->prune_blocks};
  for (int i = 0; i < 3; ++i)
    if (blk[i] < 64 || blk[i] > 65536)
      return fail(OLPT_E_INVALID_ARGUMENT, "block counts must be within [64, 65536]");
  if (t->n_streams < 1 || t->n_streams > kMaxWorkspaces)
    return fail(OLPT_E_INVALID_ARGUMENT, "n_streams must be within [1, 8]");
  if (t->frames_per_batch < 1 || t->frames_per_batch > kMaxBatch)
    return fail(OLPT_E_INVALID_ARGUMENT, "frames_per_batch must be within [1, 1024]");
  free_graphs(c);
  if (t->frames_per_batch != c->batch) {
    // workspaces hold that many frame slices: reallocate them
    HIP_TRY(hipDeviceSynchronize());
    free_workspace_memory(c);
    c->batch = t->frames_per_batch;
    c->n_ws_want = t->n_streams;
    for (int k = 0; k < c->n_ws; ++k) {
      Workspace &w = c->ws[k];
      hipStream_t keep_st = w.st;
      hipEvent_t keep_ev = w.done;
      w = Workspace{};
      w.st = keep_st;
      w.done = keep_ev;
    }
    c->n_ws = 0;
  }
  c->tune = *t;
  for (int k = 0; k < c->n_ws; ++k) {
    c->ws[k].sb.grid_build = t->build_blocks;
    c->ws[k].sb.grid_prune = t->prune_blocks;
  }
  return olpt_set_concurrency(c, t->n_streams);
}

olpt_status olpt_get_tuning(olpt_ctx *c, olpt_tuning *t) {
  if (!c || !t) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  *t = c->tune;
  t->n_streams = c->n_ws_active;
  t->frames_per_batch = c->batch;
  return OLPT_OK;
}

olpt_status olpt_set_cameras(olpt_ctx *c, const olpt_camera *cams, int n_cam) {
  if (!c || !cams) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (!c->d_cam64) return fail(OLPT_E_INVALID_ARGUMENT, "workspace not reserved");
  if (n_cam != c->n_cam) return fail(OLPT_E_INVALID_ARGUMENT, "camera count differs from reserve");
  CamT<double> h64[OLPT_MAX_CAM];
  CamT<float> h32[OLPT_MAX_CAM];
  for (int k = 0; k < n_cam; ++k) {
    const olpt_camera &q = cams[k];
    if (q.n_row != c->H || q.n_col != c->W)
      return fail(OLPT_E_INVALID_ARGUMENT, "camera image size differs from reserve");
    if (q.is_distorted && q.n_dist != 4 && q.n_dist != 5 && q.n_dist != 8 && q.n_dist != 12)
      return fail(OLPT_E_UNSUPPORTED, "number of distortion coefficients must be 4, 5, 8 or 12");
    c->cams[k] = q;
    h64[k] = make_cam<double>(q);
    h32[k] = make_cam<float>(q);
  }
  HIP_TRY(hipMemcpy(c->d_cam64, h64, sizeof(CamT<double>) * n_cam, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_cam32, h32, sizeof(CamT<float>) * n_cam, hipMemcpyHostToDevice));
  c->cams_set = 1;
  return OLPT_OK;
}

olpt_status olpt_set_otf(olpt_ctx *c, const olpt_otf *o) {
  if (!c || !o) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (o->n_cam != c->n_cam) return fail(OLPT_E_INVALID_ARGUMENT, "OTF camera count differs");
  if (o->nx < 2 || o->ny < 2 || o->nz < 2)
    return fail(OLPT_E_INVALID_ARGUMENT, "OTF grid number should be larger than 1");
  const int ng = o->nx * o->ny * o->nz;
  const long sz = (long)o->n_cam * ng;
  std::vector<double> h(4 * sz);
  std::vector<float> hf(4 * sz);
  const double *src[4] = {o->a, o->b, o->c, o->alpha};
  for (int q = 0; q < 4; ++q)
    for (long i = 0; i < sz; ++i) {
      h[q * sz + i] = src[q][i];
      hf[q * sz + i] = (float)src[q][i];
    }
  if (c->otf_ng != ng || !c->d_otf64) {
    olpt_status s;
    if ((s = dalloc_t(c, &c->d_otf64, 4 * sz))) return s;
    if ((s = dalloc_t(c, &c->d_otf32, 4 * sz))) return s;
  }
  HIP_TRY(hipMemcpy(c->d_otf64, h.data(), sizeof(double) * 4 * sz, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_otf32, hf.data(), sizeof(float) * 4 * sz, hipMemcpyHostToDevice));
  bool uniform = true;
  for (int q = 0; q < 4 && uniform; ++q)
    for (int k = 0; k < o->n_cam && uniform; ++k)
      for (int g = 1; g < ng && uniform; ++g)
        uniform = hf[q * sz + (long)k * ng + g] == hf[q * sz + (long)k * ng];
  if (uniform) {
    if (!c->d_otf_uni32) {
      olpt_status s;
      if ((s = dalloc_t(c, &c->d_otf_uni32, 5 * OLPT_MAX_CAM))) return s;
    }
    float u[5 * OLPT_MAX_CAM] = {};
    for (int k = 0; k < o->n_cam; ++k) {
      const long i = (long)k * ng;
      const float al = hf[3 * sz + i];
      u[5 * k + 0] = hf[i];
      u[5 * k + 1] = hf[sz + i];
      u[5 * k + 2] = hf[2 * sz + i];
      u[5 * k + 3] = al == 0.0f ? 1.0f : std::cos(al);
      u[5 * k + 4] = al == 0.0f ? 0.0f : std::sin(al);
    }
    HIP_TRY(hipMemcpy(c->d_otf_uni32, u, sizeof(u), hipMemcpyHostToDevice));
  }
  c->otf_uniform = uniform ? 1 : 0;
  c->otf_nx = o->nx;
  c->otf_ny = o->ny;
  c->otf_nz = o->nz;
  c->otf_ng = ng;
  for (int i = 0; i < 6; ++i) c->otf_bnd[i] = o->boundary[i];
  c->otf_set = 1;
  return OLPT_OK;
}

olpt_status olpt_diagnostics(olpt_ctx *c, int32_t out[8]) {
  if (!c || !c->n_ws) return fail(OLPT_E_INVALID_ARGUMENT, "no workspace");
  HIP_TRY(hipDeviceSynchronize());
  return read_flags(c, c->n_ws, c->st, out);
}

olpt_status olpt_build_counts(olpt_ctx *c, int32_t out[24]) {
  if (!c || !c->n_ws || !out) return fail(OLPT_E_INVALID_ARGUMENT, "no workspace");
  HIP_TRY(hipDeviceSynchronize());
  const StereoBuffers &b = c->ws[0].sb;
  int32_t v[24] = {0};
  HIP_TRY(hipMemcpy(&v[0], b.n_pairs, 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&v[1], b.n_surv, 4 * kMaxCam, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&v[1 + kMaxCam], b.n_path, 4 * kMaxCam, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&v[1 + 2 * kMaxCam], b.n_cand, 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&v[2 + 2 * kMaxCam], b.n_sel, 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&v[3 + 2 * kMaxCam], b.n_obj, 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < 24; ++i) out[i] = v[i];
  return OLPT_OK;
}

olpt_status olpt_find_peaks(olpt_ctx *c, const void *d_img, int dtype, int H, int W,
                            const olpt_ipr_params *p, double *d_xy, int cap, int32_t *n_out) {
  if (!c || !c->n_ws) return fail(OLPT_E_INVALID_ARGUMENT, "workspace not reserved");
  if (!d_img || !d_xy || !p || !n_out) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  Workspace &w = c->ws[0];
  if (H != c->H || W != c->W) return fail(OLPT_E_INVALID_ARGUMENT, "image size differs from reserve");
  CamList cl{1, {0}};
  hipStream_t st = c->st;
  fill_i32(w.flags, 16, 0, st, w.nf_cap, w.fstride);
  fill_i32(w.npts, OLPT_MAX_CAM, 0, st);
  const double mi = (double)p->min_obj_int;
  switch (dtype) {
  case OLPT_U8: launch_find_peaks<uint8_t>((const uint8_t *)d_img, 0, H, W, mi, cl, w.pk_mask, w.pk_cnt, w.pk_off, w.scan_scratch, w.totals, d_xy, cap, w.npts, w.flags + 8, st); break;
  case OLPT_U16: launch_find_peaks<uint16_t>((const uint16_t *)d_img, 0, H, W, mi, cl, w.pk_mask, w.pk_cnt, w.pk_off, w.scan_scratch, w.totals, d_xy, cap, w.npts, w.flags + 8, st); break;
  case OLPT_F32: launch_find_peaks<float>((const float *)d_img, 0, H, W, mi, cl, w.pk_mask, w.pk_cnt, w.pk_off, w.scan_scratch, w.totals, d_xy, cap, w.npts, w.flags + 8, st); break;
  case OLPT_F64: launch_find_peaks<double>((const double *)d_img, 0, H, W, mi, cl, w.pk_mask, w.pk_cnt, w.pk_off, w.scan_scratch, w.totals, d_xy, cap, w.npts, w.flags + 8, st); break;
  default: return fail(OLPT_E_UNSUPPORTED, "unsupported image dtype");
  }
  HIP_TRY(hipGetLastError());
  int32_t tot = 0;
  HIP_TRY(hipMemcpyAsync(&tot, w.totals, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *n_out = tot;
  if (tot > cap) return fail(OLPT_E_CAPACITY, "more peaks than the output capacity");
  return OLPT_OK;
}

olpt_status olpt_stereo_match(olpt_ctx *c, const int32_t *active, const double *d_xy, int cap_2d,
                              const int32_t *n_pts, const olpt_ipr_params *p, int32_t *d_cand_ids,
                              double *d_cand_e, uint8_t *d_sel, int cap_cand, double *d_obj_xyz,
                              int32_t *d_obj_ids, int cap_obj, int32_t *counts) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!d_xy || !n_pts || !p || !counts) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  hipStream_t st = c->st;
  Workspace &w = c->ws[0];
  for (int k = 0; k < c->n_cam; ++k)
    if (n_pts[k] < 0 || n_pts[k] > c->cap2d || n_pts[k] > cap_2d)
      return fail(OLPT_E_CAPACITY, "2D point count exceeds the workspace capacity");
  HIP_TRY(hipMemcpy2DAsync(w.xy, sizeof(double) * 2 * c->cap2d, d_xy, sizeof(double) * 2 * cap_2d,
                           sizeof(double) * 2 * std::min(cap_2d, c->cap2d), c->n_cam,
                           hipMemcpyDeviceToDevice, st));
  int32_t np[OLPT_MAX_CAM] = {0};
  for (int k = 0; k < c->n_cam; ++k) np[k] = (active && !active[k]) ? 0 : n_pts[k];
  HIP_TRY(hipMemcpyAsync(w.npts, np, sizeof(np), hipMemcpyHostToDevice, st));
  StereoParams sp = stereo_params(c, active, p, p->tol_2d_px, &s);
  if (s) return s;
  fill_i32(w.flags, 16, 0, st, w.nf_cap, w.fstride);
  launch_stereo(w.sb, sp, kMaxRounds, st);
  HIP_TRY(hipGetLastError());
  int32_t h[3];
  HIP_TRY(hipMemcpyAsync(&h[0], w.sb.n_cand, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&h[1], w.sb.n_sel, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&h[2], w.sb.n_obj, 4, hipMemcpyDeviceToHost, st));
  int32_t fl[8];
  HIP_TRY(hipMemcpyAsync(fl, w.flags, sizeof(fl), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  counts[0] = h[0];
  counts[1] = h[1];
  counts[2] = h[2];
  if (fl[0] || h[0] > c->cap_cand) return fail(OLPT_E_CAPACITY, "stereo candidate capacity exceeded");
  if (fl[1]) return fail(OLPT_E_CAPACITY, "stereo build list capacity exceeded");
  if (fl[4]) return fail(OLPT_E_CAPACITY, "stereo object capacity exceeded");
  if (fl[5]) return fail(OLPT_E_CAPACITY, "prune rounds exhausted");
  if (fl[2] || fl[3]) return fail(OLPT_E_PRUNE_TIES, prune_ties_msg(fl));
  const int M = h[0];
  if (d_cand_ids && d_cand_e) {
    if (M > cap_cand) return fail(OLPT_E_CAPACITY, "candidate output capacity too small");
    HIP_TRY(hipMemcpyAsync(d_cand_ids, w.sb.cand_ids_s, sizeof(int32_t) * (size_t)M * c->n_cam,
                           hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_cand_e, w.sb.cand_e_s, sizeof(double) * (size_t)M,
                           hipMemcpyDeviceToDevice, st));
  }
  if (d_sel) {
    if (M > cap_cand) return fail(OLPT_E_CAPACITY, "candidate output capacity too small");
    k_sel_flags<<<256, 256, 0, st>>>(w.sb.state, w.sb.n_cand, c->cap_cand, d_sel);
  }
  if (d_obj_xyz) {
    if (h[2] > cap_obj) return fail(OLPT_E_CAPACITY, "object output capacity too small");
    HIP_TRY(hipMemcpyAsync(d_obj_xyz, w.sb.obj_xyz, sizeof(double) * 3 * (size_t)h[2],
                           hipMemcpyDeviceToDevice, st));
    if (d_obj_ids)
      HIP_TRY(hipMemcpyAsync(d_obj_ids, w.sb.obj_ids, sizeof(int32_t) * (size_t)h[2] * c->n_cam,
                             hipMemcpyDeviceToDevice, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return OLPT_OK;
}

olpt_status olpt_otf_param(olpt_ctx *c, int cam, const double *d_xyz, int n, double *d_out) {
  if (!c || !c->n_ws) return fail(OLPT_E_INVALID_ARGUMENT, "workspace not reserved");
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (cam < 0 || cam >= c->n_cam) return fail(OLPT_E_OUT_OF_RANGE, "camera index out of range");
  if (n < 0 || (n > 0 && (!d_xyz || !d_out))) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (n == 0) return OLPT_OK;
  k_otf_param<<<std::min(4096, (n + 255) / 256), 256, 0, c->st>>>(dotf<double>(c), cam, d_xyz, n,
                                                                     d_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->st));
  return OLPT_OK;
}

olpt_status olpt_prune_match(olpt_ctx *c, const int32_t *active, const int32_t *n_pts,
                             const olpt_ipr_params *p, const int32_t *d_cand_ids,
                             const double *d_cand_e, int n_cand, uint8_t *d_sel) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!n_pts || !p || (n_cand > 0 && (!d_cand_ids || !d_cand_e || !d_sel)))
    return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (n_cand < 0) return fail(OLPT_E_INVALID_ARGUMENT, "negative candidate count");
  if (n_cand > c->cap_cand) return fail(OLPT_E_CAPACITY, "more candidates than the workspace holds");
  for (int k = 0; k < c->n_cam; ++k)
    if (n_pts[k] < 0 || n_pts[k] > c->cap2d)
      return fail(OLPT_E_CAPACITY, "2D point count exceeds the workspace capacity");
  hipStream_t st = c->st;
  Workspace &w = c->ws[0];
  StereoParams sp = stereo_params(c, active, p, p->tol_2d_px, &s);
  if (s) return s;
  const size_t nid = (size_t)n_cand * c->n_cam;
  // ids index the flat point arrays: check them on the host
  std::vector<int32_t> h_ids(nid);
  if (nid) HIP_TRY(hipMemcpy(h_ids.data(), d_cand_ids, sizeof(int32_t) * nid, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < nid; ++i) {
    const int cam = (int)(i % c->n_cam);
    if (h_ids[i] < -1 || h_ids[i] >= n_pts[cam])
      return fail(OLPT_E_OUT_OF_RANGE, "candidate id outside its camera's 2D points");
  }
  fill_i32(w.flags, 16, 0, st, w.nf_cap, w.fstride);
  if (nid) {
    HIP_TRY(hipMemcpyAsync(w.sb.cand_ids_s, d_cand_ids, sizeof(int32_t) * nid,
                           hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(w.sb.cand_e_s, d_cand_e, sizeof(double) * (size_t)n_cand,
                           hipMemcpyDeviceToDevice, st));
  }
  int32_t nc32 = n_cand;
  HIP_TRY(hipMemcpyAsync(w.sb.n_cand, &nc32, sizeof(int32_t), hipMemcpyHostToDevice, st));
  launch_prune(w.sb, sp, kMaxRounds, st);
  if (n_cand) k_sel_flags<<<256, 256, 0, st>>>(w.sb.state, w.sb.n_cand, c->cap_cand, d_sel);
  HIP_TRY(hipGetLastError());
  int32_t fl[8];
  HIP_TRY(hipMemcpyAsync(fl, w.flags, sizeof(fl), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (fl[5]) return fail(OLPT_E_CAPACITY, "prune rounds exhausted");
  if (fl[2] || fl[3]) return fail(OLPT_E_PRUNE_TIES, prune_ties_msg(fl));
  return OLPT_OK;
}

} // extern "C"

namespace {
template <typename T>
olpt_status shake_impl(olpt_ctx *c, const int32_t *active, const void *d_imgs, int dtype,
                       const olpt_ipr_params *p, const double *d_xyz_in, int n_obj,
                       double *d_xyz_out, uint8_t *d_flags, double *d_score, void *d_res) {
  hipStream_t st = c->st;
  Workspace &w = c->ws[0];
  const long n = (long)c->n_cam * c->H * c->W;
  olpt_status s = convert_image<T>(d_imgs, dtype, n, (T *)w.img_a, st);
  if (s) return s;
  HIP_TRY(hipMemcpyAsync(w.sh_xyz, d_xyz_in, sizeof(double) * 3 * (size_t)n_obj,
                         hipMemcpyDeviceToDevice, st));
  int32_t nn = n_obj;
  HIP_TRY(hipMemcpyAsync(w.sh_n, &nn, 4, hipMemcpyHostToDevice, st));
  ShakeBuffers<T> b = shake_buffers<T>(c, w, (T *)w.img_a, (T *)w.img_b);
  if constexpr (sizeof(T) == 8)
    set_grid(c, nullptr, &b, p);
  else
    set_grid(c, &b, nullptr, p);
  ShakeParams shp = shake_params(c, active, p);
  launch_shake<T>(b, shp, st);
  if (d_res) {
    launch_filter<T>(b, w.keep_xyz, (T *)w.keep_uv, w.n_keep, st);
    launch_residual_ipr<T>(b, w.n_keep, w.keep_xyz, (T *)w.keep_uv, p->radius_obj, st);
    HIP_TRY(hipMemcpyAsync(d_res, w.img_b, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  }
  HIP_TRY(hipGetLastError());
  if (d_xyz_out)
    HIP_TRY(hipMemcpyAsync(d_xyz_out, w.sh_xyz, sizeof(double) * 3 * (size_t)n_obj,
                           hipMemcpyDeviceToDevice, st));
  if (d_flags)
    HIP_TRY(hipMemcpyAsync(d_flags, w.sh_flag, (size_t)n_obj, hipMemcpyDeviceToDevice, st));
  if (d_score)
    HIP_TRY(hipMemcpyAsync(d_score, w.sh_score, sizeof(double) * (size_t)n_obj,
                           hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  return OLPT_OK;
}

template <typename T>
olpt_status residual_impl(olpt_ctx *c, const void *d_imgs, int dtype, const olpt_ipr_params *p,
                          const double *d_xyz, int n_obj, void *d_res) {
  hipStream_t st = c->st;
  Workspace &w = c->ws[0];
  const long n = (long)c->n_cam * c->H * c->W;
  olpt_status s = convert_image<T>(d_imgs, dtype, n, (T *)w.img_a, st);
  if (s) return s;
  HIP_TRY(hipMemcpyAsync(w.keep_xyz, d_xyz, sizeof(double) * 3 * (size_t)n_obj,
                         hipMemcpyDeviceToDevice, st));
  int32_t nn = n_obj;
  HIP_TRY(hipMemcpyAsync(w.n_keep, &nn, 4, hipMemcpyHostToDevice, st));
  ShakeBuffers<T> b = shake_buffers<T>(c, w, (T *)w.img_a, (T *)w.img_b);
  launch_residual_ipr<T>(b, w.n_keep, w.keep_xyz, (T *)w.keep_uv, p->radius_obj, st);
  HIP_TRY(hipMemcpyAsync(d_res, w.img_b, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  return OLPT_OK;
}

// Read and merge the 8 diagnostic flags of every frame slice of the first
// nw workspaces (flags[3] is summed, the others or-ed).
olpt_status read_flags(olpt_ctx *c, int nw, hipStream_t st, int32_t out[8]) {
  for (int i = 0; i < 8; ++i) out[i] = 0;
  std::vector<int32_t> fw;
  for (int k = 0; k < nw; ++k) {
    const Workspace &w = c->ws[k];
    if (!w.arena) continue;
    fw.assign(8 * (size_t)w.nf_cap, 0);
    HIP_TRY(hipMemcpy2DAsync(fw.data(), 8 * sizeof(int32_t), w.flags, (size_t)w.fstride,
                             8 * sizeof(int32_t), (size_t)w.nf_cap, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int f = 0; f < w.nf_cap; ++f)
      for (int i = 0; i < 8; ++i)
        out[i] = i == 3 ? out[i] + fw[8 * f + i] : (out[i] | fw[8 * f + i]);
  }
  return OLPT_OK;
}

// Frames of one call: chunks of up to c->batch consecutive frames, chunk j
// on workspace j % nw (each workspace's chunks in order on its stream).
template <typename T>
olpt_status run_chunk(olpt_ctx *c, Workspace &w, int f0, int nb, const void *d_imgs, int dtype,
                      const olpt_ipr_params *p, double *d_obj_xyz, double *d_obj_uv, int cap,
                      int32_t *d_n_obj, int32_t *d_loop_tot, int cap_loop, int *n_loop) {
  const long n = (long)c->n_cam * c->H * c->W;
  const size_t in_bytes = dtype == OLPT_U8 ? 1 : dtype == OLPT_U16 ? 2 : dtype == OLPT_F32 ? 4 : 8;
  const char *src = (const char *)d_imgs + (size_t)f0 * n * in_bytes;
  olpt_status s = convert_image<T>(src, dtype, n, (T *)w.img_a, w.st, nb, (long)(n * in_bytes),
                                   w.fstride);
  if (s) return s;
  return ipr_frame<T>(c, w, nb, p, d_obj_xyz + (size_t)f0 * cap * 3,
                      d_obj_uv ? d_obj_uv + (size_t)f0 * cap * c->n_cam * 2 : nullptr, cap,
                      d_n_obj + f0, d_loop_tot ? d_loop_tot + (size_t)f0 * cap_loop : nullptr,
                      cap_loop, n_loop);
}

template <typename T>
olpt_status ipr_run_impl(olpt_ctx *c, const void *d_imgs, int dtype, int n_frames,
                         const olpt_ipr_params *p, double *d_obj_xyz, double *d_obj_uv, int cap,
                         int32_t *d_n_obj, int32_t *loop_total, int cap_loop, bool sync) {
  hipStream_t st = c->st;
  TimerScope ts(c, !sync);
  if (loop_total) {
    const int need = std::max(1, n_frames * cap_loop);
    if (c->loop_tot_cap < need) {
      olpt_status s = dalloc_t(c, &c->loop_tot, need);
      if (s) return s;
      c->loop_tot_cap = need;
    }
  }
  const int B = std::max(1, c->batch);
  const int n_chunks = (n_frames + B - 1) / B;
  // chunks round-robin over the workspaces; the other streams start after
  // the work already queued on the context stream and join back at the end
  const int nw = std::max(1, std::min({c->n_ws_active, c->n_ws, n_chunks}));
  // loop totals past a frame's last loop read 0 (before the fork: every
  // stream writes into them)
  if (loop_total) fill_i32(c->loop_tot, (long)n_frames * cap_loop, 0, st);
  if (nw > 1) {
    HIP_TRY(hipEventRecord(c->ev_fork, st));
    for (int k = 1; k < nw; ++k) HIP_TRY(hipStreamWaitEvent(c->ws[k].st, c->ev_fork, 0));
  }
  for (int k = 0; k < nw; ++k)
    fill_i32(c->ws[k].flags, 16, 0, c->ws[k].st, c->ws[k].nf_cap, c->ws[k].fstride);
  int n_loop = 0;
  const auto t_issue = std::chrono::steady_clock::now();
  for (int j = 0; j < n_chunks; ++j) {
    const int f0 = j * B, nb = std::min(B, n_frames - f0);
    olpt_status s = run_chunk<T>(c, c->ws[j % nw], f0, nb, d_imgs, dtype, p, d_obj_xyz, d_obj_uv,
                                 cap, d_n_obj, loop_total ? c->loop_tot : nullptr, cap_loop,
                                 &n_loop);
    if (s) return s;
  }
  if (nw > 1) {
    for (int k = 1; k < nw; ++k) {
      HIP_TRY(hipEventRecord(c->ws[k].done, c->ws[k].st));
      HIP_TRY(hipStreamWaitEvent(st, c->ws[k].done, 0));
    }
  }
  if (!sync) return OLPT_OK;
  c->host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_issue).count();
  c->host_calls += 1;
  if (loop_total) {
    HIP_TRY(hipMemcpyAsync(loop_total, c->loop_tot, sizeof(int32_t) * (size_t)n_frames * cap_loop,
                           hipMemcpyDeviceToHost, st));
  }
  int32_t fl[8];
  olpt_status s = read_flags(c, nw, st, fl);
  if (s) return s;
  if (fl[7]) return fail(OLPT_E_CAPACITY, "2D peak capacity exceeded");
  if (fl[0]) return fail(OLPT_E_CAPACITY, "stereo candidate capacity exceeded");
  if (fl[1]) return fail(OLPT_E_CAPACITY, "stereo build list capacity exceeded");
  if (fl[4]) return fail(OLPT_E_CAPACITY, "object capacity exceeded");
  if (fl[5]) return fail(OLPT_E_CAPACITY, "prune rounds exhausted");
  if (fl[2] || fl[3]) return fail(OLPT_E_PRUNE_TIES, prune_ties_msg(fl));
  (void)n_loop;
  return OLPT_OK;
}

// The chunks of workspace k (j = k, k + nw, ...) on its stream, for capture:
// the same kernel sequence olpt_ipr_run issues for them, without host reads.
template <typename T>
olpt_status capture_slot(olpt_ctx *c, int k, int nw, const void *d_imgs, int dtype, int n_frames,
                         const olpt_ipr_params *p, double *d_obj_xyz, double *d_obj_uv, int cap,
                         int32_t *d_n_obj) {
  TimerScope ts(c, true);
  Workspace &w = c->ws[k];
  fill_i32(w.flags, 16, 0, w.st, w.nf_cap, w.fstride);
  const int B = std::max(1, c->batch);
  const int n_chunks = (n_frames + B - 1) / B;
  int n_loop = 0;
  for (int j = k; j < n_chunks; j += nw) {
    const int f0 = j * B, nb = std::min(B, n_frames - f0);
    olpt_status s = run_chunk<T>(c, w, f0, nb, d_imgs, dtype, p, d_obj_xyz, d_obj_uv, cap, d_n_obj,
                                 nullptr, 0, &n_loop);
    if (s) return s;
  }
  return OLPT_OK;
}
} // namespace

extern "C" {

olpt_status olpt_shake(olpt_ctx *c, const int32_t *active, const void *d_imgs, int dtype,
                       const olpt_ipr_params *p, const double *d_xyz_in, int n_obj,
                       double *d_xyz_out, uint8_t *d_flags, double *d_score, void *d_res) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!p || !d_imgs || (n_obj > 0 && !d_xyz_in)) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if ((s = check_params(c, p))) return s;
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (n_obj < 0 || n_obj > c->cap_obj) return fail(OLPT_E_CAPACITY, "object count exceeds capacity");
  if (c->prec == OLPT_PREC_F64)
    return shake_impl<double>(c, active, d_imgs, dtype, p, d_xyz_in, n_obj, d_xyz_out, d_flags, d_score, d_res);
  return shake_impl<float>(c, active, d_imgs, dtype, p, d_xyz_in, n_obj, d_xyz_out, d_flags, d_score, d_res);
}

olpt_status olpt_residual(olpt_ctx *c, const void *d_imgs, int dtype, const olpt_ipr_params *p,
                          const double *d_xyz, int n_obj, void *d_res) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!p || !d_imgs || !d_res || (n_obj > 0 && !d_xyz)) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if ((s = check_params(c, p))) return s;
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (n_obj < 0 || n_obj > c->cap_obj) return fail(OLPT_E_CAPACITY, "object count exceeds capacity");
  if (c->prec == OLPT_PREC_F64) return residual_impl<double>(c, d_imgs, dtype, p, d_xyz, n_obj, d_res);
  return residual_impl<float>(c, d_imgs, dtype, p, d_xyz, n_obj, d_res);
}

olpt_status olpt_ipr_run(olpt_ctx *c, const void *d_imgs, int dtype, int n_frames,
                         const olpt_ipr_params *p, double *d_obj_xyz, double *d_obj_uv, int cap,
                         int32_t *d_n_obj, int32_t *loop_total, int cap_loop) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!p || !d_imgs || !d_obj_xyz || !d_n_obj || n_frames < 1)
    return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if ((s = check_params(c, p))) return s;
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (c->prec == OLPT_PREC_F64)
    return ipr_run_impl<double>(c, d_imgs, dtype, n_frames, p, d_obj_xyz, d_obj_uv, cap, d_n_obj,
                                loop_total, cap_loop, true);
  return ipr_run_impl<float>(c, d_imgs, dtype, n_frames, p, d_obj_xyz, d_obj_uv, cap, d_n_obj,
                             loop_total, cap_loop, true);
}

olpt_status olpt_ipr_graph_capture(olpt_ctx *c, const void *d_imgs, int dtype, int n_frames,
                                   const olpt_ipr_params *p, double *d_obj_xyz, double *d_obj_uv,
                                   int cap, int32_t *d_n_obj) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!p || !d_imgs || !d_obj_xyz || !d_n_obj || n_frames < 1)
    return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if ((s = check_params(c, p))) return s;
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (!c->st) return fail(OLPT_E_INVALID_ARGUMENT, "graph capture needs a non-null stream");
  if (c->stats) return fail(OLPT_E_INVALID_ARGUMENT, "byte accounting cannot be captured");
  free_graphs(c);
  // One graph per workspace: the frames f = k, k + nw, ... that
  // olpt_ipr_run would send to workspace k, captured on its stream.  The
  // graphs are launched on their own streams, so frames overlap as in the
  // eager run while the host submits one graph instead of ~1000 kernels
  // per frame.
  const int n_chunks = (n_frames + std::max(1, c->batch) - 1) / std::max(1, c->batch);
  const int nw = std::max(1, std::min({c->n_ws_active, c->n_ws, n_chunks}));
  for (int k = 0; k < nw; ++k) HIP_TRY(hipStreamSynchronize(c->ws[k].st));
  c->capturing = true;
  for (int k = 0; k < nw; ++k) {
    hipStream_t st = c->ws[k].st;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    if (c->prec == OLPT_PREC_F64)
      s = capture_slot<double>(c, k, nw, d_imgs, dtype, n_frames, p, d_obj_xyz, d_obj_uv, cap, d_n_obj);
    else
      s = capture_slot<float>(c, k, nw, d_imgs, dtype, n_frames, p, d_obj_xyz, d_obj_uv, cap, d_n_obj);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (s || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      c->capturing = false;
      free_graphs(c);
      if (s) return s;
      return fail(OLPT_E_GPU, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    c->graph[k] = g;
    c->n_graph = k + 1;
    hipError_t ei = hipGraphInstantiate(&c->gexec[k], g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      c->capturing = false;
      free_graphs(c);
      return fail(OLPT_E_GPU, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
  }
  c->capturing = false;
  return OLPT_OK;
}

olpt_status olpt_stage_timing(olpt_ctx *c, int enable) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (enable && !c->timer) {
    c->timer = new StageTimer();
    for (int a = 0; a < 2; ++a)
      for (int s = 0; s < 8; ++s)
        for (int i = 0; i < kStageEvents; ++i) HIP_TRY(hipEventCreate(&c->timer->ev[a][s][i]));
  }
  if (c->timer) {
    c->timer->on = enable != 0;
    for (int s = 0; s < 8; ++s) c->timer->n[s] = 0;
  }
  return OLPT_OK;
}

olpt_status olpt_stage_times(olpt_ctx *c, double ms[8], int32_t calls[8]) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  HIP_TRY(hipStreamSynchronize(c->st));
  for (int s = 0; s < 8; ++s) {
    ms[s] = 0.0;
    calls[s] = 0;
    if (!c->timer) continue;
    for (int i = 0; i < c->timer->n[s]; ++i) {
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, c->timer->ev[0][s][i], c->timer->ev[1][s][i]));
      ms[s] += t;
    }
    calls[s] = c->timer->n[s];
    c->timer->n[s] = 0;
  }
  ms[OLPT_STAGE_HOST] = c->host_ms;
  calls[OLPT_STAGE_HOST] = c->host_calls;
  c->host_ms = 0.0;
  c->host_calls = 0;
  return OLPT_OK;
}

olpt_status olpt_stage_stats(olpt_ctx *c, int enable, double bytes[8], int32_t launches[8]) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (bytes && launches)
    for (int s = 0; s < 8; ++s) {
      bytes[s] = c->st_bytes[s];
      launches[s] = c->st_launch[s];
    }
  for (int s = 0; s < 8; ++s) {
    c->st_bytes[s] = 0;
    c->st_launch[s] = 0;
  }
  c->stats = enable != 0;
  return OLPT_OK;
}

olpt_status olpt_ipr_graph_launch(olpt_ctx *c) {
  if (!c || !c->n_graph) return fail(OLPT_E_INVALID_ARGUMENT, "no captured graph");
  const int ng = c->n_graph;
  if (ng > 1) {
    HIP_TRY(hipEventRecord(c->ev_fork, c->st));
    for (int k = 1; k < ng; ++k) HIP_TRY(hipStreamWaitEvent(c->ws[k].st, c->ev_fork, 0));
  }
  for (int k = 0; k < ng; ++k) HIP_TRY(hipGraphLaunch(c->gexec[k], c->ws[k].st));
  for (int k = 1; k < ng; ++k) {
    HIP_TRY(hipEventRecord(c->ws[k].done, c->ws[k].st));
    HIP_TRY(hipStreamWaitEvent(c->st, c->ws[k].done, 0));
  }
  return OLPT_OK;
}

} // extern "C"

// ===========================================================================
// learned shake proposal, rendering and network inputs
// ===========================================================================
namespace {
template <typename T>
olpt_status render_impl(olpt_ctx *c, const double *d_xyz, const double *d_amp, int n, double r_px,
                        void *d_out) {
  hipStream_t st = c->st;
  Workspace &w = c->ws[0];
  if (n > 0)
    HIP_TRY(hipMemcpyAsync(w.keep_xyz, d_xyz, sizeof(double) * 3 * (size_t)n,
                           hipMemcpyDeviceToDevice, st));
  fill_i32(w.n_keep, 1, n, st);
  ShakeBuffers<T> b = shake_buffers<T>(c, w, (T *)w.img_a, (T *)w.img_b);
  launch_render<T>(b, w.keep_xyz, d_amp, w.n_keep, r_px, (T *)d_out, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return OLPT_OK;
}

template <typename T>
olpt_status patches_impl(olpt_ctx *c, const int32_t *active, const void *d_imgs, int dtype,
                         const olpt_ipr_params *p, const double *d_xyz, int n, void *d_x) {
  hipStream_t st = c->st;
  Workspace &w = c->ws[0];
  const long npx = (long)c->n_cam * c->H * c->W;
  olpt_status s = convert_image<T>(d_imgs, dtype, npx, (T *)w.img_a, st);
  if (s) return s;
  if (n > 0)
    HIP_TRY(hipMemcpyAsync(w.sh_xyz, d_xyz, sizeof(double) * 3 * (size_t)n,
                           hipMemcpyDeviceToDevice, st));
  fill_i32(w.sh_n, 1, n, st);
  ShakeBuffers<T> b = shake_buffers<T>(c, w, (T *)w.img_a, (T *)w.img_b);
  ShakeParams shp = shake_params(c, active, p);
  launch_patches<T>(b, shp, d_x, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return OLPT_OK;
}
} // namespace

extern "C" {

olpt_status olpt_set_shake_model(olpt_ctx *c, const olpt_mlp *mlp, double voxel_mm,
                                 double max_step_vox) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (mlp && c->n_cam != 4)
    return fail(OLPT_E_UNSUPPORTED, "the learned shake proposal takes exactly 4 cameras");
  if (mlp && !(voxel_mm > 0)) return fail(OLPT_E_INVALID_ARGUMENT, "voxel_mm must be positive");
  c->mlp = mlp;
  c->mlp_voxel_mm = voxel_mm;
  c->mlp_max_step_vox = max_step_vox > 0 ? max_step_vox : 0.5;
  if (mlp)
    for (int k = 0; k < c->n_ws; ++k) {
      olpt_status s = alloc_learned(c, c->ws[k]);
      if (s) return s;
    }
  return OLPT_OK;
}

olpt_status olpt_set_ghost_logit(olpt_ctx *c, double logit_min) {
  if (!c) return fail(OLPT_E_INVALID_ARGUMENT, "null context");
  if (logit_min != logit_min) return fail(OLPT_E_INVALID_ARGUMENT, "logit_min is NaN");
  c->mlp_ghost_logit = logit_min;
  return OLPT_OK;
}

olpt_status olpt_render(olpt_ctx *c, const double *d_xyz, const double *d_amp, int n_obj,
                        double r_px, void *d_out) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!d_out || (n_obj > 0 && !d_xyz)) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (n_obj < 0 || n_obj > c->cap_obj) return fail(OLPT_E_CAPACITY, "object count exceeds capacity");
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (!(r_px > 0) || std::lround(2.0 * r_px) > 24)
    return fail(OLPT_E_INVALID_ARGUMENT, "r_px must be in (0, 12]");
  if (c->prec == OLPT_PREC_F64) return render_impl<double>(c, d_xyz, d_amp, n_obj, r_px, d_out);
  return render_impl<float>(c, d_xyz, d_amp, n_obj, r_px, d_out);
}

olpt_status olpt_shake_patches(olpt_ctx *c, const int32_t *active, const void *d_imgs, int dtype,
                               const olpt_ipr_params *p, const double *d_xyz, int n_obj,
                               void *d_x) {
  olpt_status s = check_ready(c);
  if (s) return s;
  if (!p || !d_imgs || !d_x || (n_obj > 0 && !d_xyz)) return fail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if ((s = check_params(c, p))) return s;
  if (!c->otf_set) return fail(OLPT_E_INVALID_ARGUMENT, "OTF not set");
  if (c->n_cam != 4) return fail(OLPT_E_UNSUPPORTED, "the learned shake proposal takes exactly 4 cameras");
  if (n_obj < 0 || n_obj > c->cap_obj) return fail(OLPT_E_CAPACITY, "object count exceeds capacity");
  if (c->prec == OLPT_PREC_F64) return patches_impl<double>(c, active, d_imgs, dtype, p, d_xyz, n_obj, d_x);
  return patches_impl<float>(c, active, d_imgs, dtype, p, d_xyz, n_obj, d_x);
}

} // extern "C"
