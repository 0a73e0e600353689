// olpt_shake.hip - Shake (tracer strategy) and residual images (Shake.cpp).
//
// Per shake iteration (Shake::runShake, Shake.cpp:36-163; Jacobi: every
// object sees the residual built from the start-of-iteration positions):
//   k_obj_project    2D projection of every object in every camera plus its
//                    OTF parameters at the 3D centre (getOTFParam), counted
//                    into the bin of its radius box (calROIBound with r_px),
//                    keyed by the box's top-left 8x8-pixel cell
//   scan, k_bin_fill CSR bins over all cameras (the scan leaves the counts
//                    zero for the next pass)
//   k_resid_gather   32x32 pixel tile per workgroup: the objects whose boxes
//                    can touch the tile are staged in LDS and scatter their
//                    box rows into an LDS image by integer max; every pixel
//                    then writes res = orig - max(0, max_k p_k)  (the
//                    reference's min fusion, Shake.cpp:226-241) - no
//                    atomics on HBM
//   k_shake          one wave64 per object, one lane group per camera: the
//                    group owns the camera's projection, OTF parameters and
//                    the pixels of its augmented ROI (buildROIInfo,
//                    :279-328), so one residual evaluation (calShakeResidue,
//                    :968-1045) covers every camera in a single pass; group
//                    butterfly sums and an in-order cross-camera mean leave
//                    the same value in every lane, so the branchy per-axis
//                    search of shakeOneObject (:409-665) stays wave-uniform;
//                    then calObjectScore (:668-782).
// After the iterations: markRepeatedObj (:785-823) on a uniform grid with
// the greedy pass over the (rare) objects that have a neighbour, the
// mean-relative ghost threshold (:118-136), order-preserving filter
// (IPR.cpp:41-63) and the IPR-mode residual for the next loop.
#include "olpt_kernels.hpp"
#include "olpt_scan.hpp"

namespace olpt {

// objects staged per tile batch: a 32 x 32 tile of a ppp 0.05 image meets
// about 65, so most tiles take one or two batches in little LDS
constexpr int kTileObjCap = 64;

static int grid_for(long n, int threads, int max_blocks) {
  long g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

// per (object, camera): u, v, box, OTF params
template <typename T> struct ObjCamRec {
  T u, v;
  T P[5];
  int32_t rmin, rmax, cmin, cmax;
};

// BIN: also count the object's box in its anchor bin (the box's top-left
// 8x8-pixel cell); bin_cnt is all zero on entry (the scan that turns the
// counts into offsets clears it again).
template <typename T, bool BIN>
__global__ void k_obj_project(ShakeBuffers<T> bb, const double *__restrict__ xyz,
                              const int32_t *__restrict__ n_ptr, ObjCamRec<T> *rec,
                              T *uv_out, int box_w) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  xyz = fshift(xyz, fo);
  n_ptr = fshift(n_ptr, fo);
  rec = fshift(rec, fo);
  uv_out = fshift(uv_out, fo);
  const int n = *n_ptr;
  const int nc = b.n_cam;
  const long nbin = (long)b.bin_rows * b.bin_cols;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T X[3] = {(T)xyz[3L * i], (T)xyz[3L * i + 1], (T)xyz[3L * i + 2]};
    for (int k = 0; k < nc; ++k) {
      ObjCamRec<T> r;
      project(b.cams[k], X, r.u, r.v);
      otf_param(b.otf, k, X, r.P);
      const PRange pr = roi_bound<T>(b.H, b.W, r.u, r.v, box_w, box_w);
      r.rmin = pr.rmin;
      r.rmax = pr.rmax;
      r.cmin = pr.cmin;
      r.cmax = pr.cmax;
      rec[(long)i * nc + k] = r;
      if (uv_out) {
        uv_out[((long)i * nc + k) * 2] = r.u;
        uv_out[((long)i * nc + k) * 2 + 1] = r.v;
      }
      if constexpr (BIN) {
        int bin = -1, rank = 0;
        if (!pr.empty()) {
          bin = (pr.rmin / b.bin_px) * b.bin_cols + (pr.cmin / b.bin_px);
          rank = atomicAdd(&b.bin_cnt[k * nbin + bin], 1);
        }
        b.obj_bin[(long)k * b.cap_obj + i] = bin;
        b.obj_rank[(long)k * b.cap_obj + i] = rank;
      }
    }
  }
}

// Copies every binned record to its place in bin order, so that a tile
// reads its objects' records contiguously.
template <typename T>
__global__ void k_bin_fill(ShakeBuffers<T> bb, const int32_t *__restrict__ n_ptr,
                           const ObjCamRec<T> *__restrict__ rec) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  n_ptr = fshift(n_ptr, fo);
  rec = fshift(rec, fo);
  ObjCamRec<T> *brec = reinterpret_cast<ObjCamRec<T> *>(b.bin_rec);
  const int n = *n_ptr;
  const int nc = b.n_cam;
  const long nbin = (long)b.bin_rows * b.bin_cols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)n * nc;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t / nc), k = (int)(t % nc);
    const int bin = b.obj_bin[(long)k * b.cap_obj + i];
    if (bin < 0) continue;
    brec[b.bin_start[k * nbin + bin] + b.obj_rank[(long)k * b.cap_obj + i]] = rec[t];
  }
}

// mode 0: shake residual (active cameras, no clamp); mode 1: IPR output
// (all cameras, clamp at 0); mode 2: rendered image max(0, max_k p_k).  The
// shake never reads an inactive camera's residual, so mode 0 skips them.
//
// calResidueImage (Shake.cpp:165-245) subtracts from each pixel the largest
// projection of the objects whose boxes cover it.  One workgroup owns a
// kResTile x kResTile tile of one camera: it stages the objects whose box
// anchors lie in the bins that can reach the tile, then each (object, box
// row) work item writes its row's projection values inside the tile into an
// LDS image with an integer max.  The values are >= 0 and the IEEE bit
// patterns of non-negative numbers order as unsigned integers, so the max is
// exact and independent of order: every pixel ends with the same value a
// per-pixel loop over the objects gives, at box-area rather than
// tile-objects work per pixel.
constexpr int kTileBinCap = 64;    // bins a tile can reach: 7 x 7 for boxes <= 24 px
constexpr int kResTile = 32;       // residual tile edge (pixels)
// Two waves per tile: a tile's work is a short chain of dependent loads
// (bin starts, then records), so many small resident workgroups hide it
// better than fewer large ones (-17% against 256 threads with 256 records)
constexpr int kResThreads = 128;
static_assert(kTileBinCap == kWave, "the bin prefix is one wave's scan");
static_assert(kTileObjCap <= kResThreads, "one thread stages each record of a batch");
static_assert((kResTile * kResTile) % kResThreads == 0, "whole pixels per thread");

template <typename T> struct MaxBits;
template <> struct MaxBits<float> {
  using U = uint32_t;
  static __device__ U enc(float v) { return __float_as_uint(v); }
  static __device__ float dec(U u) { return __uint_as_float(u); }
};
template <> struct MaxBits<double> {
  using U = unsigned long long;
  static __device__ U enc(double v) { return (U)__double_as_longlong(v); }
  static __device__ double dec(U u) { return __longlong_as_double((long long)u); }
};

template <typename T>
__global__ void __launch_bounds__(kResThreads)
k_resid_gather(ShakeBuffers<T> bb, const T *__restrict__ orig, T *__restrict__ out,
               const ObjCamRec<T> *__restrict__ rec, CamList cams, int mode, int box_w) {
  const long F = (long)(blockIdx.z / cams.n);
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  orig = fshift(orig, fo);
  out = fshift(out, fo);
  rec = fshift(rec, fo);
  using U = typename MaxBits<T>::U;
  constexpr int NPX = kResTile * kResTile;
  __shared__ U s_img[NPX];
  __shared__ ObjCamRec<T> s_obj[kTileObjCap];
  __shared__ int32_t s_bstart[kTileBinCap], s_bpre[kTileBinCap + 1];
  const int k = cams.id[blockIdx.z % cams.n];
  const int r0 = blockIdx.y * kResTile, c0 = blockIdx.x * kResTile;
  const int tid = threadIdx.x;
  const long HW = (long)b.H * b.W;
  const long nbin = (long)b.bin_rows * b.bin_cols;
  const int32_t *bstart = b.bin_start + k * nbin;  // offsets into bin_rec, all cameras
  const ObjCamRec<T> *brec = reinterpret_cast<const ObjCamRec<T> *>(b.bin_rec);
  for (int q = tid; q < NPX; q += kResThreads) s_img[q] = 0;
  // the tile's original pixels, loaded now and used at the end
  constexpr int PPT = NPX / kResThreads;
  T o_px[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int q = tid + u * kResThreads;
    const int row = r0 + q / kResTile, col = c0 + q % kResTile;
    o_px[u] = row < b.H && col < b.W ? orig[k * HW + (long)row * b.W + col] : T(0);
  }
  // anchors (box top-left) that can cover this tile, and their bins
  const int ar0 = max(0, r0 - box_w + 1), ar1 = min(b.H - 1, r0 + kResTile - 1);
  const int ac0 = max(0, c0 - box_w + 1), ac1 = min(b.W - 1, c0 + kResTile - 1);
  const int br0 = ar0 / b.bin_px, br1 = ar1 / b.bin_px;
  const int bc0 = ac0 / b.bin_px, bc1 = ac1 / b.bin_px;
  const int nbc = bc1 - bc0 + 1;
  const int nb = (br1 - br0 + 1) * nbc;  // <= kTileBinCap (box_w <= 24 checked on the host)
  if (tid < kWave) {
    int cnt = 0;
    if (tid < nb) {
      const int bin = (br0 + tid / nbc) * b.bin_cols + bc0 + tid % nbc;
      s_bstart[tid] = bstart[bin];
      cnt = bstart[bin + 1] - bstart[bin];
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int o = __shfl_up(cnt, off, kWave);
      if (tid >= off) cnt += o;
    }
    s_bpre[tid + 1] = cnt;
    if (tid == 0) s_bpre[0] = 0;
  }
  __syncthreads();
  const int total = s_bpre[nb];
  const float inv_bw = 1.0f / (float)box_w;  // items < 256 * 24: the quotient is exact
  for (int base = 0; base < total; base += kTileObjCap) {
    const int n = min(kTileObjCap, total - base);
    if (tid < n) {
      const int t = base + tid;
      int lo = 0, hi = nb - 1;  // the last bin q with s_bpre[q] <= t
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_bpre[mid] <= t)
          lo = mid;
        else
          hi = mid - 1;
      }
      s_obj[tid] = brec[s_bstart[lo] + (t - s_bpre[lo])];
    }
    __syncthreads();
    for (int it = tid; it < n * box_w; it += kResThreads) {
      const int u = (int)(((float)it + 0.5f) * inv_bw);
      const ObjCamRec<T> &r = s_obj[u];
      const int row = r.rmin + (it - u * box_w);
      if (row < r.rmax && row >= r0 && row < r0 + kResTile) {
        const int cb = max(r.cmin, c0), ce = min(r.cmax, c0 + kResTile);
        U *dst = s_img + (row - r0) * kResTile - c0;
        bool sep = false;
        if constexpr (sizeof(T) == 4) sep = r.P[3] == T(1) && r.P[4] == T(0);
        if (sep) {
          // FP32 without rotation: the row factor once, then one
          // exponential per column (rounds like k_shake's product form)
          const T dy = T(row) - r.v;
          const T gy = r.P[0] * gexp(-(r.P[2] * dy * dy));
          for (int col = cb; col < ce; ++col) {
            const T dx = T(col) - r.u;
            const T pv = gy * gexp(-(r.P[1] * dx * dx));
            if (pv > T(0)) atomicMax(dst + col, MaxBits<T>::enc(pv));
          }
        } else {
          for (int col = cb; col < ce; ++col) {
            const T pv = gauss_int<T>(r.P, r.u, r.v, row, col);
            if (pv > T(0)) atomicMax(dst + col, MaxBits<T>::enc(pv));
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int q = tid + u * kResThreads;
    const int row = r0 + q / kResTile, col = c0 + q % kResTile;
    if (row < b.H && col < b.W) {
      const long pix = (long)row * b.W + col;
      const T maxp = MaxBits<T>::dec(s_img[q]);
      T v = o_px[u] - maxp;
      if (mode == 1 && v < T(0)) v = T(0);
      if (mode == 2) v = maxp;  // rendering: max-fusion of the projections
      out[k * HW + pix] = v;
    }
  }
}

// per-object amplitude in place of the OTF's a (rendering)
template <typename T>
__global__ void k_set_amp(ObjCamRec<T> *rec, const double *__restrict__ amp,
                          const int32_t *__restrict__ n_ptr, int nc) {
  const long n = (long)*n_ptr * nc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    rec[t].P[0] = (T)amp[t / nc];
}

template <typename T>
static void residual_pass(const ShakeBuffers<T> &b, const T *orig, T *out,
                          const double *xyz, const int32_t *n_ptr,
                          ObjCamRec<T> *rec, T *uv_out, CamList cams, int mode,
                          double r_px, hipStream_t st, const double *amp = nullptr) {
  const int nc = b.n_cam;
  const long nbin = (long)b.bin_rows * b.bin_cols;
  const int box_w = box_width(r_px);
  k_obj_project<T, true><<<dim3(grid_for(b.cap_obj, 256, 2048), 1, b.nf), 256, 0, st>>>(b, xyz, n_ptr, rec,
                                                                         uv_out, box_w);
  if (amp) k_set_amp<T><<<grid_for((long)b.cap_obj * nc, 256, 2048), 256, 0, st>>>(rec, amp, n_ptr, nc);
  // one scan over all cameras' bins; bin_start[nc * nbin] receives the total
  scan_exclusive_clear(b.bin_cnt, b.bin_start, nbin * nc, b.bin_start + nbin * nc,
                       b.scan_scratch, st, b.nf, b.fstride);
  k_bin_fill<T><<<dim3(grid_for((long)b.cap_obj * nc, 256, 4096), 1, b.nf), 256, 0, st>>>(b, n_ptr,
                                                                                        rec);
  dim3 g((b.W + kResTile - 1) / kResTile, (b.H + kResTile - 1) / kResTile, cams.n * b.nf);
  stage_begin(ST_RESID, st);
  k_resid_gather<T><<<g, kResThreads, 0, st>>>(b, orig, out, rec, cams, mode, box_w);
  stage_end(ST_RESID, st);
}

template <typename T>
__global__ void k_copy_cam(const T *__restrict__ src, T *__restrict__ dst, long HW,
                           CamList cams, long fstride) {
  src = fshift(src, (long)blockIdx.z * fstride);
  dst = fshift(dst, (long)blockIdx.z * fstride);
  const int k = cams.id[blockIdx.y];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < HW;
       i += (long)gridDim.x * blockDim.x)
    dst[k * HW + i] = src[k * HW + i];
}

// Workgroups are handed to the 8 XCDs round-robin in launch order, each XCD
// with its own L2.  With a grid width that is a multiple of 8, workgroup x
// of a frame runs on XCD x % 8.  xcd_chunk numbers the chunks of grid-stride
// pass `pass` (rem chunks left from its start, each of OPW objects) XCD by
// XCD, so that each XCD takes a contiguous run of the (spatially ordered)
// objects, whose ROIs then share that XCD's L2 lines; a last, partial pass
// is spread evenly over the 8 XCDs.  Every chunk is taken once; -1: none.
__device__ __forceinline__ int xcd_chunk(int pass, int rem) {
  const int G = gridDim.x, x = blockIdx.x;
  if (G % 8 != 0) return x < rem ? pass * G + x : -1;
  const int per = rem >= G ? G / 8 : (rem + 7) / 8;
  const int k = (x % 8) * per + x / 8;
  return x / 8 < per && k < rem ? pass * G + k : -1;
}

// --------------------------------------------------------------------------
// k_shake: wave per object, lanes split by camera
// --------------------------------------------------------------------------
// A wave holds OPW objects; each object's lanes split into NC groups and
// group k owns camera k: its projection, OTF parameters and SLOTS pixels of
// the camera's augmented ROI.  A residual evaluation is one pass over all
// cameras at once; per-camera sums are butterfly reductions inside the
// group (bitwise identical in its lanes), and the cross-camera mean is
// summed in camera order from the group leaders, so every lane of an object
// holds the same value and the branchy per-axis search of shakeOneObject
// stays uniform per object.  Several objects per wave share the per-camera
// projection and OTF work of a lane over more pixels.
// Lane xor 1 or 2 inside each quad of lanes as a DPP quad permutation (a
// register move, where __shfl_xor goes through the LDS crossbar).
template <int X> __device__ __forceinline__ int quad_xor(int v) {
  static_assert(X == 1 || X == 2, "quad permutations only");
  constexpr int ctrl = X == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
  return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xF, 0xF, false);
}
template <int X> __device__ __forceinline__ float quad_xor(float v) {
  return __int_as_float(quad_xor<X>(__float_as_int(v)));
}
template <int X> __device__ __forceinline__ double quad_xor(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = quad_xor<X>((int)(u & 0xffffffffll)), hi = quad_xor<X>((int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the LPG lanes of a group (every lane of the group active); the
// additions are v += v[lane ^ off] for off = LPG/2 .. 1 in that order.
template <int LPG, typename T> __device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = LPG / 2; off > 2; off >>= 1) v += __shfl_xor(v, off, kWave);
  if constexpr (LPG >= 4) v += quad_xor<2>(v);
  if constexpr (LPG >= 2) v += quad_xor<1>(v);
  return v;
}
template <int LPG, typename T> __device__ __forceinline__ T group_max(T v) {
#pragma unroll
  for (int off = LPG / 2; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ bool in_roi_d(const PRange &r, T u, T v) {
  return T(r.cmin) <= u && u < T(r.cmax) && T(r.rmin) <= v && v < T(r.rmax);
}

// median of up to NC values (uniform across lanes)
template <typename T, int NC>
__device__ T median_small(T *v, int n) {
  if (n == 0) return T(0);
  for (int i = 1; i < n; ++i) {
    const T x = v[i];
    int j = i - 1;
    while (j >= 0 && v[j] > x) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
  const int mid = n / 2;
  T m = v[mid];
  if ((n & 1) == 0) m = T(0.5) * (m + v[mid - 1]);
  return m;
}

// myMATH::polyfit order 2 (myMATH.cpp:490-534): a = (inv(X'X) X') y, always
// in FP64 - for offsets of ~1e-2 mm the normal matrix has a condition number
// near 1e8, beyond FP32.
__device__ void polyfit2(const double *x, const double *y, int n, double c[3]) {
  double XtX[9];
  for (int a = 0; a < 3; ++a)
    for (int bb = 0; bb < 3; ++bb) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) {
        const double xa = a == 0 ? 1.0 : (a == 1 ? x[i] : x[i] * x[i]);
        const double xb = bb == 0 ? 1.0 : (bb == 1 ? x[i] : x[i] * x[i]);
        s += xa * xb;
      }
      XtX[3 * a + bb] = s;
    }
  double inv[9];
  gauss_inverse3(XtX, inv);
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double xk = k == 0 ? 1.0 : (k == 1 ? x[i] : x[i] * x[i]);
        t += inv[3 * a + k] * xk;
      }
      s += t * y[i];
    }
    c[a] = s;
  }
}

// ROI pixel layout of a camera group: LPG lanes x SLOTS slots cover a
// PW x PW box (PW = sqrt(LPG * SLOTS), a multiple of LPG) in row-major
// order, so slot sl of lane sub holds ROI pixel (ro, co) = (sl * LPG / PW,
// sl * LPG % PW + sub): a compile-time row and a lane column.  ROIs
// narrower or shorter than PW leave the slots beyond them empty.
constexpr int roi_pitch(int n) {
  int r = 1;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

// Per-lane state of one object in one camera (the lane's group camera).
template <typename T, int SLOTS> struct CamLane {
  int obase;       // first lane of this lane's object
  uint64_t omask;  // the object's lanes
  int k;           // camera of this lane's group
  bool cam_ok;     // k < n_cam
  bool act;        // camera active
  PRange roi;      // augmented ROI (zero range for inactive cameras)
  int roiw, npix;
  T aug[SLOTS];    // this lane's augmented pixels
  T R[9], Tv[3], fx, fy, cx, cy;   // the lane camera's pinhole parameters
};

// The lane camera's pinhole parameters, held in registers for the whole
// kernel (each lane evaluates only its own camera).
template <typename T, int SLOTS>
__device__ __forceinline__ void cam_lane_load(const ShakeBuffers<T> &b, int k,
                                              CamLane<T, SLOTS> &c) {
  const CamT<T> &cm = b.cams[k];
#pragma unroll
  for (int i = 0; i < 9; ++i) c.R[i] = cm.R[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) c.Tv[i] = cm.Tv[i];
  c.fx = cm.fx;
  c.fy = cm.fy;
  c.cx = cm.cx;
  c.cy = cm.cy;
}

// Camera::project with the lane's registers (the distorted model reads the
// camera record), the same operations as project().
template <typename T, int SLOTS>
__device__ __forceinline__ void project_lane(const ShakeBuffers<T> &b, const CamLane<T, SLOTS> &c,
                                             const T X[3], T &u, T &v) {
  const CamT<T> &cam = b.cams[c.k];
  if (cam.is_distorted) {
    project(cam, X, u, v);
    return;
  }
  T t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    t[i] = mad_t(c.R[3 * i + 2], X[2], mad_t(c.R[3 * i + 1], X[1], c.R[3 * i] * X[0]));
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = t[i] + c.Tv[i];
  const T x = t[2] != T(0) ? t[0] / t[2] : t[0];
  const T y = t[2] != T(0) ? t[1] / t[2] : t[1];
  u = mad_t(x, c.fx, c.cx);
  v = mad_t(y, c.fy, c.cy);
}

// calShakeResidue (Shake.cpp:968-1045) at Xe; (ue, ve) receive the lane
// camera's projection.  Wave-uniform result.
template <typename T, int NC, int OPW, int SLOTS>
__device__ __forceinline__ T eval_residue(const ShakeBuffers<T> &b, const CamLane<T, SLOTS> &c,
                                          T r_px, const T Xe[3], T &ue, T &ve) {
  constexpr int LPG = kWave / OPW / NC;
  constexpr int PW = roi_pitch(LPG * SLOTS);
  static_assert(PW * PW == LPG * SLOTS && PW % LPG == 0, "ROI slots must tile a square");
  const int sub = lane_id() % LPG;
  T sse = T(0);
  ue = T(0);
  ve = T(0);
  if (c.cam_ok) project_lane(b, c, Xe, ue, ve);
  const bool use = c.act && c.roi.rmax - c.roi.rmin > 0 && c.roi.cmax - c.roi.cmin > 0;
  if (use) {
    T P[5];
    otf_param(b.otf, c.k, Xe, P);
    const int prmin = (int)floor(ve - r_px), prmax = (int)ceil(ve + r_px + T(1));
    const int pcmin = (int)floor(ue - r_px), pcmax = (int)ceil(ue + r_px + T(1));
    const int roih = c.roi.rmax - c.roi.rmin;
    // FP32: exp2 with log2(e) folded into b and c once per evaluation
    const T b2 = P[1] * T(1.4426950408889634), c2 = P[2] * T(1.4426950408889634);
    // FP32 without rotation: the Gaussian is a product of a column factor
    // and a row factor, each zero outside the object's window and the ROI
    // (and where the intensity is not positive), so a slot costs one
    // product: PW / LPG column and PW row exponentials per lane instead of
    // one per slot.  The product rounds differently from the exponential of
    // the sum, by an ulp or two.
    bool sep = false;
    if constexpr (sizeof(T) == 4) sep = P[3] == T(1) && P[4] == T(0);
    if (sep) {
      constexpr int NCOL = PW / LPG;
      T ex[NCOL], ey[PW];
#pragma unroll
      for (int q = 0; q < NCOL; ++q) {
        const int co = q * LPG + sub, col = c.roi.cmin + co;
        const T dx = T(col) - ue;
        const T e = __builtin_amdgcn_exp2f(-(b2 * dx * dx));
        ex[q] = co < c.roiw && col >= pcmin && col < pcmax && e > T(0) ? e : T(0);
      }
#pragma unroll
      for (int ro = 0; ro < PW; ++ro) {
        const int row = c.roi.rmin + ro;
        const T dy = T(row) - ve;
        const T g = P[0] * __builtin_amdgcn_exp2f(-(c2 * dy * dy));
        ey[ro] = ro < roih && row >= prmin && row < prmax && g > T(0) ? g : T(0);
      }
      // empty slots hold aug = 0 and a zero factor: they add an exact zero
#pragma unroll
      for (int sl = 0; sl < SLOTS; ++sl) {
        const T d = c.aug[sl] - ey[sl * LPG / PW] * ex[(sl * LPG % PW) / LPG];
        sse = mad_t(d, d, sse);
      }
    } else {
#pragma unroll
      for (int sl = 0; sl < SLOTS; ++sl) {
        const int ro = sl * LPG / PW, co = sl * LPG % PW + sub;
        const int row = c.roi.rmin + ro;
        const int col = c.roi.cmin + co;
        const bool valid = co < c.roiw && ro < roih;
        const bool win = row >= prmin && row < prmax && col >= pcmin && col < pcmax;
        T pred = T(0);
        if constexpr (sizeof(T) == 4) {
          // FP32: the intensity only where some lane of the wave needs it (a
          // wave-uniform branch: the ROI's edge rows mostly lie outside every
          // object's window), selected per lane without branches
          if (ballot(win) != 0) {
            const T dx = T(col) - ue, dy = T(row) - ve;
            const T xx = mad_t(dx, P[3], dy * P[4]);
            const T yy = mad_t(dy, P[3], -dx * P[4]);
            const T g = P[0] * __builtin_amdgcn_exp2f(-mad_t(b2 * xx, xx, c2 * yy * yy));
            pred = win && g > T(0) ? g : T(0);
          }
        } else {
          if (win) pred = gauss_int<T>(P, ue, ve, row, col);
        }
        // empty slots add an exact zero
        const T d = valid ? c.aug[sl] - pred : T(0);
        sse = mad_t(d, d, sse);
      }
    }
  }
  sse = group_sum<LPG>(sse);
  const T val = use ? sse / T(c.npix) : T(0);
  const uint64_t um = ballot(use && sub == 0) >> c.obase;
  T acc = T(0);
  int used = 0;
#pragma unroll
  for (int kk = 0; kk < NC; ++kk) {
    const T vk = __shfl(val, c.obase + kk * LPG, kWave);
    if ((um >> (kk * LPG)) & 1ull) {
      acc += vk;
      ++used;
    }
  }
  return used > 1 ? acc / T(used) : T(INFINITY);
}

// Whether the projection of Xt falls outside the ROI of any camera
// (shakeOneObject's pre-evaluation check; every camera takes part).
template <typename T, int SLOTS>
__device__ __forceinline__ bool any_out(const ShakeBuffers<T> &b, const CamLane<T, SLOTS> &c,
                                        const T Xt[3]) {
  bool out = false;
  if (c.cam_ok) {
    T u, v;
    project_lane(b, c, Xt, u, v);
    out = !in_roi_d(c.roi, u, v);
  }
  return (ballot(out) & c.omask) != 0;
}

// buildROIInfo (Shake.cpp:279-328) for object i with its current
// projection: this lane's camera ROI (ratio_augimg x the object radius) and
// its SLOTS pixels of aug = max(0, residual + own projection in the box).
template <typename T, int LPG, int SLOTS>
__device__ __forceinline__ void build_roi(const ShakeBuffers<T> &b, const ShakeParams &p,
                                          const ObjCamRec<T> *__restrict__ rec, int i,
                                          CamLane<T, SLOTS> &c) {
  const int sub = lane_id() % LPG;
  const long HW = (long)b.H * b.W;
  const T r_px = (T)p.r_px;
  const T dpos = r_px > T(0) ? r_px : T(0);
  c.roi = PRange{0, 0, 0, 0};
  c.roiw = 0;
  c.npix = 0;
#pragma unroll
  for (int sl = 0; sl < SLOTS; ++sl) c.aug[sl] = T(0);
  if (!c.act) return;
  const ObjCamRec<T> &r = rec[(long)i * b.n_cam + c.k];
  const PRange roi = roi_bound<T>(b.H, b.W, r.u, r.v, p.roi_w, p.roi_w);
  c.roi = roi;
  if (roi.empty()) return;
  c.roiw = roi.cmax - roi.cmin;
  const int roih = roi.rmax - roi.rmin;
  c.npix = roih * c.roiw;
  if (c.npix <= 0) return;
  constexpr int PW = roi_pitch(LPG * SLOTS);
  const T rP[5] = {r.P[0], r.P[1], r.P[2], r.P[3], r.P[4]};
#pragma unroll
  for (int sl = 0; sl < SLOTS; ++sl) {
    const int ro = sl * LPG / PW, co = sl * LPG % PW + sub;
    if (co < c.roiw && ro < roih) {
      const int row = roi.rmin + ro;
      const int col = roi.cmin + co;
      T a = b.res[c.k * HW + (long)row * b.W + col];
      if (row >= r.rmin && row < r.rmax && col >= r.cmin && col < r.cmax)
        a += gauss_int<T>(rP, r.u, r.v, row, col);
      c.aug[sl] = a > T(0) ? a : T(0);
    }
  }
}

template <typename T, int NC, int OPW, int SLOTS>
__global__ void __launch_bounds__(kWave)
k_shake(ShakeBuffers<T> bb, ShakeParams p, const ObjCamRec<T> *__restrict__ rec,
        T delta) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  rec = fshift(rec, fo);
  constexpr int LPO = kWave / OPW;   // lanes per object
  constexpr int LPG = LPO / NC;      // lanes per camera
  const int n = *b.n_obj;
  const int nc = b.n_cam;
  const int lane = lane_id();
  const int sub = lane % LPG;
  const long HW = (long)b.H * b.W;
  const T r_px = (T)p.r_px;
  const T dpos = r_px > T(0) ? r_px : T(0);
  CamLane<T, SLOTS> c;
  c.obase = (lane / LPO) * LPO;
  c.omask = (LPO == kWave ? ~0ull : ((1ull << LPO) - 1)) << c.obase;
  c.k = (lane % LPO) / LPG;
  c.cam_ok = c.k < nc;
  c.act = c.cam_ok && p.active[c.k];
  const int ksafe = c.cam_ok ? c.k : 0;
  cam_lane_load(b, ksafe, c);
  const int n_chunk = (n + OPW - 1) / OPW;
  for (int pass = 0; pass * (int)gridDim.x < n_chunk; ++pass) {
    const int chunk = xcd_chunk(pass, n_chunk - pass * (int)gridDim.x);
    const int ii = chunk * OPW + lane / LPO;
    if (chunk < 0 || ii >= n) continue;   // wave-uniform except in the last chunk
    const int i = b.order ? b.order[ii] : ii;
    T X[3] = {(T)b.xyz[3L * i], (T)b.xyz[3L * i + 1], (T)b.xyz[3L * i + 2]};
    build_roi<T, LPG, SLOTS>(b, p, rec, i, c);
    // shakeOneObject (:409-665); tmp keeps the last evaluated coordinate on
    // each axis, as the reference's reused temporary object does.  The
    // centre sample of an axis is tmp itself (the axis's own coordinate is
    // still X[ax]), so when the previous axis ended on an evaluation of tmp
    // its residue and projection are reused: the evaluation is a pure
    // function of the point and the object's fixed ROI, so the value is the
    // one a repeated evaluation would give, bit for bit.
    T tmp[3] = {X[0], X[1], X[2]};
    bool have_last = false;  // (f_last, u_last, v_last) evaluated at tmp
    T f_last = T(0), u_last = T(0), v_last = T(0);
    for (int ax = 0; ax < 3; ++ax) {
      const T x0 = X[ax];
      const T DL = -delta, DR = delta;
      const T rel[3] = {DL, T(0), DR};
      T f[3], ue[3], ve[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        tmp[ax] = x0 + rel[j];
        if (j == 1 && have_last) {
          f[1] = f_last;
          ue[1] = u_last;
          ve[1] = v_last;
        } else {
          f[j] = eval_residue<T, NC, OPW, SLOTS>(b, c, r_px, tmp, ue[j], ve[j]);
        }
      }
      have_last = true;
      f_last = f[2];
      u_last = ue[2];
      v_last = ve[2];
      const T fL = f[0], f0 = f[1], fR = f[2];
      const T tol = T(1e-3) * fmax(T(1), f0);
      const bool mono_dec = (fL > f0 + tol) && (f0 > fR + tol);
      const bool mono_inc = (fR > f0 + tol) && (f0 > fL + tol);
      const bool mono = mono_dec || mono_inc;
      int sgn = mono_dec ? 1 : (mono_inc ? -1 : 0);
      const T kappa = (fL - T(2) * f0 + fR);
      const T g = (fR - fL) / fmax(T(2) * delta, T(1e-12));
      T cx[5] = {DL, T(0), DR, T(0), T(0)}, cf[5] = {fL, f0, fR, T(0), T(0)};
      T cabs[5] = {x0 + DL, x0, x0 + DR, T(0), T(0)};
      int ncand = 3;
      bool skip_expand = false, allow_xhat = true;
      if (kappa > T(0)) {
        const T x3 = delta * (fL - fR) / (T(2) * kappa);
        if (fabs(x3) <= T(0.95) * delta) {
          const T a = kappa / (delta * delta + T(1e-12));
          const T rho = (g * g) / (T(2) * a * fmax(f0, T(1)));
          skip_expand = true;
          if (rho < T(8e-4)) allow_xhat = false;
        }
      }
      if (!skip_expand && mono) {
        // per-camera pixel speed and ROI margin, gathered in camera order
        const bool ok0 = in_roi_d(c.roi, ue[0], ve[0]);
        const bool ok1 = in_roi_d(c.roi, ue[1], ve[1]);
        const bool ok2 = in_roi_d(c.roi, ue[2], ve[2]);
        T Jk = T(0), mk = T(0);
        if (ok0 && ok2)
          Jk = hypot(ue[2] - ue[0], ve[2] - ve[0]) / fmax(T(2) * delta, T(1e-12));
        if (ok1) {
          const T mu = fmin(ue[1] - T(c.roi.cmin), T(c.roi.cmax) - T(1) - ue[1]);
          const T mv = fmin(ve[1] - T(c.roi.rmin), T(c.roi.rmax) - T(1) - ve[1]);
          mk = fmin(mu, mv);
        }
        const uint64_t mJ = ballot(c.cam_ok && sub == 0 && ok0 && ok2) >> c.obase;
        const uint64_t mM = ballot(c.cam_ok && sub == 0 && ok1) >> c.obase;
        T Jv[NC], mg[NC];
        int nJ = 0, nm = 0;
#pragma unroll
        for (int kk = 0; kk < NC; ++kk) {
          const T jv = __shfl(Jk, c.obase + kk * LPG, kWave);
          const T mv = __shfl(mk, c.obase + kk * LPG, kWave);
          if ((mJ >> (kk * LPG)) & 1ull) Jv[nJ++] = jv;
          if ((mM >> (kk * LPG)) & 1ull) mg[nm++] = mv;
        }
        const T Jx = median_small<T, NC>(Jv, nJ);
        const T mpx = median_small<T, NC>(mg, nm);
        const T Droi = (Jx > T(0) && mpx > T(0)) ? (T(0.5) * mpx / Jx) : T(0);
        T Dtry = T(0);
        if (kappa > T(0)) {
          const T x3 = delta * (fL - fR) / (T(2) * kappa);
          if (fabs(x3) > delta) {
            sgn = x3 > T(0) ? 1 : -1;
            if (Jx > T(0) && fabs(x3) > (T(1) / Jx))
              Dtry = T(1) / Jx;
            else
              Dtry = T(1.15) * fabs(x3);
          }
        }
        if (Dtry <= T(0) && Jx > T(0)) Dtry = T(1) / Jx;
        const T Dp = (Dtry > T(0) && Droi > T(0)) ? fmin(Dtry, Droi) : T(0);
        if (Dp > T(0)) {
          const T dxe = sgn >= 0 ? Dp : -Dp;
          tmp[ax] = x0 + dxe;
          have_last = !any_out<T, SLOTS>(b, c, tmp);
          if (have_last) {
            cx[ncand] = dxe;
            cf[ncand] = eval_residue<T, NC, OPW, SLOTS>(b, c, r_px, tmp, u_last, v_last);
            f_last = cf[ncand];
            cabs[ncand] = x0 + dxe;
            ++ncand;
          }
        }
      }
      double coef[3];
      if constexpr (sizeof(T) == 4) {
        // FP32 mode: the vertex of the least-squares parabola from the
        // normal equations in offsets scaled by 1/delta (well conditioned),
        // by cofactors: only c1 / c2 and the sign of c2 are needed, and the
        // determinant of the Gram matrix of >= 3 distinct offsets is > 0
        const double sc = 1.0 / (double)delta;
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0, t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
          if (t < ncand) {
            const double x = (double)cx[t] * sc, y = (double)cf[t];
            const double x2 = x * x;
            s1 += x;
            s2 += x2;
            s3 += x2 * x;
            s4 += x2 * x2;
            t0 += y;
            t1 += x * y;
            t2 += x2 * y;
          }
        }
        const double s0 = (double)ncand;
        const double c01 = s2 * s3 - s1 * s4, c02 = s1 * s3 - s2 * s2;
        const double c11 = s0 * s4 - s2 * s2, c12 = s1 * s2 - s0 * s3;
        const double c22 = s0 * s2 - s1 * s1;
        const double n1 = c01 * t0 + c11 * t1 + c12 * t2;
        const double n2 = c02 * t0 + c12 * t1 + c22 * t2;
        // c1 and c2 of the unscaled fit times the (positive) determinant
        coef[0] = 0.0;
        coef[1] = n1 * sc;
        coef[2] = n2 * sc * sc;
      } else {
        double dcx[5], dcf[5];
        for (int t = 0; t < ncand; ++t) {
          dcx[t] = (double)cx[t];
          dcf[t] = (double)cf[t];
        }
        polyfit2(dcx, dcf, ncand, coef);
      }
      if (allow_xhat && coef[2] > 0.0) {
        const T xh = (T)(-coef[1] / (2.0 * coef[2]));
        T xmin = cx[0], xmax = cx[0];
        for (int t = 0; t < ncand; ++t) {
          xmin = fmin(xmin, cx[t]);
          xmax = fmax(xmax, cx[t]);
        }
        if (xh >= xmin && xh <= xmax) {
          tmp[ax] = x0 + xh;
          have_last = !any_out<T, SLOTS>(b, c, tmp);
          if (have_last) {
            cx[ncand] = xh;
            cf[ncand] = eval_residue<T, NC, OPW, SLOTS>(b, c, r_px, tmp, u_last, v_last);
            f_last = cf[ncand];
            cabs[ncand] = x0 + xh;
            ++ncand;
          }
        }
      }
      int bi = 0;
      T bf = cf[0];
      for (int t = 1; t < ncand; ++t)
        if (cf[t] < bf) {
          bf = cf[t];
          bi = t;
        }
      X[ax] = cabs[bi];
    }
    // final projection and calObjectScore (:668-782)
    T uf = T(0), vf = T(0), Pf[5] = {T(0), T(0), T(0), T(0), T(0)};
    if (c.cam_ok) {
      project_lane(b, c, X, uf, vf);
      otf_param(b.otf, c.k, X, Pf);
    }
    bool usek = false;
    PRange rect{1, 0, 1, 0};
    if (c.act && dpos > T(0)) {
      rect = roi_bound<T>(b.H, b.W, uf, vf, p.box_w, p.box_w);
      usek = !rect.empty();
    }
    const int rw = usek ? rect.cmax - rect.cmin : 0;
    const int np = usek ? (rect.rmax - rect.rmin) * rw : 0;
    const int np_all = group_max<LPO>(np);
    T pk = T(0), nm = T(0), dn = T(0);
    for (int base = 0; base < np_all; base += LPG) {
      const int idx = base + sub;
      const bool valid = idx < np;
      int row = 0, col = 0;
      if (valid) {
        row = rect.rmin + idx / rw;
        col = rect.cmin + idx % rw;
      }
      // augmented value if inside this camera's ROI, else the residual
      const PRange &ro = c.roi;
      const bool inr = valid && ro.rmax - ro.rmin > 0 && ro.cmax - ro.cmin > 0 &&
                       row >= ro.rmin && row < ro.rmax && col >= ro.cmin && col < ro.cmax;
      const int aidx = inr ? (row - ro.rmin) * roi_pitch(LPG * SLOTS) + (col - ro.cmin) : 0;
      const int src = (lane / LPG) * LPG + aidx % LPG;   // same camera group
      T av = T(0);
#pragma unroll
      for (int sl = 0; sl < SLOTS; ++sl) {
        const T got = __shfl(c.aug[sl], src, kWave);
        if (aidx / LPG == sl) av = got;
      }
      T meas = T(0), pred = T(0);
      if (valid) {
        meas = inr ? av : b.res[ksafe * HW + (long)row * b.W + col];
        pred = gauss_int<T>(Pf, uf, vf, row, col);
      }
      const T mpos = meas > T(0) ? meas : T(0);
      pk = fmax(pk, valid ? meas : T(0));
      nm += valid ? mpos : T(0);
      dn += valid ? pred : T(0);
    }
    pk = group_max<LPG>(pk);
    nm = group_sum<LPG>(nm);
    dn = group_sum<LPG>(dn);
    const uint64_t mu = ballot(usek && sub == 0) >> c.obase;
    T peak[NC], num[NC], den[NC];
    int use[NC];
    int n_used = 0;
#pragma unroll
    for (int kk = 0; kk < NC; ++kk) {
      peak[kk] = __shfl(pk, c.obase + kk * LPG, kWave);
      num[kk] = __shfl(nm, c.obase + kk * LPG, kWave);
      den[kk] = __shfl(dn, c.obase + kk * LPG, kWave);
      use[kk] = (int)((mu >> (kk * LPG)) & 1ull);
      n_used += use[kk];
    }
    T score;
    if (n_used < 2) {
      score = T(1e-12);
    } else {
      if (n_used > 2) {
        int mi = -1;
        T mv = -T(INFINITY);
#pragma unroll
        for (int kk = 0; kk < NC; ++kk)
          if (use[kk] && peak[kk] > mv) {
            mv = peak[kk];
            mi = kk;
          }
#pragma unroll
        for (int kk = 0; kk < NC; ++kk)
          if (kk == mi) use[kk] = 0;
      }
      T tn = T(0), td = T(0);
#pragma unroll
      for (int kk = 0; kk < NC; ++kk)
        if (use[kk]) {
          tn += num[kk];
          td += den[kk];
        }
      T ratio = (td > T(1e-12)) ? (tn / td) : T(1);
      ratio = fmax(T(2) / T(3), fmin(T(1.5), ratio));
      score = sqrt(ratio);
    }
    if (lane == c.obase) {
      b.xyz[3L * i] = (double)X[0];
      b.xyz[3L * i + 1] = (double)X[1];
      b.xyz[3L * i + 2] = (double)X[2];
      b.score[i] = b.score[i] * (double)score;
    }
    if (c.cam_ok && sub == 0) {
      b.uv[((long)i * nc + c.k) * 2] = uf;
      b.uv[((long)i * nc + c.k) * 2 + 1] = vf;
    }
  }
}

// --------------------------------------------------------------------------
// learned shake proposal
// --------------------------------------------------------------------------
// Network input of every object: per camera the 11x11 window of the
// augmented residual max(0, residual + own projection in its box) centred on
// the rounded projection, divided by the OTF amplitude there; 4 x 121 values
// in camera order, zero outside the image and for inactive cameras, then 28
// zeros of padding (bf16 [n][512]).
constexpr int kPatch = 11, kPatchHalf = 5, kPatchIn = 512;

// One wave per object: camera by camera the object's record is the same for
// every lane (scalar loads), and the 121 window pixels are spread over the
// lanes; each 512-value row is written whole by its wave.
constexpr int kPatchWaves = 4;   // waves (objects) per workgroup

template <typename T>
__global__ void __launch_bounds__(kPatchWaves * kWave)
k_patch_gather(ShakeBuffers<T> bb, ShakeParams p, const ObjCamRec<T> *__restrict__ rec,
               __bf16 *__restrict__ X) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  rec = fshift(rec, fo);
  X = fshift(X, (long)F * bb.lstride);
  const int n = *b.n_obj;
  const int nc = b.n_cam;
  const long HW = (long)b.H * b.W;
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
  for (int i = blockIdx.x * kPatchWaves + wave; i < n; i += gridDim.x * kPatchWaves) {
    __bf16 *row_out = X + (long)i * kPatchIn;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool cam = k < nc && p.active[k];
      ObjCamRec<T> r{};
      bool inside = false;
      int cu = 0, cv = 0;
      if (cam) {
        r = rec[(long)i * nc + k];
        inside = fabs((double)r.u) < 1e6 && fabs((double)r.v) < 1e6;
        if (inside) {
          cu = (int)round((double)r.u);
          cv = (int)round((double)r.v);
        }
      }
      const T amp = r.P[0] > T(1e-6) ? r.P[0] : T(1e-6);
      for (int pix = lane; pix < kPatch * kPatch; pix += kWave) {
        float v = 0.0f;
        if (inside) {
          const int row = cv + pix / kPatch - kPatchHalf, col = cu + pix % kPatch - kPatchHalf;
          if (row >= 0 && row < b.H && col >= 0 && col < b.W) {
            T a = b.res[k * HW + (long)row * b.W + col];
            if (row >= r.rmin && row < r.rmax && col >= r.cmin && col < r.cmax)
              a += gauss_int<T>(r.P, r.u, r.v, row, col);
            a = a > T(0) ? a : T(0);
            v = (float)(a / amp);
          }
        }
        row_out[k * kPatch * kPatch + pix] = (__bf16)v;
      }
    }
    // the padding after the 4 x 121 values
    if (lane < kPatchIn - 4 * kPatch * kPatch) row_out[4 * kPatch * kPatch + lane] = (__bf16)0.0f;
  }
}

// Take the network's step X + clamp(dX) * voxel when it lowers the shake
// residue (calShakeResidue on the same ROI); otherwise keep X.
template <typename T, int NC, int OPW, int SLOTS>
__global__ void __launch_bounds__(kWave)
k_propose(ShakeBuffers<T> bb, ShakeParams p, const ObjCamRec<T> *__restrict__ rec,
          const float *__restrict__ prop) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  rec = fshift(rec, fo);
  prop = fshift(prop, (long)F * bb.pstride);
  constexpr int LPO = kWave / OPW;
  const int n = *b.n_obj;
  const int lane = lane_id();
  const T r_px = (T)p.r_px;
  CamLane<T, SLOTS> c;
  c.obase = (lane / LPO) * LPO;
  c.omask = (LPO == kWave ? ~0ull : ((1ull << LPO) - 1)) << c.obase;
  c.k = (lane % LPO) / (LPO / NC);
  c.cam_ok = c.k < b.n_cam;
  c.act = c.cam_ok && p.active[c.k];
  cam_lane_load(b, c.cam_ok ? c.k : 0, c);
  const T lim = (T)p.max_step_vox, vox = (T)p.voxel_mm;
  const int n_chunk = (n + OPW - 1) / OPW;
  for (int pass = 0; pass * (int)gridDim.x < n_chunk; ++pass) {
    const int chunk = xcd_chunk(pass, n_chunk - pass * (int)gridDim.x);
    const int ii = chunk * OPW + lane / LPO;
    if (chunk < 0 || ii >= n) continue;   // wave-uniform except in the last chunk
    const int i = b.order ? b.order[ii] : ii;
    T X[3] = {(T)b.xyz[3L * i], (T)b.xyz[3L * i + 1], (T)b.xyz[3L * i + 2]};
    build_roi<T, LPO / NC, SLOTS>(b, p, rec, i, c);
    T ue, ve;
    const T f0 = eval_residue<T, NC, OPW, SLOTS>(b, c, r_px, X, ue, ve);
    T Xn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T s = (T)prop[4L * i + d];
      s = s > lim ? lim : (s < -lim ? -lim : s);
      Xn[d] = X[d] + s * vox;
    }
    const T f1 = eval_residue<T, NC, OPW, SLOTS>(b, c, r_px, Xn, ue, ve);
    if (f1 < f0 && lane == c.obase) {
      b.xyz[3L * i] = (double)Xn[0];
      b.xyz[3L * i + 1] = (double)Xn[1];
      b.xyz[3L * i + 2] = (double)Xn[2];
    }
  }
}

// --------------------------------------------------------------------------
// repeated objects + ghost threshold
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int grid_cell_of(const ShakeBuffers<T> &b, const double *X,
                                            int d) {
  int c = (int)floor((X[d] - b.grid_lo[d]) / b.grid_cell);
  return c < 0 ? 0 : (c >= b.grid_n ? b.grid_n - 1 : c);
}

template <typename T>
__global__ void k_grid_count(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  const int G = b.grid_n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double *X = b.xyz + 3L * i;
    const int cell = (grid_cell_of(b, X, 0) * G + grid_cell_of(b, X, 1)) * G + grid_cell_of(b, X, 2);
    b.g_cellof[i] = cell;
    b.g_rank[i] = atomicAdd(&b.g_cnt[cell], 1);
  }
}

template <typename T>
__global__ void k_grid_fill(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b.g_obj[b.g_start[b.g_cellof[i]] + b.g_rank[i]] = i;
}

// Processing order of the per-object shake kernels: objects sorted by the
// Morton code of their cell in a 64^3 grid over the repeated-object grid's
// box (a counting sort through the g_* arrays, free until markRepeatedObj
// rebuilds them).  Objects close in 3D project close in every camera.
constexpr int kOrderBits = 6;
__device__ __forceinline__ unsigned spread3(unsigned v) {
  unsigned r = 0;
#pragma unroll
  for (int k = 0; k < kOrderBits; ++k) r |= ((v >> k) & 1u) << (3 * k);
  return r;
}

template <typename T>
__global__ void k_order_count(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  const double cell = b.grid_cell * b.grid_n / (double)(1 << kOrderBits);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    unsigned key = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int c = (int)floor((b.xyz[3L * i + d] - b.grid_lo[d]) / cell);
      c = c < 0 ? 0 : (c >= (1 << kOrderBits) ? (1 << kOrderBits) - 1 : c);
      key |= spread3((unsigned)c) << d;
    }
    b.g_cellof[i] = (int)key;
    b.g_rank[i] = atomicAdd(&b.g_cnt[key], 1);
  }
}

// nanoflann L2_Simple_Adaptor distance with diff = query - point
__device__ __forceinline__ double dist2(const double *a, const double *c) {
  const double d0 = a[0] - c[0], d1 = a[1] - c[1], d2 = a[2] - c[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// Objects closer than sqrt(t2) to object i (marked with mark_val when mark
// is given); with first, 1 at the first one (only whether there is one).
template <typename T>
__device__ __forceinline__ int count_nb(const ShakeBuffers<T> &b, int i, double t2,
                                        int8_t *mark, int mark_val, bool first = false) {
  const double *X = b.xyz + 3L * i;
  const int G = b.grid_n;
  const int cx = grid_cell_of(b, X, 0), cy = grid_cell_of(b, X, 1), cz = grid_cell_of(b, X, 2);
  int found = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int x = cx + dx, y = cy + dy, z = cz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) continue;
        const int cell = (x * G + y) * G + z;
        const int s0 = b.g_start[cell], ncell = b.g_start[cell + 1] - s0;
        for (int t = 0; t < ncell; ++t) {
          const int j = b.g_obj[s0 + t];
          if (j == i) continue;
          if (dist2(X, b.xyz + 3L * j) < t2) {
            if (first) return 1;
            ++found;
            if (mark) mark[j] = (int8_t)mark_val;
          }
        }
      }
  return found;
}

template <typename T>
__global__ void k_has_nb(ShakeBuffers<T> bb, double t2) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  // has_nb is scanned over the first n entries only
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    b.has_nb[i] = count_nb(b, i, t2, nullptr, 0, true);
    b.flag[i] = OLPT_FLAG_NONE;
  }
}

// greedy pass (index order) over the objects that have a neighbour; the
// last object is never a query (Shake.cpp:797: i < n_obj - 1).  Marks are
// written as OLPT_FLAG_REPEATED into flag[].
template <typename T>
__global__ void k_repeat_greedy(ShakeBuffers<T> bb, double t2) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = *b.n_obj;
  const int m = b.totals[0];
  int8_t *rep = (int8_t *)b.flag;
  for (int q = 0; q < m; ++q) {
    const int i = b.keep_off[q];
    if (i >= n - 1) break;
    if (rep[i]) continue;
    count_nb(b, i, t2, rep, OLPT_FLAG_REPEATED);
  }
}

// compact indices of objects with neighbours (order preserved)
template <typename T>
__global__ void k_nb_list(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (b.has_nb[i]) b.keep_off[b.g_rank[i]] = i;
}

// deterministic two-level reduction of the non-repeated scores
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 256;
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_score_sum(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  __shared__ double sd[kRedThreads];
  __shared__ int sc[kRedThreads];
  const int n = *b.n_obj;
  double s = 0.0;
  int c = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (b.flag[i] == OLPT_FLAG_NONE) {
      s += b.score[i];
      ++c;
    }
  sd[threadIdx.x] = s;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sd[threadIdx.x] += sd[threadIdx.x + o];
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    b.partial[2 * blockIdx.x] = sd[0];
    b.partial[2 * blockIdx.x + 1] = (double)sc[0];
  }
}

// Ghost: score < thred_ghost * mean(non-repeated) or score == 0
// (Shake.cpp:118-136); keep_off <- 1 for unflagged objects.  With the
// learned proposal run in this shake (use_logit), also when its ghost logit
// exceeds ghost_logit (no counterpart in the reference).
template <typename T>
__global__ void k_flags(ShakeBuffers<T> bb, double thr, bool use_logit, double ghost_logit) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  double sum = 0.0, cnt = 0.0;
  for (int k = 0; k < kRedBlocks; ++k) {
    sum += b.partial[2 * k];
    cnt += b.partial[2 * k + 1];
  }
  const double mean = cnt > 0 ? sum / cnt : 0.0;
  const int n = *b.n_obj;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint8_t f = b.flag[i];
    const double sc = b.score[i];
    if (sc < mean * thr || sc == 0.0) f |= OLPT_FLAG_GHOST;
    if (use_logit && (double)b.prop[4L * i + 3] > ghost_logit) f |= OLPT_FLAG_GHOST;
    b.flag[i] = f;
  }
}

template <typename T>
__global__ void k_init_score(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    b.score[i] = 1.0;
    b.flag[i] = 0;
  }
}

template <typename T>
__global__ void k_zero_score(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const int n = *b.n_obj;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b.score[i] = 0.0;
}

template <typename T, int NC, int OPW, int SLOTS>
static void shake_iter(const ShakeBuffers<T> &b, const ShakeParams &p,
                       const ObjCamRec<T> *rec, T delta, hipStream_t st) {
  stage_begin(ST_SHAKE, st);
  k_shake<T, NC, OPW, SLOTS><<<dim3(launch_blocks(b.grid_shake, b.nf, 4 * kMinLaunchBlocks), 1, b.nf), kWave, 0, st>>>(
      b, p, rec, delta);
  stage_end(ST_SHAKE, st);
}

// Objects per wave for up to 4 cameras and an 8x8 ROI: 8 in FP32 (measured
// fastest with 16-frame launches: the per-camera projection and OTF work
// covers 32 pixels per lane; 4 objects per wave took 23% longer, 16 need
// 277 registers), 2 in FP64, whose doubled registers would otherwise spill.
template <typename T> constexpr int shake_opw() { return sizeof(T) == 4 ? 8 : 2; }

template <typename T>
void launch_shake(const ShakeBuffers<T> &b, const ShakeParams &p, hipStream_t st) {
  const int nc = b.n_cam;
  CamList act{0, {0}};
  for (int k = 0; k < nc; ++k)
    if (p.active[k]) act.id[act.n++] = k;
  CamList inact{0, {0}};
  for (int k = 0; k < nc; ++k)
    if (!p.active[k]) inact.id[inact.n++] = k;
  ObjCamRec<T> *rec = reinterpret_cast<ObjCamRec<T> *>(b.rec);
  const double r_px = p.r_px;
  // the ROI edge, as every kernel uses it (calROIBound, decided on the host)
  const int roi_px = p.roi_w > 0 ? p.roi_w * p.roi_w : 0;
  const int gO = grid_for(b.cap_obj, 256, 2048);
  k_init_score<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
  // the spatial processing order (g_cnt is all zero here; the scan clears
  // it), used by the per-object kernels until markRepeatedObj
  ShakeBuffers<T> bo = b;
  if ((long)b.grid_n * b.grid_n * b.grid_n >= (1L << (3 * kOrderBits))) {
    const long nkey = 1L << (3 * kOrderBits);
    k_order_count<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
    scan_exclusive_clear(b.g_cnt, b.g_start, nkey, b.g_start + nkey, b.scan_scratch, st, b.nf,
                         b.fstride);
    k_grid_fill<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
    bo.order = b.g_obj;
  }
  const bool learned = p.mlp && b.patches && b.prop && nc == 4 && roi_px <= 64;
  if (learned) {
    // learned proposal from the matched positions, then the plain iterations
    residual_pass<T>(b, b.img, b.res, b.xyz, b.n_obj, rec, b.uv, act, 0, r_px, st);
    k_patch_gather<T><<<dim3(grid_for(b.cap_obj, kPatchWaves, 8192), 1, b.nf), kPatchWaves * kWave,
                        0, st>>>(b, p, rec, (__bf16 *)b.patches);
    mlp_infer_device(p.mlp, b.patches, b.n_obj, b.cap_obj, b.prop, st, b.nf, b.lstride,
                     b.fstride, b.pstride);
    constexpr int O = shake_opw<T>();
    k_propose<T, 4, O, 4 * O><<<dim3(launch_blocks(b.grid_shake, b.nf, 4 * kMinLaunchBlocks), 1, b.nf), kWave, 0, st>>>(
        bo, p, rec, b.prop);
  }
  double delta = p.shake_width;
  double dmin = p.shakewidth_min > 0 ? p.shakewidth_min : delta / 20;
  for (int it = 0; it < p.n_shake_loop; ++it) {
    // residual from the start-of-iteration positions (active cameras); the
    // projection pass also refreshes uv (runShake reprojects first)
    residual_pass<T>(b, b.img, b.res, b.xyz, b.n_obj, rec, b.uv, act, 0, r_px, st);
    if (nc < 2) {
      k_zero_score<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
    } else if (nc <= 4 && roi_px <= 64) {
      shake_iter<T, 4, shake_opw<T>(), 4 * shake_opw<T>()>(bo, p, rec, (T)delta, st);
    } else if (nc <= 4) {
      shake_iter<T, 4, 1, 16>(bo, p, rec, (T)delta, st);
    } else if (roi_px <= 64) {
      shake_iter<T, 8, 1, 8>(bo, p, rec, (T)delta, st);
    } else {
      shake_iter<T, 8, 1, 32>(bo, p, rec, (T)delta, st);
    }
    delta *= 0.5;
    if (delta < dmin) delta = dmin;
  }
  if (p.n_shake_loop <= 0) {
    k_obj_project<T, false><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b, b.xyz, b.n_obj, rec, b.uv, p.box_w);
  }
  // markRepeatedObj
  const double t2 = p.tol3d * p.tol3d;
  const long ncell = (long)b.grid_n * b.grid_n * b.grid_n;
  k_grid_count<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);  // g_cnt is all zero here; the scan clears it
  scan_exclusive_clear(b.g_cnt, b.g_start, ncell, b.g_start + ncell, b.scan_scratch, st, b.nf, b.fstride);
  k_grid_fill<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
  k_has_nb<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b, t2);
  scan_exclusive(b.has_nb, b.g_rank, b.cap_obj, 1, b.cap_obj, b.totals, b.scan_scratch, st, b.nf,
                 b.fstride, b.n_obj);
  k_nb_list<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
  k_repeat_greedy<T><<<dim3(1, 1, b.nf), 64, 0, st>>>(b, t2);
  // ghost threshold: mean over non-repeated scores
  k_score_sum<T><<<dim3(kRedBlocks, 1, b.nf), kRedThreads, 0, st>>>(b);
  k_flags<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b, p.thred_ghost,
                                                  learned && p.ghost_logit < INFINITY,
                                                  p.ghost_logit);
  (void)inact;
}

// --------------------------------------------------------------------------
// filter + IPR-mode residual
// --------------------------------------------------------------------------
template <typename T>
__global__ void k_keep_write(ShakeBuffers<T> bb, const int32_t *__restrict__ keep,
                             const int32_t *__restrict__ off, double *xyz_keep,
                             T *uv_keep) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  const long fo = F * bb.fstride;
  keep = fshift(keep, fo);
  off = fshift(off, fo);
  xyz_keep = fshift(xyz_keep, fo);
  uv_keep = fshift(uv_keep, fo);
  const int n = *b.n_obj;
  const int nc = b.n_cam;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!keep[i]) continue;
    const int o = off[i];
    for (int d = 0; d < 3; ++d) xyz_keep[3L * o + d] = b.xyz[3L * i + d];
    for (int k = 0; k < nc; ++k) {
      uv_keep[((long)o * nc + k) * 2] = b.uv[((long)i * nc + k) * 2];
      uv_keep[((long)o * nc + k) * 2 + 1] = b.uv[((long)i * nc + k) * 2 + 1];
    }
  }
}

template <typename T>
__global__ void k_keep_flags(ShakeBuffers<T> bb) {
  const long F = (long)blockIdx.z;
  const ShakeBuffers<T> b = bb.frame(F);
  // scanned over the first n_obj entries only
  const int n = *b.n_obj;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b.has_nb[i] = b.flag[i] == OLPT_FLAG_NONE ? 1 : 0;
}

template <typename T>
__global__ void k_keep_count(const int32_t *tot, int32_t *n_keep, long fstride) {
  tot = fshift(tot, (long)blockIdx.z * fstride);
  n_keep = fshift(n_keep, (long)blockIdx.z * fstride);
  if (threadIdx.x == 0 && blockIdx.x == 0) *n_keep = tot[0];
}

template <typename T>
void launch_filter(const ShakeBuffers<T> &b, double *xyz_keep, T *uv_keep,
                   int32_t *n_keep, hipStream_t st) {
  const int gO = grid_for(b.cap_obj, 256, 2048);
  k_keep_flags<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b);
  scan_exclusive(b.has_nb, b.keep_off, b.cap_obj, 1, b.cap_obj, b.totals,
                 b.scan_scratch, st, b.nf, b.fstride, b.n_obj);
  k_keep_write<T><<<dim3(gO, 1, b.nf), 256, 0, st>>>(b, b.has_nb, b.keep_off, xyz_keep, uv_keep);
  k_keep_count<T><<<dim3(1, 1, b.nf), 64, 0, st>>>(b.totals, n_keep, b.fstride);
}

template <typename T>
void launch_residual_ipr(const ShakeBuffers<T> &b, const int32_t *n_keep,
                         const double *xyz_keep, T *uv_keep, double r_px,
                         hipStream_t st) {
  CamList all{0, {0}};
  for (int k = 0; k < b.n_cam; ++k) all.id[all.n++] = k;
  ObjCamRec<T> *rec = reinterpret_cast<ObjCamRec<T> *>(b.rec);
  residual_pass<T>(b, b.img, b.img_next, xyz_keep, n_keep, rec, uv_keep, all, 1, r_px, st);
}

template <typename T>
void launch_patches(const ShakeBuffers<T> &b, const ShakeParams &p, void *X, hipStream_t st) {
  CamList act{0, {0}};
  for (int k = 0; k < b.n_cam; ++k)
    if (p.active[k]) act.id[act.n++] = k;
  ObjCamRec<T> *rec = reinterpret_cast<ObjCamRec<T> *>(b.rec);
  residual_pass<T>(b, b.img, b.res, b.xyz, b.n_obj, rec, b.uv, act, 0, p.r_px, st);
  k_patch_gather<T><<<dim3(grid_for(b.cap_obj, kPatchWaves, 8192), 1, b.nf), kPatchWaves * kWave, 0,
                      st>>>(b, p, rec, (__bf16 *)X);
}

template <typename T>
void launch_render(const ShakeBuffers<T> &b, const double *xyz, const double *amp,
                   const int32_t *n_ptr, double r_px, T *out, hipStream_t st) {
  CamList all{0, {0}};
  for (int k = 0; k < b.n_cam; ++k) all.id[all.n++] = k;
  ObjCamRec<T> *rec = reinterpret_cast<ObjCamRec<T> *>(b.rec);
  residual_pass<T>(b, out, out, xyz, n_ptr, rec, nullptr, all, 2, r_px, st, amp);
}

size_t shake_rec_bytes(bool f64) {
  return f64 ? sizeof(ObjCamRec<double>) : sizeof(ObjCamRec<float>);
}

#define INST(T)                                                                   \
  template void launch_shake<T>(const ShakeBuffers<T> &, const ShakeParams &,     \
                                hipStream_t);                                     \
  template void launch_residual_ipr<T>(const ShakeBuffers<T> &, const int32_t *,  \
                                       const double *, T *, double, hipStream_t); \
  template void launch_filter<T>(const ShakeBuffers<T> &, double *, T *,          \
                                 int32_t *, hipStream_t);                          \
  template void launch_patches<T>(const ShakeBuffers<T> &, const ShakeParams &,   \
                                  void *, hipStream_t);                            \
  template void launch_render<T>(const ShakeBuffers<T> &, const double *,         \
                                 const double *, const int32_t *, double, T *,     \
                                 hipStream_t);
INST(float)
INST(double)
#undef INST

} // namespace olpt
