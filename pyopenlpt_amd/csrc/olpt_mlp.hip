// olpt_mlp.hip - learned shake proposal: a BF16 MLP on MFMA.
//
// No counterpart in the reference.  A candidate's four 11x11 augmented
// residual patches (484 values, zero-padded to 512) map to a 3D correction
// (in voxels) and a ghost logit:
//     512 -> 256 -> 256 -> 256 -> 4 (padded to 32),  GELU between layers.
// Weights keep an FP32 master copy (Adam) and two BF16 copies: WT [N][K]
// (forward: the B fragments of v_mfma_f32_32x32x16_bf16 are 8 consecutive k
// of one output column, one 16-byte load) and W [K][N] (backward data pass,
// 8 consecutive n).  Accumulation is FP32 throughout.
//
// Fragment layout of v_mfma_f32_32x32x16_bf16 (lane l, element i / j):
//   A[32][16]: row l%32, k (l/32)*8+i      B[16][32]: k (l/32)*8+i, col l%32
//   D[32][32]: row (j/4)*8 + (l/32)*4 + j%4, col l%32
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "olpt_c.h"

namespace olpt_mlp_impl {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 484, K0 = 512, H = 256, NOut = 4, N4 = 32;
constexpr int kRows = 64;            // batch rows per workgroup
constexpr int kLdsLd = H + 8;        // padded LDS row (bf16) against bank conflicts

// parameter blob offsets (floats): W1 [K0][H], b1, W2 [H][H], b2, W3, b3, W4 [H][N4], b4
constexpr long oW1 = 0, ob1 = oW1 + (long)K0 * H, oW2 = ob1 + H, ob2 = oW2 + (long)H * H,
               oW3 = ob2 + H, ob3 = oW3 + (long)H * H, oW4 = ob3 + H, ob4 = oW4 + (long)H * N4,
               kParams = ob4 + N4;

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ bf16x8 load8(const bf16 *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// acc[r][c] += A[r*32.., k] * B[k, c*32..] over K (multiple of 16).  A rows
// are contiguous in k (lda), Bt rows (one per output column) contiguous in k
// (ldb).  Rows of A at or beyond a_rows read as zero.
template <int R, int C>
__device__ __forceinline__ void mma_block(f32x16 (&acc)[R][C], const bf16 *A, int lda, int a_rows,
                                          const bf16 *Bt, int ldb, int K) {
  const int l = threadIdx.x % 64;
  const int kk = (l / 32) * 8;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x8 a[R], b[C];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = r * 32 + l % 32;
      if (row < a_rows)
        a[r] = load8(A + (long)row * lda + k0 + kk);
      else
        a[r] = bf16x8{};
    }
#pragma unroll
    for (int c = 0; c < C; ++c) b[c] = load8(Bt + (long)(c * 32 + l % 32) * ldb + k0 + kk);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < C; ++c)
        acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[r], b[c], acc[r][c], 0, 0, 0);
  }
}

__device__ __forceinline__ int acc_row(int j) {
  const int l = threadIdx.x % 64;
  return (j / 4) * 8 + (l / 32) * 4 + (j % 4);
}

struct Net {
  const bf16 *WT1, *WT2, *WT3, *WT4;   // [N][K]
  const bf16 *Wb2, *Wb3, *Wb4;          // [K][N]
  const float *b1, *b2, *b3, *b4;
};

struct TrainBuf {
  float *Z1, *Z2, *Z3;      // [B][H] pre-activations
  bf16 *AT0, *AT1, *AT2, *AT3;  // transposed layer inputs: [K0][B], [H][B] x3
  bf16 *dZ, *dZb;           // [B][H] ping-pong, bf16
  bf16 *dZT;                // [H][B] transposed gradient of the current layer
  bf16 *dY;                 // [B][N4]
  bf16 *dYT;                // [N4][B]
  float *grad;              // [kParams]
  int B;                    // padded batch (multiple of kRows)
};

// Forward pass of kRows rows.  Hidden activations stay in LDS; in training
// mode pre-activations and transposed inputs are kept for the backward pass.
// The row count is n_host, or *n_dev when given; blocks stride over the
// 64-row tiles.
template <bool TRAIN>
__global__ void __launch_bounds__(256) k_fwd(Net net, const bf16 *X, int n_host, const int32_t *n_dev,
                                             float *Y, TrainBuf tb, long xs = 0, long ns = 0,
                                             long ys = 0) {
  __shared__ bf16 hA[kRows * kLdsLd], hB[kRows * kLdsLd];
  // inference over a batch of frames: frame blockIdx.z's rows, count and
  // outputs lie xs, ns and ys bytes per frame after frame 0's
  X = (const bf16 *)((const char *)X + (long)blockIdx.z * xs);
  if (n_dev) n_dev = (const int32_t *)((const char *)n_dev + (long)blockIdx.z * ns);
  Y = (float *)((char *)Y + (long)blockIdx.z * ys);
  const int n = n_dev ? *n_dev : n_host;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  for (int m0 = blockIdx.x * kRows; m0 < n; m0 += gridDim.x * kRows) {
  const int rows = min(kRows, n - m0);
  if (TRAIN) {
    // transposed input for the first layer's weight gradient
    for (int i = threadIdx.x; i < kRows * K0; i += 256) {
      const int r = i / K0, k = i % K0;
      tb.AT0[(long)k * tb.B + m0 + r] = r < rows ? X[(long)(m0 + r) * K0 + k] : bf16(0.0f);
    }
  }
  const int ncol = w * 64;  // this wave's 64 output columns of a 256-wide layer
  // ---- three hidden layers
  const bf16 *WT[3] = {net.WT1, net.WT2, net.WT3};
  const float *bias[3] = {net.b1, net.b2, net.b3};
  float *Zs[3] = {tb.Z1, tb.Z2, tb.Z3};
  bf16 *ATs[3] = {tb.AT1, tb.AT2, tb.AT3};
  bf16 *out_lds[3] = {hA, hB, hA};
#pragma unroll
  for (int layer = 0; layer < 3; ++layer) {
    f32x16 acc[2][2] = {};
    if (layer == 0)
      mma_block<2, 2>(acc, X + (long)m0 * K0, K0, rows, WT[0] + (long)ncol * K0, K0, K0);
    else
      mma_block<2, 2>(acc, layer == 1 ? hA : hB, kLdsLd, kRows, WT[layer] + (long)ncol * H, H, H);
    __syncthreads();  // the next store may overwrite the LDS buffer just read
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int row = r * 32 + acc_row(j), col = ncol + c * 32 + l % 32;
          const float z = acc[r][c][j] + bias[layer][col];
          const float a = gelu(z);
          out_lds[layer][row * kLdsLd + col] = bf16(a);
          if (TRAIN) {
            const bool live = row < rows;
            Zs[layer][(long)(m0 + row) * H + col] = live ? z : 0.0f;
            ATs[layer][(long)col * tb.B + m0 + row] = live ? bf16(a) : bf16(0.0f);
          }
        }
    __syncthreads();
  }
  // ---- output layer: 256 -> 32 (4 used); waves 0 and 1 take 32 rows each
  if (w < 2) {
    f32x16 acc[1][1] = {};
    mma_block<1, 1>(acc, hA + (w * 32) * kLdsLd, kLdsLd, 32, net.WT4, H, H);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int row = w * 32 + acc_row(j), col = l % 32;
      if (col < NOut && row < rows) Y[(long)(m0 + row) * NOut + col] = acc[0][0][j] + net.b4[col];
    }
  }
  __syncthreads();  // the next tile overwrites the LDS activations
  }
}

// Loss on the outputs: (1 - g) * |y_xyz - t|^2 + wg * BCE(y_ghost, g), mean
// over rows; writes dY (bf16, padded) and its transpose, accumulates loss.
__global__ void k_loss(const float *Y, const float *T, int n, TrainBuf tb, float wg,
                       float *loss) {
  float part = 0.0f;
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < tb.B; m += gridDim.x * blockDim.x) {
    float d[N4];
#pragma unroll
    for (int c = 0; c < N4; ++c) d[c] = 0.0f;
    if (m < n) {
      const float g = T[(long)m * 4 + 3];
      const float inv = 1.0f / (float)n;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float e = Y[(long)m * NOut + c] - T[(long)m * 4 + c];
        part += (1.0f - g) * e * e * inv;
        d[c] = 2.0f * (1.0f - g) * e * inv;
      }
      const float y = Y[(long)m * NOut + 3];
      const float sp = fmaxf(y, 0.0f) + log1pf(expf(-fabsf(y)));  // log(1 + e^y)
      part += wg * (sp - g * y) * inv;
      d[3] = wg * (1.0f / (1.0f + expf(-y)) - g) * inv;
    }
#pragma unroll
    for (int c = 0; c < N4; ++c) {
      tb.dY[(long)m * N4 + c] = bf16(d[c]);
      tb.dYT[(long)c * tb.B + m] = bf16(d[c]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if (threadIdx.x % 64 == 0) atomicAdd(loss, part);
}

// dA = dZ_up . W^T for kRows rows (K_red = n_up), then dZ = dA * gelu'(Z)
// for a 256-wide layer; writes dZ [B][H] and dZT [H][B].
template <int NUP>
__global__ void __launch_bounds__(256) k_bwd_data(const bf16 *dZup, const bf16 *Wb /*[H][NUP]*/,
                                                  const float *Z, bf16 *dZ, bf16 *dZT, int B) {
  const int m0 = blockIdx.x * kRows;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int ncol = w * 64;
  f32x16 acc[2][2] = {};
  mma_block<2, 2>(acc, dZup + (long)m0 * NUP, NUP, kRows, Wb + (long)ncol * NUP, NUP, NUP);
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int row = m0 + r * 32 + acc_row(j), col = ncol + c * 32 + l % 32;
        const float g = acc[r][c][j] * gelu_grad(Z[(long)row * H + col]);
        dZ[(long)row * H + col] = bf16(g);
        dZT[(long)col * B + row] = bf16(g);
      }
}

// dW[k][n] += sum over a batch chunk of AT[k][m] * dZT[n][m]; one workgroup
// per (64-row block of k, batch chunk), all n (NC 32-wide column tiles).
template <int NC>
__global__ void __launch_bounds__(256) k_wgrad(const bf16 *AT, const bf16 *dZT, int B, int chunk,
                                               float *dW /*[K][NC*32]*/, float *db) {
  const int k0 = blockIdx.x * kRows;
  const int b0 = blockIdx.y * chunk;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  constexpr int N = NC * 32;
  // waves split the n tiles
  for (int ct = w; ct < NC; ct += 4) {
    f32x16 acc[2][1] = {};
    mma_block<2, 1>(acc, AT + (long)k0 * B + b0, B, kRows, dZT + (long)(ct * 32) * B + b0, B, chunk);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int k = k0 + r * 32 + acc_row(j), nn = ct * 32 + l % 32;
        atomicAdd(&dW[(long)k * N + nn], acc[r][0][j]);
      }
  }
  // bias gradient from the first k block only
  if (blockIdx.x == 0)
    for (int nn = threadIdx.x; nn < N; nn += 256) {
      float s = 0.0f;
      for (int m = b0; m < b0 + chunk; ++m) s += (float)dZT[(long)nn * B + m];
      atomicAdd(&db[nn], s);
    }
}

__global__ void k_zero(float *p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Adam on the FP32 master copy; padded input rows of W1 and padded output
// columns of W4/b4 stay zero.
__global__ void k_adam(float *P, float *Mo, float *Ve, const float *G, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    bool frozen = false;
    if (i < ob1) frozen = (i / H) >= kIn;
    else if (i >= oW4 && i < ob4) frozen = ((i - oW4) % N4) >= NOut;
    else if (i >= ob4) frozen = (i - ob4) >= NOut;
    if (frozen) continue;
    const float g = G[i];
    const float m = b1 * Mo[i] + (1.0f - b1) * g;
    const float v = b2 * Ve[i] + (1.0f - b2) * g * g;
    Mo[i] = m;
    Ve[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  }
}

// BF16 copies: WT [N][K] for the forward pass, Wb [K][N] for the backward one
__global__ void k_refresh(const float *P, bf16 *WT1, bf16 *WT2, bf16 *WT3, bf16 *WT4, bf16 *Wb2,
                          bf16 *Wb3, bf16 *Wb4) {
  const long t0 = blockIdx.x * (long)blockDim.x + threadIdx.x, st = (long)gridDim.x * blockDim.x;
  for (long i = t0; i < (long)K0 * H; i += st) {
    const int k = (int)(i / H), nn = (int)(i % H);
    WT1[(long)nn * K0 + k] = bf16(P[oW1 + i]);
  }
  for (long i = t0; i < (long)H * H; i += st) {
    const int k = (int)(i / H), nn = (int)(i % H);
    WT2[(long)nn * H + k] = bf16(P[oW2 + i]);
    WT3[(long)nn * H + k] = bf16(P[oW3 + i]);
    Wb2[i] = bf16(P[oW2 + i]);
    Wb3[i] = bf16(P[oW3 + i]);
  }
  for (long i = t0; i < (long)H * N4; i += st) {
    const int k = (int)(i / N4), nn = (int)(i % N4);
    WT4[(long)nn * H + k] = bf16(P[oW4 + i]);
    Wb4[i] = bf16(P[oW4 + i]);
  }
}

}  // namespace olpt_mlp_impl

using namespace olpt_mlp_impl;

struct olpt_mlp {
  int device = 0;
  hipStream_t st = nullptr;
  int max_batch = 0, B = 0;
  float *P = nullptr, *Mo = nullptr, *Ve = nullptr, *G = nullptr;
  bf16 *WT1 = nullptr, *WT2 = nullptr, *WT3 = nullptr, *WT4 = nullptr;
  bf16 *Wb2 = nullptr, *Wb3 = nullptr, *Wb4 = nullptr;
  TrainBuf tb{};
  float *Y = nullptr, *loss = nullptr;
  long step = 0;
  std::vector<void *> allocs;
};

namespace {
thread_local std::string g_mlp_err;

olpt_status mfail(olpt_status s, const std::string &m) {
  g_mlp_err = m;
  return s;
}

#define MLP_TRY(x)                                                                   \
  do {                                                                               \
    hipError_t _e = (x);                                                             \
    if (_e != hipSuccess) return mfail(OLPT_E_GPU, std::string(#x ": ") + hipGetErrorString(_e)); \
  } while (0)

template <typename T> olpt_status mal(olpt_mlp *m, T **p, size_t count) {
  hipError_t e = hipMalloc((void **)p, count * sizeof(T) + 16);
  if (e != hipSuccess) return mfail(OLPT_E_ALLOC, "hipMalloc failed in the MLP workspace");
  m->allocs.push_back(*p);
  return OLPT_OK;
}

Net net_of(const olpt_mlp *m) {
  Net n;
  n.WT1 = m->WT1; n.WT2 = m->WT2; n.WT3 = m->WT3; n.WT4 = m->WT4;
  n.Wb2 = m->Wb2; n.Wb3 = m->Wb3; n.Wb4 = m->Wb4;
  n.b1 = m->P + ob1; n.b2 = m->P + ob2; n.b3 = m->P + ob3; n.b4 = m->P + ob4;
  return n;
}

void refresh(olpt_mlp *m) {
  k_refresh<<<512, 256, 0, m->st>>>(m->P, m->WT1, m->WT2, m->WT3, m->WT4, m->Wb2, m->Wb3, m->Wb4);
}
}  // namespace

namespace olpt {
// Forward pass for rows counted on the device (inside an IPR run, on the
// caller's stream; inference only reads the weights).
void mlp_infer_device(const void *mlp, const void *X, const int32_t *d_n, int cap, float *Y,
                      hipStream_t st, int nf, long xs, long ns, long ys) {
  const olpt_mlp *m = static_cast<const olpt_mlp *>(mlp);
  int blocks = (cap + kRows - 1) / kRows;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  TrainBuf none{};
  k_fwd<false><<<dim3(blocks, 1, nf), 256, 0, st>>>(net_of(m), (const bf16 *)X, 0, d_n, Y, none,
                                                    xs, ns, ys);
}
}  // namespace olpt

extern "C" {

const char *olpt_mlp_last_error(void) { return g_mlp_err.c_str(); }

int32_t olpt_mlp_num_params(void) { return (int32_t)kParams; }

olpt_status olpt_mlp_create(int device, void *stream, int max_batch, olpt_mlp **out) {
  if (!out || max_batch < 1) return mfail(OLPT_E_INVALID_ARGUMENT, "bad argument");
  MLP_TRY(hipSetDevice(device));
  olpt_mlp *m = new olpt_mlp();
  m->device = device;
  m->st = (hipStream_t)stream;
  m->max_batch = max_batch;
  m->B = (max_batch + kRows - 1) / kRows * kRows;
  const int B = m->B;
  olpt_status s;
#define M_A(p, T, n) if ((s = mal(m, &(p), (size_t)(n)))) { olpt_mlp_destroy(m); return s; }
  M_A(m->P, float, kParams);
  M_A(m->Mo, float, kParams);
  M_A(m->Ve, float, kParams);
  M_A(m->G, float, kParams);
  M_A(m->WT1, bf16, (long)K0 * H);
  M_A(m->WT2, bf16, (long)H * H);
  M_A(m->WT3, bf16, (long)H * H);
  M_A(m->WT4, bf16, (long)H * N4);
  M_A(m->Wb2, bf16, (long)H * H);
  M_A(m->Wb3, bf16, (long)H * H);
  M_A(m->Wb4, bf16, (long)H * N4);
  M_A(m->tb.Z1, float, (long)B * H);
  M_A(m->tb.Z2, float, (long)B * H);
  M_A(m->tb.Z3, float, (long)B * H);
  M_A(m->tb.AT0, bf16, (long)K0 * B);
  M_A(m->tb.AT1, bf16, (long)H * B);
  M_A(m->tb.AT2, bf16, (long)H * B);
  M_A(m->tb.AT3, bf16, (long)H * B);
  M_A(m->tb.dZ, bf16, (long)B * H);
  M_A(m->tb.dZb, bf16, (long)B * H);
  M_A(m->tb.dZT, bf16, (long)H * B);
  M_A(m->tb.dY, bf16, (long)B * N4);
  M_A(m->tb.dYT, bf16, (long)N4 * B);
  M_A(m->Y, float, (long)B * NOut);
  M_A(m->loss, float, 4);
#undef M_A
  m->tb.grad = m->G;
  m->tb.B = B;
  k_zero<<<256, 256, 0, m->st>>>(m->P, kParams);
  k_zero<<<256, 256, 0, m->st>>>(m->Mo, kParams);
  k_zero<<<256, 256, 0, m->st>>>(m->Ve, kParams);
  refresh(m);
  MLP_TRY(hipGetLastError());
  *out = m;
  return OLPT_OK;
}

olpt_status olpt_mlp_destroy(olpt_mlp *m) {
  if (!m) return OLPT_OK;
  (void)hipSetDevice(m->device);
  if (m->st)
    (void)hipStreamSynchronize(m->st);
  else
    (void)hipDeviceSynchronize();
  for (void *p : m->allocs) (void)hipFree(p);
  delete m;
  return OLPT_OK;
}

// Parameters as one FP32 array in blob order (W1 [512][256], b1, W2
// [256][256], b2, W3, b3, W4 [256][32], b4 [32]); rows 484..511 of W1 and
// columns 4..31 of W4/b4 are padding and are kept at zero.
olpt_status olpt_mlp_set_params(olpt_mlp *m, const float *host, int32_t n) {
  if (!m || !host || n != kParams) return mfail(OLPT_E_INVALID_ARGUMENT, "parameter count mismatch");
  std::vector<float> h(host, host + kParams);
  for (long i = 0; i < kParams; ++i) {
    if (i < ob1 && (i / H) >= kIn) h[i] = 0.0f;
    if (i >= oW4 && i < ob4 && ((i - oW4) % N4) >= NOut) h[i] = 0.0f;
    if (i >= ob4 && (i - ob4) >= NOut) h[i] = 0.0f;
  }
  MLP_TRY(hipMemcpyAsync(m->P, h.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, m->st));
  k_zero<<<256, 256, 0, m->st>>>(m->Mo, kParams);
  k_zero<<<256, 256, 0, m->st>>>(m->Ve, kParams);
  refresh(m);
  m->step = 0;
  MLP_TRY(hipStreamSynchronize(m->st));
  return OLPT_OK;
}

olpt_status olpt_mlp_get_params(olpt_mlp *m, float *host, int32_t n) {
  if (!m || !host || n != kParams) return mfail(OLPT_E_INVALID_ARGUMENT, "parameter count mismatch");
  MLP_TRY(hipMemcpyAsync(host, m->P, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  return OLPT_OK;
}

// Gradient of the last olpt_mlp_train_step's loss, same layout as the
// parameters (a step with lr = 0 leaves the parameters unchanged).
olpt_status olpt_mlp_get_grads(olpt_mlp *m, float *host, int32_t n) {
  if (!m || !host || n != kParams) return mfail(OLPT_E_INVALID_ARGUMENT, "parameter count mismatch");
  MLP_TRY(hipMemcpyAsync(host, m->G, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  return OLPT_OK;
}

// Y [n][4] = MLP(X [n][512] bf16), device buffers.
olpt_status olpt_mlp_infer(olpt_mlp *m, const void *d_x, int32_t n, float *d_y) {
  if (!m || (n > 0 && (!d_x || !d_y))) return mfail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (n <= 0) return OLPT_OK;
  TrainBuf none{};
  k_fwd<false><<<(n + kRows - 1) / kRows, 256, 0, m->st>>>(net_of(m), (const bf16 *)d_x, n, nullptr,
                                                         d_y, none);
  MLP_TRY(hipGetLastError());
  return OLPT_OK;
}

// One Adam step on a batch: X [n][512] bf16, T [n][4] = (dx, dy, dz in
// voxels, ghost label 0/1).  *loss_out (host) receives the batch loss.
olpt_status olpt_mlp_train_step(olpt_mlp *m, const void *d_x, const float *d_t, int32_t n,
                                float lr, float ghost_weight, float *loss_out) {
  if (!m || !d_x || !d_t || n < 1) return mfail(OLPT_E_INVALID_ARGUMENT, "null argument");
  if (n > m->max_batch) return mfail(OLPT_E_CAPACITY, "batch larger than max_batch");
  hipStream_t st = m->st;
  TrainBuf tb = m->tb;
  const int B = m->B;
  const int nb = (n + kRows - 1) / kRows;
  k_zero<<<512, 256, 0, st>>>(m->G, kParams);
  k_zero<<<1, 64, 0, st>>>(m->loss, 4);
  k_fwd<true><<<nb, 256, 0, st>>>(net_of(m), (const bf16 *)d_x, n, nullptr, m->Y, tb);
  k_loss<<<256, 256, 0, st>>>(m->Y, d_t, n, tb, ghost_weight, m->loss);
  // The backward pass covers the nb*kRows rows the forward pass wrote (rows
  // past n hold zeros there); batch chunks of the weight gradient tile them
  // exactly.
  const int Bn = nb * kRows;
  const int chunk = Bn % 256 == 0 ? 256 : kRows;
  const int nchunk = Bn / chunk;
  // output layer: dW4 = A3^T dY, dZ3 = (dY W4^T) * gelu'(Z3)
  k_wgrad<1><<<dim3(H / kRows, nchunk), 256, 0, st>>>(tb.AT3, tb.dYT, B, chunk, m->G + oW4, m->G + ob4);
  k_bwd_data<N4><<<nb, 256, 0, st>>>(tb.dY, m->Wb4, tb.Z3, tb.dZ, tb.dZT, B);
  // layer 3
  k_wgrad<8><<<dim3(H / kRows, nchunk), 256, 0, st>>>(tb.AT2, tb.dZT, B, chunk, m->G + oW3, m->G + ob3);
  k_bwd_data<H><<<nb, 256, 0, st>>>(tb.dZ, m->Wb3, tb.Z2, tb.dZb, tb.dZT, B);
  // layer 2
  k_wgrad<8><<<dim3(H / kRows, nchunk), 256, 0, st>>>(tb.AT1, tb.dZT, B, chunk, m->G + oW2, m->G + ob2);
  k_bwd_data<H><<<nb, 256, 0, st>>>(tb.dZb, m->Wb2, tb.Z1, tb.dZ, tb.dZT, B);
  // layer 1
  k_wgrad<8><<<dim3(K0 / kRows, nchunk), 256, 0, st>>>(tb.AT0, tb.dZT, B, chunk, m->G + oW1, m->G + ob1);
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
  k_adam<<<512, 256, 0, st>>>(m->P, m->Mo, m->Ve, m->G, kParams, lr, b1, b2, 1e-8f, c1, c2);
  refresh(m);
  MLP_TRY(hipGetLastError());
  if (loss_out) {
    MLP_TRY(hipMemcpyAsync(loss_out, m->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    MLP_TRY(hipStreamSynchronize(st));
  }
  return OLPT_OK;
}

}  // extern "C"
