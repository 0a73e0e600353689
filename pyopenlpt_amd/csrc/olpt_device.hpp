// olpt_device.hpp - device-side geometry shared by the IPR kernels (gfx950).
//
// Camera projection / line of sight / triangulation / OTF interpolation /
// ROI bounds, written for the GPU: templated on the arithmetic type T
// (float for the fast path, double for the parity path) where the reference
// computes in FP64, and always double where the tier design keeps FP64
// (line-of-sight inversion of distortion, triangulation normal equations).
// The library is compiled with -ffp-contract=off so FP64 results follow the
// reference's rounding (the reference is built without FMA contraction).
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "olpt_types.h"
#include "olpt_gi3.hpp"

namespace olpt {

constexpr int kWave = 64;
constexpr int kMaxCam = OLPT_MAX_CAM;

// --------------------------------------------------------------------------
// Camera in the arithmetic type T (converted once on the host).
// --------------------------------------------------------------------------
// a * b + c.  FP32 fuses it (one instruction, one rounding); FP64 keeps the
// separately rounded product and sum of the reference's expressions (the
// build turns off contraction), so FP64 results do not change where it is
// used in place of the written-out form.
__host__ __device__ __forceinline__ float mad_t(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__host__ __device__ __forceinline__ double mad_t(double a, double b, double c) {
  return a * b + c;
}

template <typename T> struct CamT {
  T fx, fy, cx, cy;
  T k[12];
  T R[9];
  T Tv[3];
  T Rinv[9];
  T Tinv[3];
  int32_t n_dist, is_distorted, n_row, n_col;
};

template <typename T> CamT<T> make_cam(const olpt_camera &c) {
  CamT<T> o;
  o.fx = (T)c.fx; o.fy = (T)c.fy; o.cx = (T)c.cx; o.cy = (T)c.cy;
  for (int i = 0; i < 12; ++i) o.k[i] = (T)c.dist[i];
  for (int i = 0; i < 9; ++i) { o.R[i] = (T)c.R[i]; o.Rinv[i] = (T)c.Rinv[i]; }
  for (int i = 0; i < 3; ++i) { o.Tv[i] = (T)c.T[i]; o.Tinv[i] = (T)c.Tinv[i]; }
  o.n_dist = c.is_distorted ? c.n_dist : 0;
  o.is_distorted = c.is_distorted;
  o.n_row = c.n_row;
  o.n_col = c.n_col;
  return o;
}

// Camera::distort (Camera.cpp:913-966): normalized (x,y) -> pixel (u,v)
template <typename T>
__host__ __device__ __forceinline__ void distort(const CamT<T> &c, T x, T y, T &u, T &v) {
  if (c.is_distorted) {
    const T *k = c.k;
    const int nd = c.n_dist;
    T r2 = x * x + y * y;
    T r4 = r2 * r2;
    T r6 = r4 * r2;
    T a1 = T(2) * x * y;
    T a2 = r2 + T(2) * x * x;
    T a3 = r2 + T(2) * y * y;
    T cd = T(1) + k[0] * r2 + k[1] * r4;
    if (nd > 4) cd += k[4] * r6;
    T ic = T(1);
    if (nd > 5) ic = T(1) / (T(1) + k[5] * r2 + k[6] * r4 + k[7] * r6);
    T xd = x * cd * ic + k[2] * a1 + k[3] * a2;
    T yd = y * cd * ic + k[2] * a3 + k[3] * a1;
    if (nd > 8) {
      xd += k[8] * r2 + k[9] * r4;
      yd += k[10] * r2 + k[11] * r4;
    }
    x = xd;
    y = yd;
  }
  u = mad_t(x, c.fx, c.cx);
  v = mad_t(y, c.fy, c.cy);
}

// Camera::project, pinhole branch = worldToUndistImg (Camera.cpp:680-689)
// followed by distort.
template <typename T>
__host__ __device__ __forceinline__ void project(const CamT<T> &c, const T X[3], T &u,
                                                 T &v) {
  T t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    t[i] = mad_t(c.R[3 * i + 2], X[2], mad_t(c.R[3 * i + 1], X[1], c.R[3 * i] * X[0]));
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = t[i] + c.Tv[i];
  const T x = t[2] != T(0) ? t[0] / t[2] : t[0];
  const T y = t[2] != T(0) ? t[1] / t[2] : t[1];
  distort(c, x, y, u, v);
}

// Camera::undistort (Camera.cpp:1018-1094) + pinholeLine (:1096-1108).
// Returns the unit direction; the line passes through Tinv.  FP64 only.
__device__ inline void line_of_sight(const CamT<double> &c, double u,
                                     double v, double out[3]) {
  double x = (u - c.cx) / c.fx;
  double y = (v - c.cy) / c.fy;
  double ux = x, uy = y;
  if (c.is_distorted) {
    const double *k = c.k;
    const int nd = c.n_dist;
    const double x0 = x, y0 = y;
    double err = 1.7976931348623157e308;
    int it = 0;
    while (err > 1e-5 && it < 50) {
      double r2 = x * x + y * y;
      double a1 = 2 * x * y;
      double a2 = r2 + 2 * x * x;
      double a3 = r2 + 2 * y * y;
      double ic;
      if (nd == 4)
        ic = 1 / (1 + (k[1] * r2 + k[0]) * r2);
      else if (nd == 5)
        ic = 1 / (1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2);
      else
        ic = (1 + ((k[7] * r2 + k[6]) * r2 + k[5]) * r2) /
             (1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2);
      if (ic < 0) {
        ux = (u - c.cx) / c.fx;
        uy = (v - c.cy) / c.fy;
        break;
      }
      double dX = k[2] * a1 + k[3] * a2;
      double dY = k[2] * a3 + k[3] * a1;
      if (nd > 8) {
        dX += k[8] * r2 + k[9] * r2 * r2;
        dY += k[10] * r2 + k[11] * r2 * r2;
      }
      x = (x0 - dX) * ic;
      y = (y0 - dY) * ic;
      ux = x;
      uy = y;
      double pu, pv;
      distort(c, ux, uy, pu, pv);
      const double du = pu - u, dv = pv - v;
      err = sqrt(du * du + dv * dv);
      ++it;
    }
  }
  double p[3] = {ux, uy, 1.0}, w[3], d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    w[i] = c.Rinv[3 * i] * p[0] + c.Rinv[3 * i + 1] * p[1] + c.Rinv[3 * i + 2] * p[2];
#pragma unroll
  for (int i = 0; i < 3; ++i) w[i] = w[i] + c.Tinv[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) d[i] = w[i] - c.Tinv[i];
  const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = d[i] / n;
}

// myMATH::gaussInverse for 3x3 (myMATH.h:183-289), incl. the signed row
// maximum used for pivot scaling: gauss_inverse3 in olpt_gi3.hpp.

// myMATH::triangulation (myMATH.cpp:357-419) over n lines through pt[i]
// with unit direction u[i].  Returns false where the reference throws.
// Fixed line count: every loop unrolls, so the lines stay in registers.
template <int N>
__host__ __device__ __forceinline__ bool triangulate_n(const double (*pt)[3], const double (*u)[3],
                                              double X[3], double &err) {
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, P[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double nx = u[i][0], ny = u[i][1], nz = u[i][2];
    const double t[9] = {1 - nx * nx, -nx * ny, -nx * nz, -nx * ny, 1 - ny * ny,
                         -ny * nz, -nx * nz, -ny * nz, 1 - nz * nz};
#pragma unroll
    for (int r = 0; r < 3; ++r)
      P[r] += t[3 * r] * pt[i][0] + t[3 * r + 1] * pt[i][1] + t[3 * r + 2] * pt[i][2];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] += t[k];
  }
  double inv[9];
  gauss_inverse3(M, inv);
#pragma unroll
  for (int r = 0; r < 3; ++r) X[r] = inv[3 * r] * P[0] + inv[3 * r + 1] * P[1] + inv[3 * r + 2] * P[2];
  double e = 0.0;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double d0 = X[0] - pt[i][0], d1 = X[1] - pt[i][1], d2 = X[2] - pt[i][2];
    const double nr = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double dt = d0 * u[i][0] + d1 * u[i][1] + d2 * u[i][2];
    double h = nr * nr - dt * dt;
    if (h < 0.0 && h >= -1e-8)
      h = 0.0;
    else if (h < -1e-8)
      ok = false;  // the reference throws here
    else
      h = sqrt(h);
    e += h * h;
  }
  err = sqrt(e / N);
  return ok;
}

// The decision triangulate_n<N>(pt, u, X, err) && err <= tq, taken without
// the Gauss inverse where the outcome is certain: 1 (true), 0 (false), or -1
// when it has to be computed exactly.  The least-squares point comes from
// the adjugate of the same normal matrix, and the squared residual sum
// E = sum |(X - pt_i) x u_i|^2 from cross products.  triangulate_n's err
// is sqrt(sum h_i^2 / N) with h_i^2 = |d_i|^2 - (d_i . u_i)^2 (d_i = X -
// pt_i), whose rounding is at most 13 eps |d_i|^2 per line (the
// cancellation of two squares of size |d_i|^2); the two least-squares
// points differ by rounding only, which moves E in second order.  So the
// decision is certain when E and N tq^2 differ by more than
// N * 16 eps max|d_i|^2 plus 1e-6 N tq^2 (the final division, square roots
// and comparison).  Taken only where the normal matrix is well conditioned
// (every pair of lines at least 1 degree apart: max_dot below cos 1 deg)
// and |d_i| < 1000 (then h_i^2 >= -13 eps |d_i|^2 > -1e-8, the bound below
// which triangulate_n fails).
template <int N>
__host__ __device__ __forceinline__ int tri_decide(const double (*pt)[3], const double (*u)[3],
                                          double tq, double max_dot) {
  constexpr double kCos1Deg = 0.99984769515639124;
  if (!(max_dot < kCos1Deg)) return -1;
  double m00 = 0, m01 = 0, m02 = 0, m11 = 0, m12 = 0, m22 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double nx = u[i][0], ny = u[i][1], nz = u[i][2];
    const double t00 = 1 - nx * nx, t01 = -nx * ny, t02 = -nx * nz;
    const double t11 = 1 - ny * ny, t12 = -ny * nz, t22 = 1 - nz * nz;
    m00 += t00;
    m01 += t01;
    m02 += t02;
    m11 += t11;
    m12 += t12;
    m22 += t22;
    b0 += t00 * pt[i][0] + t01 * pt[i][1] + t02 * pt[i][2];
    b1 += t01 * pt[i][0] + t11 * pt[i][1] + t12 * pt[i][2];
    b2 += t02 * pt[i][0] + t12 * pt[i][1] + t22 * pt[i][2];
  }
  const double c00 = m11 * m22 - m12 * m12, c01 = m02 * m12 - m01 * m22,
               c02 = m01 * m12 - m02 * m11, c11 = m00 * m22 - m02 * m02,
               c12 = m01 * m02 - m00 * m12, c22 = m00 * m11 - m01 * m01;
  const double det = m00 * c00 + m01 * c01 + m02 * c02;
  if (!(det > 1e-12)) return -1;
  const double inv = 1.0 / det;
  const double X0 = (c00 * b0 + c01 * b1 + c02 * b2) * inv;
  const double X1 = (c01 * b0 + c11 * b1 + c12 * b2) * inv;
  const double X2 = (c02 * b0 + c12 * b1 + c22 * b2) * inv;
  double E = 0.0, D2 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double d0 = X0 - pt[i][0], d1 = X1 - pt[i][1], d2 = X2 - pt[i][2];
    const double q0 = d1 * u[i][2] - d2 * u[i][1];
    const double q1 = d2 * u[i][0] - d0 * u[i][2];
    const double q2 = d0 * u[i][1] - d1 * u[i][0];
    E += q0 * q0 + q1 * q1 + q2 * q2;
    D2 = fmax(D2, d0 * d0 + d1 * d1 + d2 * d2);
  }
  if (!(D2 < 1e6) || !(E == E)) return -1;
  const double T = (double)N * tq * tq;
  const double delta = (double)N * 16.0 * 2.220446049250313e-16 * D2 + 1e-6 * T;
  if (E + delta < T) return 1;
  if (E - delta > T) return 0;
  return -1;
}

// triangulate_n<n> for a runtime n <= MC on arrays of MC rows (rows past MC
// are never named, so register arrays of MC rows stay in registers)
template <int MC>
__device__ __forceinline__ bool triangulate_upto(const double (*pt)[3], const double (*u)[3],
                                                 int n, double X[3], double &err) {
  if (n == 2) return triangulate_n<2>(pt, u, X, err);
  if constexpr (MC >= 3) if (n == 3) return triangulate_n<3>(pt, u, X, err);
  if constexpr (MC >= 4) if (n == 4) return triangulate_n<4>(pt, u, X, err);
  if constexpr (MC >= 5) if (n == 5) return triangulate_n<5>(pt, u, X, err);
  if constexpr (MC >= 6) if (n == 6) return triangulate_n<6>(pt, u, X, err);
  if constexpr (MC >= 7) if (n == 7) return triangulate_n<7>(pt, u, X, err);
  if constexpr (MC >= 8) if (n == 8) return triangulate_n<8>(pt, u, X, err);
  return false;
}

__device__ __forceinline__ bool triangulate(const double (*pt)[3], const double (*u)[3],
                                            int n, double X[3], double &err) {
  switch (n) {
  case 2: return triangulate_n<2>(pt, u, X, err);
  case 3: return triangulate_n<3>(pt, u, X, err);
  case 4: return triangulate_n<4>(pt, u, X, err);
  case 5: return triangulate_n<5>(pt, u, X, err);
  case 6: return triangulate_n<6>(pt, u, X, err);
  case 7: return triangulate_n<7>(pt, u, X, err);
  case 8: return triangulate_n<8>(pt, u, X, err);
  default: return false;
  }
}

// --------------------------------------------------------------------------
// OTF (OTF.cpp:273-338 + myMATH.cpp:131-214), arithmetic type T.
// --------------------------------------------------------------------------
template <typename T> struct OtfT {
  int32_t nx, ny, nz, ng;
  T bnd[6];
  // grid spacing (bnd[2a+1] - bnd[2a]) / (n_a - 1), rounded in T exactly as
  // the device would, and its reciprocal (FP32 lookups only)
  T step[3], inv_step[3];
  // FP32 only: the grid holds one (a, b, c, alpha) per camera; uni is
  // [n_cam][5] = (a, b, c, cos alpha, sin alpha)
  int32_t uniform;
  const T *uni;
  const T *a, *b, *c, *alpha; // [n_cam][ng]
};

// myMATH::triLinearInterp (myMATH.cpp:131-214) in two parts: the clamped
// cell fractions of p (shared by the four OTF parameters), then the
// interpolation of one parameter's 8 corner values.
template <typename T>
__device__ __forceinline__ void trilin_frac(const T lim[6], const T p[3], T d[3]) {
  T xd = (p[0] - lim[0]) / (lim[1] - lim[0]);
  T yd = (p[1] - lim[2]) / (lim[3] - lim[2]);
  T zd = (p[2] - lim[4]) / (lim[5] - lim[4]);
  d[0] = fmax(T(0), fmin(T(1), xd));
  d[1] = fmax(T(0), fmin(T(1), yd));
  d[2] = fmax(T(0), fmin(T(1), zd));
}

template <typename T>
__device__ __forceinline__ T trilin(const T v[8], const T d[3]) {
  const T xd = d[0], yd = d[1], zd = d[2];
  const T c00 = mad_t(v[0], 1 - xd, v[1] * xd);
  const T c01 = mad_t(v[3], 1 - xd, v[2] * xd);
  const T c10 = mad_t(v[4], 1 - xd, v[5] * xd);
  const T c11 = mad_t(v[7], 1 - xd, v[6] * xd);
  const T c0 = mad_t(c00, 1 - yd, c10 * yd);
  const T c1 = mad_t(c01, 1 - yd, c11 * yd);
  return mad_t(c0, 1 - zd, c1 * zd);
}

// -> {a, b, c, cos(alpha), sin(alpha)}
template <typename T>
__device__ inline void otf_param(const OtfT<T> &o, int cam, const T X[3], T out[5]) {
  if constexpr (sizeof(T) == 4) {
    // a uniform grid interpolates to its node values wherever X lies
    if (o.uniform) {
      const T *u = o.uni + 5 * cam;
#pragma unroll
      for (int q = 0; q < 5; ++q) out[q] = u[q];
      return;
    }
  }
  const T *B = o.bnd;
  const T px = fmin(fmax(X[0], B[0]), B[1]);
  const T py = fmin(fmax(X[1], B[2]), B[3]);
  const T pz = fmin(fmax(X[2], B[4]), B[5]);
  const T gx = o.step[0], gy = o.step[1], gz = o.step[2];
  int ix, iy, iz;
  if constexpr (sizeof(T) == 4) {
    // FP32: products with the reciprocal spacing (the cell index can differ
    // from the division's only for a point within rounding of a cell face,
    // where the interpolation is continuous)
    ix = (int)floor((px - B[0]) * o.inv_step[0]);
    iy = (int)floor((py - B[2]) * o.inv_step[1]);
    iz = (int)floor((pz - B[4]) * o.inv_step[2]);
  } else {
    ix = (int)floor((px - B[0]) / gx);
    iy = (int)floor((py - B[2]) / gy);
    iz = (int)floor((pz - B[4]) / gz);
  }
  ix = ix < 0 ? 0 : (ix > o.nx - 2 ? o.nx - 2 : ix);
  iy = iy < 0 ? 0 : (iy > o.ny - 2 ? o.ny - 2 : iy);
  iz = iz < 0 ? 0 : (iz > o.nz - 2 ? o.nz - 2 : iz);
  const T lim[6] = {B[0] + gx * ix, B[0] + gx * (ix + 1), B[2] + gy * iy,
                    B[2] + gy * (iy + 1), B[4] + gz * iz, B[4] + gz * (iz + 1)};
  const int nyz = o.ny * o.nz;
  const int g[8] = {ix * nyz + iy * o.nz + iz,           (ix + 1) * nyz + iy * o.nz + iz,
                    (ix + 1) * nyz + iy * o.nz + iz + 1, ix * nyz + iy * o.nz + iz + 1,
                    ix * nyz + (iy + 1) * o.nz + iz,     (ix + 1) * nyz + (iy + 1) * o.nz + iz,
                    (ix + 1) * nyz + (iy + 1) * o.nz + iz + 1, ix * nyz + (iy + 1) * o.nz + iz + 1};
  const T p[3] = {px, py, pz};
  T va[8], vb[8], vc[8], vl[8];
  const int base = cam * o.ng;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    va[k] = o.a[base + g[k]];
    vb[k] = o.b[base + g[k]];
    vc[k] = o.c[base + g[k]];
    vl[k] = o.alpha[base + g[k]];
  }
  T d[3];
  if constexpr (sizeof(T) == 4) {
    // FP32: the cell edge taken as the grid spacing
#pragma unroll
    for (int a = 0; a < 3; ++a)
      d[a] = fmax(T(0), fmin(T(1), (p[a] - lim[2 * a]) * o.inv_step[a]));
  } else {
    trilin_frac(lim, p, d);
  }
  out[0] = trilin(va, d);
  out[1] = trilin(vb, d);
  out[2] = trilin(vc, d);
  const T al = trilin(vl, d);
  // an unrotated OTF (alpha 0 everywhere) skips the trigonometry:
  // cos(0) = 1 and sin(0) = 0 exactly
  if (__builtin_expect(al == T(0), 1)) {
    out[3] = T(1);
    out[4] = T(0);
  } else {
    out[3] = cos(al);
    out[4] = sin(al);
  }
}

// Gaussian OTF intensity at pixel (col,row) for a particle at (u,v)
// (TracerShakeStrategy::gaussIntensity, Shake.cpp:901-914).
// exp of the Gaussian intensity: FP64 keeps the correctly rounded exp (the
// reference's); FP32 uses the hardware exp2 (v_exp_f32) on x log2(e), whose
// relative error (~1e-6 over the ROI's exponent range) is far below the
// FP32 loop's own residual rounding.
__device__ __forceinline__ double gexp(double x) { return exp(x); }
__device__ __forceinline__ float gexp(float x) { return __expf(x); }

template <typename T>
__device__ __forceinline__ T gauss_int(const T P[5], T u, T v, int row, int col) {
  const T dx = T(col) - u, dy = T(row) - v;
  const T xx = mad_t(dx, P[3], dy * P[4]);
  const T yy = mad_t(dy, P[3], -dx * P[4]);
  const T val = P[0] * gexp(-mad_t(P[1] * xx, xx, P[2] * yy * yy));
  return val > T(0) ? val : T(0);
}

// Shake::calROIBound (Shake.cpp:246-277).  [rmin,rmax) x [cmin,cmax);
// empty is {1,0,1,0}.
struct PRange {
  int rmin, rmax, cmin, cmax;
  __device__ bool empty() const { return rmax < rmin || cmax < cmin; }
};

// The box size is decided on the host, in the reference's double
// arithmetic: width = round(2 dx), height = round(2 dy) (box_width below);
// a negative width or height stands for dx <= 0 or dy <= 0 (empty).
template <typename T>
__device__ __forceinline__ PRange roi_bound(int H, int W, T cx, T cy, int width, int height) {
  PRange e{1, 0, 1, 0};
  if (H <= 0 || W <= 0 || width < 0 || height < 0 || cx <= T(0) || cy <= T(0) ||
      cx >= T(W) || cy >= T(H))
    return e;
  int cmin = (int)round(cx - T(width - 1) / T(2));
  int rmin = (int)round(cy - T(height - 1) / T(2));
  int cmax = cmin + width, rmax = rmin + height;
  rmin = rmin > 0 ? rmin : 0;
  cmin = cmin > 0 ? cmin : 0;
  rmax = rmax < H ? rmax : H;
  cmax = cmax < W ? cmax : W;
  if (rmax < rmin || cmax < cmin) return e;
  return PRange{rmin, rmax, cmin, cmax};
}

// calROIBound's box size for half-size d (Shake.cpp:251-257): round(2 d)
// with C++ std::round (half away from zero), -1 when d <= 0.
inline int box_width(double d) { return d > 0.0 ? (int)std::round(2.0 * d) : -1; }

// --------------------------------------------------------------------------
// Wavefront helpers (wave64).
// --------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits below this lane
__device__ __forceinline__ int mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

} // namespace olpt
