/*
 * olpt_oracle.c - CPU restatement of the reference IPR hot path (FP64) and
 * of the tracking functions predictNext, checkLinearFit and the predictive field.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against; it is never linked into the product.  Each function cites the
 * reference file:line it restates; arithmetic is written in the reference's
 * operation order (the reference is built without FMA contraction, and so is
 * this file: -ffp-contract=off).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   camera project / LOS      reference KAT files test/solutions/test_Camera
 *                             + the compiled reference (oracle/_ref)
 *   triangulation, polyfit    closed forms, reference KATs, compiled reference
 *   OTF trilinear             constant-grid closed form, compiled reference
 *   peaks                     brute force on tiny images, compiled reference
 *   stereo match / prune      compiled reference on seeded frames
 *   shake / flags / residual  compiled reference on seeded frames
 *   runIPR                    compiled reference, per-loop counts
 *   STB predictNext,          compiled reference (predictNext / checkLinearFit
 *   checkLinearFit,           / calPredField golden vectors, bit-exact), the
 *   predictive field          one-step LMS closed form, exact recovery of
 *                             isolated displacements
 * Exact duplicates at zero distance in the repeat search are resolved by
 * nanoflann's unstable distance sort in the reference; here every neighbour
 * other than the query itself is marked ("parity unpinned" for that tie).
 */
#include "olpt_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define SMALLNUMBER 1e-8     /* STBCommons.h:10 */
#define LOGSMALLNUMBER 1e-4  /* STBCommons.h:16 */
#define UNDISTORT_MAX_ITER 50 /* STBCommons.h:19 */
#define UNDISTORT_EPS 1e-5   /* STBCommons.h:20 */

enum { OK = 0, ERR_CAP = 1, ERR_ARG = 2, ERR_UNSUP = 3, ERR_ALLOC = 4 };

/* ------------------------------------------------------------------------ */
/* growable arrays                                                          */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t *v; int n, cap; } ivec;
typedef struct { double *v; int n, cap; } dvec;

static int ivec_push(ivec *a, int32_t x) {
  if (a->n == a->cap) {
    int nc = a->cap ? 2 * a->cap : 16;
    int32_t *nv = (int32_t *)realloc(a->v, sizeof(int32_t) * (size_t)nc);
    if (!nv) return ERR_ALLOC;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  return OK;
}
static int dvec_push(dvec *a, double x) {
  if (a->n == a->cap) {
    int nc = a->cap ? 2 * a->cap : 16;
    double *nv = (double *)realloc(a->v, sizeof(double) * (size_t)nc);
    if (!nv) return ERR_ALLOC;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  return OK;
}

/* ------------------------------------------------------------------------ */
/* Camera model: Camera.cpp:648-689, 913-966, 996-1108                      */
/* ------------------------------------------------------------------------ */
typedef struct { double pt[3], u[3]; } line3;
typedef struct { double pt[2], u[2]; } line2;

/* Camera::worldToUndistImg (Camera.cpp:680-689): temp = R*X + T */
static void world_to_undist(const olpt_camera *c, const double X[3],
                            double *x, double *y) {
  double t[3];
  for (int i = 0; i < 3; ++i)
    t[i] = c->R[3 * i] * X[0] + c->R[3 * i + 1] * X[1] + c->R[3 * i + 2] * X[2];
  for (int i = 0; i < 3; ++i) t[i] = t[i] + c->T[i];
  *x = t[2] != 0.0 ? t[0] / t[2] : t[0];
  *y = t[2] != 0.0 ? t[1] / t[2] : t[1];
}

/* Camera::distort (Camera.cpp:913-966) incl. the intrinsic mapping */
static void distort(const olpt_camera *c, double x, double y, double *u,
                    double *v) {
  double xd, yd;
  if (c->is_distorted) {
    const double *k = c->dist;
    const int nd = c->n_dist;
    double r2 = x * x + y * y;
    double r4 = r2 * r2;
    double r6 = r4 * r2;
    double a1 = 2 * x * y;
    double a2 = r2 + 2 * x * x;
    double a3 = r2 + 2 * y * y;
    double cdist = 1 + k[0] * r2 + k[1] * r4;
    if (nd > 4) cdist += k[4] * r6;
    double icdist2 = 1.0;
    if (nd > 5) icdist2 = 1.0 / (1 + k[5] * r2 + k[6] * r4 + k[7] * r6);
    xd = x * cdist * icdist2 + k[2] * a1 + k[3] * a2;
    yd = y * cdist * icdist2 + k[2] * a3 + k[3] * a1;
    if (nd > 8) {
      xd += k[8] * r2 + k[9] * r4;
      yd += k[10] * r2 + k[11] * r4;
    }
  } else {
    xd = x;
    yd = y;
  }
  *u = xd * c->fx + c->cx;
  *v = yd * c->fy + c->cy;
}

/* Camera::project, pinhole branch (Camera.cpp:648-652) */
static void project(const olpt_camera *c, const double X[3], double uv[2]) {
  double x, y;
  world_to_undist(c, X, &x, &y);
  distort(c, x, y, &uv[0], &uv[1]);
}

/* Camera::undistort (Camera.cpp:1018-1094): fixed-point iteration */
static void undistort(const olpt_camera *c, double u, double v, double *xo,
                      double *yo) {
  const double fx = c->fx, fy = c->fy, cx = c->cx, cy = c->cy;
  double x = (u - cx) / fx;
  double y = (v - cy) / fy;
  double ux = x, uy = y;
  if (c->is_distorted) {
    const double *k = c->dist;
    const int nd = c->n_dist;
    double x0 = x, y0 = y;
    double error = 1.7976931348623157e308;
    int iter = 0;
    while (error > UNDISTORT_EPS && iter < UNDISTORT_MAX_ITER) {
      double r2 = x * x + y * y;
      double a1 = 2 * x * y;
      double a2 = r2 + 2 * x * x;
      double a3 = r2 + 2 * y * y;
      double icdist;
      if (nd == 4)
        icdist = 1 / (1 + (k[1] * r2 + k[0]) * r2);
      else if (nd == 5)
        icdist = 1 / (1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2);
      else
        icdist = (1 + ((k[7] * r2 + k[6]) * r2 + k[5]) * r2) /
                 (1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2);
      if (icdist < 0) {
        ux = (u - cx) / fx;
        uy = (v - cy) / fy;
        break;
      }
      double deltaX = k[2] * a1 + k[3] * a2;
      double deltaY = k[2] * a3 + k[3] * a1;
      if (nd > 8) {
        deltaX += k[8] * r2 + k[9] * r2 * r2;
        deltaY += k[10] * r2 + k[11] * r2 * r2;
      }
      x = (x0 - deltaX) * icdist;
      y = (y0 - deltaY) * icdist;
      ux = x;
      uy = y;
      double du, dv;
      distort(c, ux, uy, &du, &dv);
      error = sqrt((du - u) * (du - u) + (dv - v) * (dv - v));
      iter++;
    }
  }
  *xo = ux;
  *yo = uy;
}

/* Camera::pinholeLine (Camera.cpp:1096-1108) + createUnitVector
 * (myMATH.cpp:218-223) */
static void pinhole_line(const olpt_camera *c, double x, double y, line3 *L) {
  double p[3] = {x, y, 1.0}, w[3], d[3];
  for (int i = 0; i < 3; ++i)
    w[i] = c->Rinv[3 * i] * p[0] + c->Rinv[3 * i + 1] * p[1] +
           c->Rinv[3 * i + 2] * p[2];
  for (int i = 0; i < 3; ++i) w[i] = w[i] + c->Tinv[i];
  for (int i = 0; i < 3; ++i) d[i] = w[i] - c->Tinv[i];
  double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int i = 0; i < 3; ++i) {
    L->pt[i] = c->Tinv[i];
    L->u[i] = d[i] / nrm;
  }
}

/* Camera::lineOfSight, pinhole branch (Camera.cpp:996-1001) */
static void line_of_sight(const olpt_camera *c, double u, double v,
                          line3 *L) {
  double x, y;
  undistort(c, u, v, &x, &y);
  pinhole_line(c, x, y, L);
}

/* ------------------------------------------------------------------------ */
/* Linear algebra: myMATH.h:183-289 (gaussForward/Backward/Inverse)         */
/* ------------------------------------------------------------------------ */
static void gauss_inverse3(const double A[9], double out[9]) {
  double u[9];
  int idx[3];
  double maxrow[3];
  memcpy(u, A, sizeof(u));
  for (int i = 0; i < 3; ++i) {
    idx[i] = i;
    double m = u[3 * i];
    for (int j = 1; j < 3; ++j)
      if (u[3 * i + j] > m) m = u[3 * i + j];
    maxrow[i] = m; /* std::max_element: signed maximum (myMATH.h:200) */
  }
  for (int y = 0; y < 2; ++y) {
    double maxp = 0;
    int t = 0;
    for (int x = y; x < 3; ++x) {
      double pv = fabs(u[3 * idx[x] + y] / maxrow[idx[x]]);
      if (pv > maxp) {
        maxp = pv;
        t = x;
      }
    }
    int tmp = idx[t];
    idx[t] = idx[y];
    idx[y] = tmp;
    for (int x = y + 1; x < 3; ++x) {
      double m = u[3 * idx[x] + y] / u[3 * idx[y] + y];
      u[3 * idx[x] + y] = m;
      for (int it = y + 1; it < 3; ++it)
        u[3 * idx[x] + it] -= m * u[3 * idx[y] + it];
    }
  }
  double b[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  memcpy(out, b, sizeof(b));
  for (int col = 0; col < 3; ++col) {
    for (int y = 0; y < 2; ++y)
      for (int x = y + 1; x < 3; ++x)
        b[3 * idx[x] + col] -= u[3 * idx[x] + y] * b[3 * idx[y] + col];
    out[3 * 2 + col] = b[3 * idx[2] + col] / u[3 * idx[2] + 2];
    for (int x = 1; x > -1; --x) {
      double z = b[3 * idx[x] + col];
      for (int y = x + 1; y < 3; ++y) z -= u[3 * idx[x] + y] * out[3 * y + col];
      out[3 * x + col] = z / u[3 * idx[x] + x];
    }
  }
}

/* myMATH::gaussInverse for n x n (myMATH.h:183-289), row-major in/out */
static void gauss_inverse_n(const double *A, int n, double *out) {
  double *u = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *b = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *mr = (double *)malloc(sizeof(double) * (size_t)n);
  int *idx = (int *)malloc(sizeof(int) * (size_t)n);
  memcpy(u, A, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; ++i) {
    idx[i] = i;
    double m = u[(size_t)i * n];
    for (int j = 1; j < n; ++j)
      if (u[(size_t)i * n + j] > m) m = u[(size_t)i * n + j];
    mr[i] = m;
  }
  for (int y = 0; y < n - 1; ++y) {
    double maxp = 0;
    int t = 0;
    for (int x = y; x < n; ++x) {
      double pv = fabs(u[(size_t)idx[x] * n + y] / mr[idx[x]]);
      if (pv > maxp) {
        maxp = pv;
        t = x;
      }
    }
    int tmp = idx[t];
    idx[t] = idx[y];
    idx[y] = tmp;
    for (int x = y + 1; x < n; ++x) {
      double m = u[(size_t)idx[x] * n + y] / u[(size_t)idx[y] * n + y];
      u[(size_t)idx[x] * n + y] = m;
      for (int it = y + 1; it < n; ++it)
        u[(size_t)idx[x] * n + it] -= m * u[(size_t)idx[y] * n + it];
    }
  }
  for (int i = 0; i < n * n; ++i) b[i] = (i % (n + 1) == 0) ? 1.0 : 0.0;
  memcpy(out, b, sizeof(double) * (size_t)n * n);
  for (int col = 0; col < n; ++col) {
    for (int y = 0; y < n - 1; ++y)
      for (int x = y + 1; x < n; ++x)
        b[(size_t)idx[x] * n + col] -= u[(size_t)idx[x] * n + y] * b[(size_t)idx[y] * n + col];
    out[(size_t)(n - 1) * n + col] = b[(size_t)idx[n - 1] * n + col] / u[(size_t)idx[n - 1] * n + n - 1];
    for (int x = n - 2; x > -1; --x) {
      double z = b[(size_t)idx[x] * n + col];
      for (int y = x + 1; y < n; ++y) z -= u[(size_t)idx[x] * n + y] * out[(size_t)y * n + col];
      out[(size_t)x * n + col] = z / u[(size_t)idx[x] * n + x];
    }
  }
  free(u);
  free(b);
  free(mr);
  free(idx);
}

/* myMATH::polyfit of any order (myMATH.cpp:490-534) */
static void polyfit_n(const double *x, const double *y, int n, int order, double *c) {
  const int m = order + 1;
  double *X = (double *)malloc(sizeof(double) * (size_t)n * m);
  double *XtX = (double *)malloc(sizeof(double) * (size_t)m * m);
  double *inv = (double *)malloc(sizeof(double) * (size_t)m * m);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) X[(size_t)i * m + j] = pow(x[i], (double)j);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += X[(size_t)i * m + a] * X[(size_t)i * m + b];
      XtX[a * m + b] = s;
    }
  gauss_inverse_n(XtX, m, inv);
  for (int a = 0; a < m; ++a) {
    double s = 0;
    for (int i = 0; i < n; ++i) {
      double t = 0;
      for (int k = 0; k < m; ++k) t += inv[a * m + k] * X[(size_t)i * m + k];
      s += t * y[i];
    }
    c[a] = s;
  }
  free(X);
  free(XtX);
  free(inv);
}

/* myMATH::triangulation (myMATH.cpp:357-419).  Returns -1 where the
 * reference throws (h < -1e-8). */
static int triangulate(const line3 *const *L, int n, double X[3],
                       double *err) {
  double M[9] = {0}, P[3] = {0};
  for (int i = 0; i < n; ++i) {
    const double *p = L[i]->pt, *u = L[i]->u;
    double nx = u[0], ny = u[1], nz = u[2];
    double t[9] = {1 - nx * nx, -nx * ny, -nx * nz,
                   -nx * ny, 1 - ny * ny, -ny * nz,
                   -nx * nz, -ny * nz, 1 - nz * nz};
    for (int r = 0; r < 3; ++r)
      P[r] += t[3 * r] * p[0] + t[3 * r + 1] * p[1] + t[3 * r + 2] * p[2];
    for (int k = 0; k < 9; ++k) M[k] += t[k];
  }
  double inv[9];
  gauss_inverse3(M, inv);
  for (int r = 0; r < 3; ++r)
    X[r] = inv[3 * r] * P[0] + inv[3 * r + 1] * P[1] + inv[3 * r + 2] * P[2];
  double e = 0.0;
  for (int i = 0; i < n; ++i) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = X[k] - L[i]->pt[k];
    double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double dt = d[0] * L[i]->u[0] + d[1] * L[i]->u[1] + d[2] * L[i]->u[2];
    double h = nrm * nrm - dt * dt;
    if (h < 0.0 && h >= -SMALLNUMBER)
      h = 0.0;
    else if (h < -SMALLNUMBER)
      return -1;
    else
      h = sqrt(h);
    e += h * h;
  }
  *err = sqrt(e / n);
  return 0;
}

/* myMATH::polyfit (myMATH.cpp:490-534): a = (inv(X'X) X') y */
static void polyfit2(const double *x, const double *y, int n, double c[3]) {
  const int m = 3;
  double *X = (double *)malloc(sizeof(double) * (size_t)n * m);
  double XtX[9], inv[9];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) X[i * m + j] = pow(x[i], (double)j);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += X[i * m + a] * X[i * m + b];
      XtX[a * m + b] = s;
    }
  gauss_inverse3(XtX, inv);
  for (int a = 0; a < m; ++a) {
    double s = 0;
    for (int i = 0; i < n; ++i) {
      double t = 0;
      for (int k = 0; k < m; ++k) t += inv[a * m + k] * X[i * m + k];
      s += t * y[i];
    }
    c[a] = s;
  }
  free(X);
}

/* ------------------------------------------------------------------------ */
/* OTF: OTF.cpp:145-157 (setGrid), :273-338; myMATH.cpp:9-30, :131-214     */
/* ------------------------------------------------------------------------ */
static double linspace_at(double lo, double hi, int n, int i) {
  double delta = (hi - lo) / (n - 1);
  return lo + delta * i;
}

static double trilinear(const double lim[6], const double v[8],
                        const double p[3]) {
  double xd = (p[0] - lim[0]) / (lim[1] - lim[0]);
  double yd = (p[1] - lim[2]) / (lim[3] - lim[2]);
  double zd = (p[2] - lim[4]) / (lim[5] - lim[4]);
  xd = fmax(0.0, fmin(1.0, xd));
  yd = fmax(0.0, fmin(1.0, yd));
  zd = fmax(0.0, fmin(1.0, zd));
  double c00 = v[0] * (1 - xd) + v[1] * xd;
  double c01 = v[3] * (1 - xd) + v[2] * xd;
  double c10 = v[4] * (1 - xd) + v[5] * xd;
  double c11 = v[7] * (1 - xd) + v[6] * xd;
  double c0 = c00 * (1 - yd) + c10 * yd;
  double c1 = c01 * (1 - yd) + c11 * yd;
  return c0 * (1 - zd) + c1 * zd;
}

/* OTF::getOTFParam -> {a, b, c, cos(alpha), sin(alpha)} */
static void otf_param(const olpt_otf *o, int cam, const double X[3],
                      double out[5]) {
  const double *B = o->boundary;
  double px = fmin(fmax(X[0], B[0]), B[1]);
  double py = fmin(fmax(X[1], B[2]), B[3]);
  double pz = fmin(fmax(X[2], B[4]), B[5]);
  double dx = (B[1] - B[0]) / (o->nx - 1);
  double dy = (B[3] - B[2]) / (o->ny - 1);
  double dz = (B[5] - B[4]) / (o->nz - 1);
  int ix = (int)floor((px - B[0]) / dx);
  int iy = (int)floor((py - B[2]) / dy);
  int iz = (int)floor((pz - B[4]) / dz);
  ix = ix < 0 ? 0 : ix;
  iy = iy < 0 ? 0 : iy;
  iz = iz < 0 ? 0 : iz;
  if (ix > o->nx - 2) ix = o->nx - 2;
  if (iy > o->ny - 2) iy = o->ny - 2;
  if (iz > o->nz - 2) iz = o->nz - 2;
  double lim[6] = {linspace_at(B[0], B[1], o->nx, ix),
                   linspace_at(B[0], B[1], o->nx, ix + 1),
                   linspace_at(B[2], B[3], o->ny, iy),
                   linspace_at(B[2], B[3], o->ny, iy + 1),
                   linspace_at(B[4], B[5], o->nz, iz),
                   linspace_at(B[4], B[5], o->nz, iz + 1)};
  const int nyz = o->ny * o->nz, ng = o->nx * nyz;
  int g[8] = {ix * nyz + iy * o->nz + iz,
              (ix + 1) * nyz + iy * o->nz + iz,
              (ix + 1) * nyz + iy * o->nz + iz + 1,
              ix * nyz + iy * o->nz + iz + 1,
              ix * nyz + (iy + 1) * o->nz + iz,
              (ix + 1) * nyz + (iy + 1) * o->nz + iz,
              (ix + 1) * nyz + (iy + 1) * o->nz + iz + 1,
              ix * nyz + (iy + 1) * o->nz + iz + 1};
  double p[3] = {px, py, pz}, va[8], vb[8], vc[8], vl[8];
  for (int k = 0; k < 8; ++k) {
    va[k] = o->a[cam * ng + g[k]];
    vb[k] = o->b[cam * ng + g[k]];
    vc[k] = o->c[cam * ng + g[k]];
    vl[k] = o->alpha[cam * ng + g[k]];
  }
  out[0] = trilinear(lim, va, p);
  out[1] = trilinear(lim, vb, p);
  out[2] = trilinear(lim, vc, p);
  double al = trilinear(lim, vl, p);
  out[3] = cos(al);
  out[4] = sin(al);
}

/* ------------------------------------------------------------------------ */
/* ObjectFinder2D::findTracer2D (ObjectFinder.cpp:37-111) with             */
/* myMATH::isLocalMax (myMATH.h:398-439)                                    */
/* ------------------------------------------------------------------------ */
static int is_local_max(const double *img, int H, int W, int r, int c) {
  static const int mr[4] = {-1, 0, 1, 0}, mc[4] = {0, -1, 0, 1};
  double val = img[(size_t)r * W + c];
  for (int i = 0; i < 4; ++i) {
    int rr = r + mr[i], cc = c + mc[i];
    if (rr < 0 || rr >= H || cc < 0 || cc >= W) continue;
    double nv = img[(size_t)rr * W + cc];
    if (i < 2) {
      if (nv > val) return 0;
    } else {
      if (nv >= val) return 0;
    }
  }
  return 1;
}

static double safe_ln(double v) { return log(v < LOGSMALLNUMBER ? LOGSMALLNUMBER : v); }

static int find_peaks(const double *img, int H, int W, double min_int,
                      dvec *xy) {
  for (int row = 1; row < H - 1; ++row)
    for (int col = 1; col < W - 1; ++col) {
      const double cI = img[(size_t)row * W + col];
      if (cI < min_int) continue;
      if (!is_local_max(img, H, W, row, col)) continue;
      const int x1 = col - 1, x2 = col, x3 = col + 1;
      const int y1 = row - 1, y2 = row, y3 = row + 1;
      const double l1x = safe_ln(img[(size_t)y2 * W + x1]);
      const double l2 = safe_ln(cI);
      const double l3x = safe_ln(img[(size_t)y2 * W + x3]);
      const double nx = l1x * ((x2 * x2) - (x3 * x3)) - l2 * ((x1 * x1) - (x3 * x3)) +
                        l3x * ((x1 * x1) - (x2 * x2));
      const double dx = l1x * (x3 - x2) - l3x * (x1 - x2) + l2 * (x1 - x3);
      if (dx == 0.0) continue;
      const double xc = -0.5 * (nx / dx);
      if (!isfinite(xc)) continue;
      const double l1y = safe_ln(img[(size_t)y1 * W + x2]);
      const double l3y = safe_ln(img[(size_t)y3 * W + x2]);
      const double ny = l1y * ((y2 * y2) - (y3 * y3)) - l2 * ((y1 * y1) - (y3 * y3)) +
                        l3y * ((y1 * y1) - (y2 * y2));
      const double dy = l1y * (y3 - y2) - l3y * (y1 - y2) + l2 * (y1 - y3);
      if (dy == 0.0) continue;
      const double yc = -0.5 * (ny / dy);
      if (!isfinite(yc)) continue;
      if (dvec_push(xy, xc) || dvec_push(xy, yc)) return ERR_ALLOC;
    }
  return OK;
}

/* ------------------------------------------------------------------------ */
/* StereoMatch (StereoMatch.cpp)                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  int rows, cols, cell;
  int *start;   /* rows*cols+1 */
  int *pid;     /* bucketed point ids, ascending within a cell */
} idmap;

typedef struct {
  const olpt_camera *cams;
  int n_cam;
  int active[OLPT_MAX_CAM];
  const double *pts[OLPT_MAX_CAM]; /* xy per camera */
  int npts[OLPT_MAX_CAM];
  line3 *los[OLPT_MAX_CAM];        /* LOS per 2D point (deterministic, so
                                      computed once instead of per visit) */
  idmap map[OLPT_MAX_CAM];
  double tol2d, tol3d;
  double limit[6];
  int build_count, n_active;
} smctx;

/* IDMap ctor + rebuild (StereoMatch.cpp:1413-1438) */
static int idmap_build(idmap *m, int H, int W, int cell, const double *xy,
                       int n) {
  cell = cell < 1 ? 1 : cell;
  m->cell = cell;
  m->rows = H / cell + (H % cell ? 1 : 0);
  m->cols = W / cell + (W % cell ? 1 : 0);
  if (m->rows < 1) m->rows = 1;
  if (m->cols < 1) m->cols = 1;
  const int nc = m->rows * m->cols;
  m->start = (int *)calloc((size_t)nc + 1, sizeof(int));
  m->pid = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *cellof = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!m->start || !m->pid || !cellof) return ERR_ALLOC;
  for (int i = 0; i < n; ++i) {
    int cx = (int)floor(xy[2 * i] / cell);
    int cy = (int)floor(xy[2 * i + 1] / cell);
    if (cx < 0 || cy < 0 || cx >= m->cols || cy >= m->rows) {
      cellof[i] = -1;
      continue;
    }
    cellof[i] = cy * m->cols + cx;
    m->start[cellof[i] + 1]++;
  }
  for (int k = 0; k < nc; ++k) m->start[k + 1] += m->start[k];
  int *fill = (int *)malloc(sizeof(int) * (size_t)nc);
  memcpy(fill, m->start, sizeof(int) * (size_t)nc);
  for (int i = 0; i < n; ++i)
    if (cellof[i] >= 0) m->pid[fill[cellof[i]]++] = i;
  free(fill);
  free(cellof);
  return OK;
}

/* StereoMatch::makeLine2DFromLOS3D (StereoMatch.cpp:1053-1066) */
static void make_line2d(const olpt_camera *c, const line3 *L, line2 *out) {
  double a[2], b[2];
  project(c, L->pt, a);
  double q[3] = {L->pt[0] + L->u[0], L->pt[1] + L->u[1], L->pt[2] + L->u[2]};
  project(c, q, b);
  double d[2] = {b[0] - a[0], b[1] - a[1]};
  double nrm = sqrt(d[0] * d[0] + d[1] * d[1]);
  out->pt[0] = a[0];
  out->pt[1] = a[1];
  out->u[0] = d[0] / nrm;
  out->u[1] = d[1] / nrm;
}

/* IDMap::normalized (StereoMatch.cpp:1442-1451) */
static void normalized(const double u[2], double o[2]) {
  double n2 = u[0] * u[0] + u[1] * u[1];
  if (n2 <= 1e-24) {
    o[0] = 1.0;
    o[1] = 0.0;
    return;
  }
  double inv = 1.0 / sqrt(n2);
  o[0] = u[0] * inv;
  o[1] = u[1] * inv;
}

/* IDMap::computeStripIntersection (StereoMatch.cpp:1475-1566) then
 * IDMap::visitPointsInRowSpans (:1568-1627).  Output order: cell rows, then
 * cells left to right, then bucket order. */
static int strip_candidates(const smctx *S, int cam, const line2 *lines,
                            int nl, ivec *out) {
  const idmap *m = &S->map[cam];
  int *xmin = (int *)malloc(sizeof(int) * (size_t)m->rows);
  int *xmax = (int *)malloc(sizeof(int) * (size_t)m->rows);
  if (!xmin || !xmax) return ERR_ALLOC;
  for (int r = 0; r < m->rows; ++r) {
    xmin[r] = 0;
    xmax[r] = m->cols - 1;
  }
  const double hw = 0.5 * m->cell, hh = 0.5 * m->cell;
  const double to_index = 1.0 / m->cell;
  const double eps = 1e-12, tolg = S->tol2d + 1e-6, eps_c = 1e-9;
  for (int l = 0; l < nl; ++l) {
    double u[2];
    normalized(lines[l].u, u);
    const double n0 = -u[1], n1 = u[0];
    const double d = -(n0 * lines[l].pt[0] + n1 * lines[l].pt[1]);
    for (int cy = 0; cy < m->rows; ++cy) {
      if (xmin[cy] > xmax[cy]) continue;
      const double yc = (cy + 0.5) * m->cell;
      const double C = n1 * yc + d;
      const double pe = fabs(n0) * hw + fabs(n1) * hh;
      const double T = tolg + pe;
      int a = 0, b = m->cols - 1;
      if (fabs(n0) < eps) {
        if (fabs(C) > T) {
          xmin[cy] = 1;
          xmax[cy] = 0;
          continue;
        }
      } else {
        double xl = (-T - C) / n0;
        double xr = (T - C) / n0;
        if (xl > xr) {
          double t = xl;
          xl = xr;
          xr = t;
        }
        a = (int)ceil(xl * to_index - 0.5 - eps_c);
        b = (int)floor(xr * to_index - 0.5 + eps_c);
        if (a < 0) a = 0;
        if (b > m->cols - 1) b = m->cols - 1;
        if (a > b) {
          xmin[cy] = 1;
          xmax[cy] = 0;
          continue;
        }
      }
      if (a > xmin[cy]) xmin[cy] = a;
      if (b < xmax[cy]) xmax[cy] = b;
    }
  }
  double U[OLPT_MAX_CAM][2];
  for (int l = 0; l < nl; ++l) normalized(lines[l].u, U[l]);
  const double tol2 = S->tol2d * S->tol2d;
  const double *q = S->pts[cam];
  for (int cy = 0; cy < m->rows; ++cy) {
    if (xmin[cy] > xmax[cy]) continue;
    for (int cx = xmin[cy]; cx <= xmax[cy]; ++cx) {
      const int cell = cy * m->cols + cx;
      for (int k = m->start[cell]; k < m->start[cell + 1]; ++k) {
        const int pid = m->pid[k];
        int ok = 1;
        for (int l = 0; l < nl && ok; ++l) {
          const double dx = q[2 * pid] - lines[l].pt[0];
          const double dy = q[2 * pid + 1] - lines[l].pt[1];
          const double z = dy * U[l][0] - dx * U[l][1];
          if (z * z > tol2) ok = 0;
        }
        if (ok && ivec_push(out, pid)) return ERR_ALLOC;
      }
    }
  }
  free(xmin);
  free(xmax);
  return OK;
}

/* StereoMatch::checkBackProjection (StereoMatch.cpp:1206-1227) */
static int back_projection_ok(const smctx *S, int target, int pid,
                              const int *ccam, const double (*cpts)[2],
                              int k) {
  const line3 *lt = &S->los[target][pid];
  const double tol2 = S->tol2d * S->tol2d;
  for (int i = 0; i < k; ++i) {
    line2 L;
    make_line2d(&S->cams[ccam[i]], lt, &L);
    const double d0 = cpts[i][0] - L.pt[0], d1 = cpts[i][1] - L.pt[1];
    const double z = d1 * L.u[0] - d0 * L.u[1];
    if (z * z > tol2) return 0;
  }
  return 1;
}

/* StereoMatch::enumerateCandidatesOnCam (StereoMatch.cpp:449-498) */
static int enumerate_on_cam(const smctx *S, const line3 *const *los, int nl,
                            int target, const int *ccam,
                            const double (*cpts)[2], ivec *out) {
  out->n = 0;
  line2 lines[OLPT_MAX_CAM];
  for (int i = 0; i < nl; ++i) make_line2d(&S->cams[target], los[i], &lines[i]);
  int rc = strip_candidates(S, target, lines, nl, out);
  if (rc) return rc;
  int w = 0;
  for (int i = 0; i < out->n; ++i)
    if (back_projection_ok(S, target, out->v[i], ccam, cpts, nl))
      out->v[w++] = out->v[i];
  out->n = w;
  return OK;
}

/* StereoMatch::computeMinParallaxDeg (StereoMatch.cpp:1082-1099) */
static double min_parallax_deg(const line3 *const *los, int n) {
  if (n < 2) return 180.0;
  double mn = 180.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double *a = los[i]->u, *b = los[j]->u;
      double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
      dot = fmax(-1.0, fmin(1.0, dot));
      double deg = acos(dot) * 180.0 / M_PI;
      if (deg < mn) mn = deg;
    }
  return mn;
}

/* StereoMatch::calTriangulateTol (StereoMatch.cpp:1103-1126) */
static double tri_tol(double tol, int k, int m, double th_deg) {
  const int kk = k > 2 ? k : 2, mm = m > 2 ? m : 2;
  const double ad = sqrt((double)mm / (double)kk);
  const double th = fmax(th_deg, 1.0);
  const double num = sin(6.0 * M_PI / 180.0);
  const double den = fmax(1e-6, sin(th * M_PI / 180.0));
  const double aa = fmax(1.0, fmin(2.0, num / den));
  return tol * ad * aa;
}

/* StereoMatch::selectSecondCameraByLineLength (StereoMatch.cpp:1307-1370) */
static int select_second(const smctx *S, const line3 *lr, const int *rem,
                         int nrem) {
  int best = -1;
  double best_len = INFINITY;
  for (int ri = 0; ri < nrem; ++ri) {
    const int cam = rem[ri];
    line2 L;
    make_line2d(&S->cams[cam], lr, &L);
    const double px = L.pt[0], py = L.pt[1], ux = L.u[0], uy = L.u[1];
    const double W = (double)S->cams[cam].n_col, H = (double)S->cams[cam].n_row;
    double pts[4][2];
    int np = 0;
#define PUSH_IF(X, Y)                                                         \
  do {                                                                        \
    double _x = (X), _y = (Y);                                                \
    if (_x >= 0.0 && _x <= W && _y >= 0.0 && _y <= H) {                       \
      pts[np][0] = _x;                                                        \
      pts[np][1] = _y;                                                        \
      np++;                                                                   \
    }                                                                         \
  } while (0)
    if (fabs(ux) > 1e-12) {
      double t = (0.0 - px) / ux;
      PUSH_IF(0.0, py + t * uy);
      t = (W - 1 - px) / ux;
      PUSH_IF(W - 1, py + t * uy);
    }
    if (fabs(uy) > 1e-12) {
      double t = (0.0 - py) / uy;
      PUSH_IF(px + t * ux, 0.0);
      t = (H - 1 - py) / uy;
      PUSH_IF(px + t * ux, H - 1);
    }
#undef PUSH_IF
    double uq[4][2];
    int nu = 0;
    for (int i = 0; i < np; ++i) {
      int dup = 0;
      for (int j = 0; j < nu; ++j)
        if (fabs(uq[j][0] - pts[i][0]) < 1e-9 && fabs(uq[j][1] - pts[i][1]) < 1e-9) {
          dup = 1;
          break;
        }
      if (!dup) {
        uq[nu][0] = pts[i][0];
        uq[nu][1] = pts[i][1];
        nu++;
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

/* StereoMatch::selectNextCameraByMaxPairAngle (StereoMatch.cpp:1375-1411) */
static int select_next(const smctx *S, const line3 *const *los, int n,
                       const int *rem, int nrem) {
  int best = -1;
  double best_score = -1.0;
  for (int ri = 0; ri < nrem; ++ri) {
    const int cam = rem[ri];
    line2 L[OLPT_MAX_CAM];
    for (int i = 0; i < n; ++i) make_line2d(&S->cams[cam], los[i], &L[i]);
    double score = 0.0;
    if (n >= 2) {
      for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
          const double s = fabs(L[i].u[0] * L[j].u[1] - L[i].u[1] * L[j].u[0]);
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

static int in_limit(const double lim[6], const double X[3]) {
  return !(X[0] > lim[1] || X[0] < lim[0] || X[1] > lim[3] || X[1] < lim[2] ||
           X[2] > lim[5] || X[2] < lim[4]);
}

/* DFS of StereoMatch::buildMatch (StereoMatch.cpp:184-444); the emitted set
 * does not depend on visiting order (candidates are sorted afterwards). */
typedef struct {
  int cams[OLPT_MAX_CAM], ids[OLPT_MAX_CAM];
  double pts[OLPT_MAX_CAM][2];
  const line3 *los[OLPT_MAX_CAM];
  int k;
  int rem[OLPT_MAX_CAM], nrem;
} path_t;

static int dfs(const smctx *S, path_t *P, int target, const ivec *cands,
               ivec *emit) {
  for (int ci = 0; ci < cands->n; ++ci) {
    const int pid = cands->v[ci];
    path_t N = *P;
    N.cams[N.k] = target;
    N.ids[N.k] = pid;
    N.pts[N.k][0] = S->pts[target][2 * pid];
    N.pts[N.k][1] = S->pts[target][2 * pid + 1];
    N.los[N.k] = &S->los[target][pid];
    N.k++;
    int w = 0;
    for (int i = 0; i < N.nrem; ++i)
      if (N.rem[i] != target) N.rem[w++] = N.rem[i];
    N.nrem = w;
    {
      const double th = min_parallax_deg(N.los, N.k);
      const double tq = tri_tol(S->tol3d, N.k, S->build_count, th);
      double X[3], e;
      if (triangulate(N.los, N.k, X, &e) != 0) continue;
      if (!(e <= tq)) continue;
    }
    if (N.k == S->build_count) {
      for (int i = 0; i < S->n_cam; ++i) {
        int id = -1;
        for (int j = 0; j < N.k; ++j)
          if (N.cams[j] == i) id = N.ids[j];
        if (ivec_push(emit, id)) return ERR_ALLOC;
      }
      continue;
    }
    const int nt = select_next(S, N.los, N.k, N.rem, N.nrem);
    if (nt < 0) continue;
    ivec nc = {0};
    int rc = enumerate_on_cam(S, N.los, N.k, nt, N.cams,
                              (const double(*)[2])N.pts, &nc);
    if (rc) {
      free(nc.v);
      return rc;
    }
    if (nc.n > 0) rc = dfs(S, &N, nt, &nc, emit);
    free(nc.v);
    if (rc) return rc;
  }
  return OK;
}

/* StereoMatch::checkMatch, all-cameras-built branch (StereoMatch.cpp:504-540) */
static int check_match(const smctx *S, const int *ids, double *e) {
  const line3 *los[OLPT_MAX_CAM];
  int n = 0;
  for (int c = 0; c < S->n_cam; ++c)
    if (ids[c] >= 0) los[n++] = &S->los[c][ids[c]];
  if (n < 2) return 0;
  double X[3];
  if (triangulate(los, n, X, e) != 0) return -1; /* reference would throw */
  if (*e > S->tol3d) return 0;
  if (!in_limit(S->limit, X)) return 0;
  return 1;
}

/* eq_eps of StereoMatch::pruneMatch (StereoMatch.cpp:653-657) */
static int eq_eps(double a, double b) {
  const double s = fmax(fabs(a), fabs(b));
  return fabs(a - b) <= (s * 1e-9 + 1e-12);
}

typedef struct { double e; int j; } ej_t;

static int cmp_pool(const ej_t *A, const ej_t *B) {
  if (!eq_eps(A->e, B->e)) return A->e < B->e ? -1 : 1;
  return A->j < B->j ? -1 : (A->j > B->j ? 1 : 0);
}

static void msort_pool(ej_t *a, ej_t *tmp, int n) {
  if (n < 2) return;
  int h = n / 2;
  msort_pool(a, tmp, h);
  msort_pool(a + h, tmp, n - h);
  int i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = (cmp_pool(&a[j], &a[i]) < 0) ? a[j++] : a[i++];
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, sizeof(ej_t) * (size_t)n);
}

typedef struct { const double *pct, *e; } ordctx;

static int less_order(const ordctx *o, int aj, int bj) {
  const double pa = o->pct[aj], pb = o->pct[bj];
  const int af = isfinite(pa), bf = isfinite(pb);
  if (af && bf) {
    if (!eq_eps(pa, pb)) return pa < pb;
  } else if (af) {
    return 1;
  } else if (bf) {
    return 0;
  }
  const double ea = o->e[aj], eb = o->e[bj];
  const int aef = isfinite(ea), bef = isfinite(eb);
  if (aef && bef) {
    if (!eq_eps(ea, eb)) return ea < eb;
  } else if (aef) {
    return 1;
  } else if (bef) {
    return 0;
  }
  return aj < bj;
}

static void msort_order(const ordctx *o, int *a, int *tmp, int n) {
  if (n < 2) return;
  int h = n / 2;
  msort_order(o, a, tmp, h);
  msort_order(o, a + h, tmp, n - h);
  int i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = less_order(o, a[j], a[i]) ? a[j++] : a[i++];
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, sizeof(int) * (size_t)n);
}

/* StereoMatch::pruneMatch (StereoMatch.cpp:631-980).  ids: M x n_cam in
 * candidate order; sel[j] set to 1 for selected candidates. */
static int prune_match(const smctx *S, const int32_t *ids, const double *e,
                       int M, uint8_t *sel) {
  const int nc = S->n_cam;
  memset(sel, 0, (size_t)M);
  if (M == 0) return OK;
  size_t off[OLPT_MAX_CAM + 1];
  off[0] = 0;
  for (int c = 0; c < nc; ++c) off[c + 1] = off[c] + (size_t)S->npts[c];
  const size_t total = off[nc];
  int *nfl = (int *)calloc((size_t)M, sizeof(int));
  size_t *fl = (size_t *)malloc(sizeof(size_t) * (size_t)M * nc);
  int *pstart = (int *)calloc(total + 1, sizeof(int));
  double *psum = (double *)calloc((size_t)M, sizeof(double));
  int *pcnt = (int *)calloc((size_t)M, sizeof(int));
  double *pct = (double *)malloc(sizeof(double) * (size_t)M);
  int *order = (int *)malloc(sizeof(int) * (size_t)M);
  int *tmpi = (int *)malloc(sizeof(int) * (size_t)M);
  int *owner = (int *)malloc(sizeof(int) * (total + 1));
  if (!nfl || !fl || !pstart || !psum || !pcnt || !pct || !order || !tmpi || !owner)
    return ERR_ALLOC;
  for (int j = 0; j < M; ++j)
    for (int c = 0; c < nc; ++c) {
      if (!S->active[c]) continue;
      const int pid = ids[(size_t)j * nc + c];
      if (pid < 0 || pid >= S->npts[c]) continue;
      fl[(size_t)j * nc + nfl[j]++] = off[c] + (size_t)pid;
    }
  for (int j = 0; j < M; ++j)
    for (int k = 0; k < nfl[j]; ++k) pstart[fl[(size_t)j * nc + k] + 1]++;
  for (size_t f = 0; f < total; ++f) pstart[f + 1] += pstart[f];
  int *plist = (int *)malloc(sizeof(int) * (size_t)(pstart[total] > 0 ? pstart[total] : 1));
  int *pfill = (int *)malloc(sizeof(int) * (total + 1));
  memcpy(pfill, pstart, sizeof(int) * (total + 1));
  for (int j = 0; j < M; ++j)
    for (int k = 0; k < nfl[j]; ++k) {
      size_t f = fl[(size_t)j * nc + k];
      plist[pfill[f]++] = j;
    }
  ej_t *pool = (ej_t *)malloc(sizeof(ej_t) * (size_t)M);
  ej_t *ptmp = (ej_t *)malloc(sizeof(ej_t) * (size_t)M);
  for (size_t f = 0; f < total; ++f) {
    int n = 0;
    for (int k = pstart[f]; k < pstart[f + 1]; ++k) {
      const int j = plist[k];
      if (isfinite(e[j])) {
        pool[n].e = e[j];
        pool[n].j = j;
        n++;
      }
    }
    if (n == 0) continue;
    if (n == 1) {
      pcnt[pool[0].j] += 1;
      continue;
    }
    msort_pool(pool, ptmp, n);
    int k = 0;
    while (k < n) {
      int k0 = k, k1 = k + 1;
      while (k1 < n && eq_eps(pool[k1].e, pool[k0].e)) ++k1;
      const double avg_rank = 0.5 * (double)(k0 + (k1 - 1));
      const double denom = (double)(n - 1);
      const double pv = (denom > 0.0) ? (avg_rank / denom) : 0.0;
      for (int t = k0; t < k1; ++t) {
        psum[pool[t].j] += pv;
        pcnt[pool[t].j] += 1;
      }
      k = k1;
    }
  }
  for (int j = 0; j < M; ++j) pct[j] = pcnt[j] > 0 ? psum[j] / (double)pcnt[j] : INFINITY;
  for (int j = 0; j < M; ++j) order[j] = j;
  ordctx oc = {pct, e};
  msort_order(&oc, order, tmpi, M);
  for (size_t f = 0; f <= total; ++f) owner[f] = -1;
  for (int oi = 0; oi < M; ++oi) {
    const int j = order[oi];
    if (!isfinite(pct[j]) || !isfinite(e[j])) continue;
    const size_t *fj = &fl[(size_t)j * nc];
    const int nf = nfl[j];
    if (nf == 0) continue;
    int conflict = 0;
    for (int k = 0; k < nf; ++k)
      if (owner[fj[k]] != -1) {
        conflict = 1;
        break;
      }
    if (!conflict) {
      for (int k = 0; k < nf; ++k) owner[fj[k]] = j;
      sel[j] = 1;
      continue;
    }
    int Sset[OLPT_MAX_CAM], ns = 0;
    for (int k = 0; k < nf; ++k) {
      const int o = owner[fj[k]];
      if (o == -1) continue;
      int seen = 0;
      for (int t = 0; t < ns; ++t)
        if (Sset[t] == o) seen = 1;
      if (!seen) Sset[ns++] = o;
    }
    double ap = 0.0, ae = 0.0;
    int cnt = 0;
    for (int t = 0; t < ns; ++t) {
      const double po = pct[Sset[t]], eo = e[Sset[t]];
      if (!isfinite(po) || !isfinite(eo)) continue;
      ap += po;
      ae += eo;
      ++cnt;
    }
    if (cnt == 0) continue;
    ap /= (double)cnt;
    ae /= (double)cnt;
    const double pj = pct[j], ej = e[j];
    const int better_pct = (!eq_eps(pj, ap)) && ((pj + 1e-12) < ap);
    const int tie_better = eq_eps(pj, ap) && (!eq_eps(ej, ae)) && ((ej + 1e-12) < ae);
    if (better_pct || tie_better) {
      for (int t = 0; t < ns; ++t) {
        const int o = Sset[t];
        const size_t *fo = &fl[(size_t)o * nc];
        for (int k = 0; k < nfl[o]; ++k)
          if (owner[fo[k]] == o) owner[fo[k]] = -1;
        sel[o] = 0;
      }
      for (int k = 0; k < nf; ++k) owner[fj[k]] = j;
      sel[j] = 1;
    }
  }
  free(nfl); free(fl); free(pstart); free(psum); free(pcnt); free(pct);
  free(order); free(tmpi); free(owner); free(plist); free(pfill);
  free(pool); free(ptmp);
  return OK;
}

typedef struct { int32_t *ids; double e; } cand_t;
static int g_ncam_sort;
static int cmp_cand(const void *a, const void *b) {
  const cand_t *A = (const cand_t *)a, *B = (const cand_t *)b;
  for (int c = 0; c < g_ncam_sort; ++c) {
    if (A->ids[c] != B->ids[c]) return A->ids[c] < B->ids[c] ? -1 : 1;
  }
  return 0;
}

static void sm_free(smctx *S) {
  for (int c = 0; c < S->n_cam; ++c) {
    free(S->los[c]);
    free(S->map[c].start);
    free(S->map[c].pid);
  }
}

static int sm_init(smctx *S, const olpt_camera *cams, int n_cam,
                   const int32_t *active, const double *xy,
                   const int32_t *n_pts, const olpt_ipr_params *p) {
  memset(S, 0, sizeof(*S));
  if (n_cam < 1 || n_cam > OLPT_MAX_CAM) return ERR_ARG;
  S->cams = cams;
  S->n_cam = n_cam;
  S->tol2d = p->tol_2d_px;
  S->tol3d = p->tol_3d_mm;
  memcpy(S->limit, p->limit, sizeof(S->limit));
  size_t off = 0;
  for (int c = 0; c < n_cam; ++c) {
    S->active[c] = active ? (active[c] != 0) : 1;
    S->pts[c] = xy + 2 * off;
    S->npts[c] = n_pts[c];
    off += (size_t)n_pts[c];
    if (S->active[c]) S->n_active++;
  }
  /* StereoMatch ctor pre-checks (StereoMatch.cpp:27-47) */
  if (S->n_active < 2) return ERR_ARG;
  if (p->match_cam_count < 2 || p->match_cam_count > S->n_active) return ERR_ARG;
  int bc = p->match_cam_count > 2 ? p->match_cam_count : 2;
  S->build_count = bc < S->n_active ? bc : S->n_active;
  /* With check cameras (build_count < n_active) the reference's checkMatch
   * leaves e_check NaN (StereoMatch.cpp:522, :603-604) and match() aborts
   * (:126-127); only the all-cameras-built configuration is defined. */
  if (S->build_count != S->n_active) return ERR_UNSUP;
  for (int c = 0; c < n_cam; ++c) {
    if (!S->active[c]) continue;
    const int n = S->npts[c];
    S->los[c] = (line3 *)malloc(sizeof(line3) * (size_t)(n > 0 ? n : 1));
    if (!S->los[c]) return ERR_ALLOC;
    for (int i = 0; i < n; ++i)
      line_of_sight(&cams[c], S->pts[c][2 * i], S->pts[c][2 * i + 1], &S->los[c][i]);
    int rc = idmap_build(&S->map[c], cams[c].n_row, cams[c].n_col,
                         p->idmap_cell_px, S->pts[c], n);
    if (rc) return rc;
  }
  return OK;
}

/* Build + check for one reference observation (StereoMatch.cpp:108-136). */
static int build_one(const smctx *S, int ref_cam, int r, ivec *ids_out,
                     dvec *e_out) {
  path_t P;
  memset(&P, 0, sizeof(P));
  P.cams[0] = ref_cam;
  P.ids[0] = r;
  P.pts[0][0] = S->pts[ref_cam][2 * r];
  P.pts[0][1] = S->pts[ref_cam][2 * r + 1];
  P.los[0] = &S->los[ref_cam][r];
  P.k = 1;
  for (int c = 0; c < S->n_cam; ++c)
    if (S->active[c] && c != ref_cam) P.rem[P.nrem++] = c;
  if (P.nrem == 0) return OK;
  const int t = select_second(S, P.los[0], P.rem, P.nrem);
  if (t < 0) return OK;
  ivec c2 = {0}, em = {0};
  int rc = enumerate_on_cam(S, P.los, 1, t, P.cams, (const double(*)[2])P.pts, &c2);
  if (!rc && c2.n > 0) rc = dfs(S, &P, t, &c2, &em);
  free(c2.v);
  if (rc) {
    free(em.v);
    return rc;
  }
  for (int i = 0; i < em.n / S->n_cam; ++i) {
    const int *ids = &em.v[i * S->n_cam];
    double e = NAN;
    int ok = check_match(S, ids, &e);
    if (ok < 0) {
      free(em.v);
      return ERR_UNSUP;
    }
    if (!ok) continue;
    for (int c = 0; c < S->n_cam; ++c)
      if (ivec_push(ids_out, ids[c])) return ERR_ALLOC;
    if (dvec_push(e_out, e)) return ERR_ALLOC;
  }
  free(em.v);
  return OK;
}

/* Full StereoMatch::match.  Results are written into freshly allocated
 * arrays (caller frees): sorted candidates, selection flags, objects. */
typedef struct {
  int M;
  int32_t *ids;  /* M x n_cam */
  double *e;
  uint8_t *sel;
  int n_obj;
  double *xyz;   /* n_obj x 3 */
  int32_t *oids; /* n_obj x n_cam */
} match_out;

static void match_free(match_out *o) {
  free(o->ids); free(o->e); free(o->sel); free(o->xyz); free(o->oids);
  memset(o, 0, sizeof(*o));
}

static int stereo_match(const smctx *S, match_out *out) {
  memset(out, 0, sizeof(*out));
  const int nc = S->n_cam;
  int ref_cam = -1;
  long best = -1;
  for (int i = 0; i < nc; ++i)
    if (S->active[i] && (best < 0 || S->npts[i] < best)) {
      best = S->npts[i];
      ref_cam = i;
    }
  const int nr = S->npts[ref_cam];
  ivec *pid = (ivec *)calloc((size_t)(nr > 0 ? nr : 1), sizeof(ivec));
  dvec *pe = (dvec *)calloc((size_t)(nr > 0 ? nr : 1), sizeof(dvec));
  int err = OK;
#pragma omp parallel for schedule(dynamic, 32)
  for (int r = 0; r < nr; ++r) {
    int rc = build_one(S, ref_cam, r, &pid[r], &pe[r]);
    if (rc) {
#pragma omp critical
      err = rc;
    }
  }
  if (err) return err;
  int M = 0;
  for (int r = 0; r < nr; ++r) M += pe[r].n;
  cand_t *cs = (cand_t *)malloc(sizeof(cand_t) * (size_t)(M > 0 ? M : 1));
  int32_t *allids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(M > 0 ? M : 1) * nc);
  int m = 0;
  for (int r = 0; r < nr; ++r) {
    for (int i = 0; i < pe[r].n; ++i) {
      memcpy(&allids[(size_t)m * nc], &pid[r].v[(size_t)i * nc], sizeof(int32_t) * nc);
      cs[m].ids = &allids[(size_t)m * nc];
      cs[m].e = pe[r].v[i];
      ++m;
    }
    free(pid[r].v);
    free(pe[r].v);
  }
  free(pid);
  free(pe);
  g_ncam_sort = nc;
  qsort(cs, (size_t)M, sizeof(cand_t), cmp_cand);
  out->M = M;
  out->ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(M > 0 ? M : 1) * nc);
  out->e = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
  out->sel = (uint8_t *)malloc((size_t)(M > 0 ? M : 1));
  for (int j = 0; j < M; ++j) {
    memcpy(&out->ids[(size_t)j * nc], cs[j].ids, sizeof(int32_t) * nc);
    out->e[j] = cs[j].e;
  }
  free(cs);
  free(allids);
  int rc = prune_match(S, out->ids, out->e, M, out->sel);
  if (rc) return rc;
  /* StereoMatch::triangulateMatch (StereoMatch.cpp:985-1050) */
  int ns = 0;
  for (int j = 0; j < M; ++j) ns += out->sel[j];
  out->xyz = (double *)malloc(sizeof(double) * 3 * (size_t)(ns > 0 ? ns : 1));
  out->oids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ns > 0 ? ns : 1) * nc);
  int no = 0;
  for (int j = 0; j < M; ++j) {
    if (!out->sel[j]) continue;
    const int32_t *ids = &out->ids[(size_t)j * nc];
    const line3 *los[OLPT_MAX_CAM];
    int n = 0;
    for (int c = 0; c < nc; ++c)
      if (ids[c] >= 0) los[n++] = &S->los[c][ids[c]];
    if (n < 2) continue;
    double X[3], e;
    if (triangulate(los, n, X, &e) != 0) continue;
    if (e > S->tol3d) continue;
    memcpy(&out->xyz[3 * no], X, sizeof(X));
    memcpy(&out->oids[(size_t)no * nc], ids, sizeof(int32_t) * nc);
    ++no;
  }
  out->n_obj = no;
  return OK;
}

/* ------------------------------------------------------------------------ */
/* Shake (Shake.cpp), tracer strategy                                       */
/* ------------------------------------------------------------------------ */
typedef struct { int rmin, rmax, cmin, cmax; } prange; /* [min, max) */

typedef struct {
  const olpt_camera *cams;
  int n_cam;
  int active[OLPT_MAX_CAM];
  int H, W;
  const olpt_otf *otf;
  double r_px;
  int ratio;
  double tol3d;
} shctx;

typedef struct {
  double X[3];
  double uv[OLPT_MAX_CAM][2];
} sobj;

static void sobj_project(const shctx *S, sobj *o) {
  for (int k = 0; k < S->n_cam; ++k) project(&S->cams[k], o->X, o->uv[k]);
}

/* Shake::calROIBound (Shake.cpp:246-277) */
static prange roi_bound(const shctx *S, int k, double cx, double cy,
                        double dx, double dy) {
  const int H = S->cams[k].n_row, W = S->cams[k].n_col;
  prange e = {1, 0, 1, 0};
  if (H <= 0 || W <= 0 || dx <= 0.0 || dy <= 0.0 || cx <= 0 || cy <= 0 ||
      cx >= W || cy >= H)
    return e;
  const int width = (int)round(2.0 * dx);
  const int height = (int)round(2.0 * dy);
  int cmin = (int)round(cx - (width - 1) / 2.0);
  int rmin = (int)round(cy - (height - 1) / 2.0);
  int cmax = cmin + width, rmax = rmin + height;
  rmin = rmin > 0 ? rmin : 0;
  cmin = cmin > 0 ? cmin : 0;
  rmax = rmax < H ? rmax : H;
  cmax = cmax < W ? cmax : W;
  if (rmax < rmin || cmax < cmin) return e;
  prange r = {rmin, rmax, cmin, cmax};
  return r;
}

static int pr_empty(prange r) { return r.rmax < r.rmin || r.cmax < r.cmin; }

/* TracerShakeStrategy::gaussIntensity / project2DInt (Shake.cpp:901-953):
 * OTF parameters taken at the object's 3D centre. */
static double project_int(const shctx *S, const sobj *o, int k, int row,
                          int col) {
  double P[5];
  otf_param(S->otf, k, o->X, P);
  const double dx = col - o->uv[k][0];
  const double dy = row - o->uv[k][1];
  const double xx = dx * P[3] + dy * P[4];
  const double yy = -dx * P[4] + dy * P[3];
  const double v = P[0] * exp(-(P[1] * xx * xx + P[2] * yy * yy));
  return v > 0.0 ? v : 0.0;
}

/* Shake::calResidueImage (Shake.cpp:165-244).  res = copy of orig, then min
 * fusion of orig - projection over each object's radius box. */
static void residue_image(const shctx *S, const sobj *objs, int n,
                          const uint8_t *flags, int output_ipr,
                          const double *orig, double *res) {
  const size_t HW = (size_t)S->H * S->W;
  memcpy(res, orig, sizeof(double) * HW * S->n_cam);
  for (int k = 0; k < S->n_cam; ++k) {
    if (!S->active[k] && !output_ipr) continue;
    double *R = res + HW * k;
    const double *O = orig + HW * k;
    for (int i = 0; i < n; ++i) {
      if (flags && flags[i] != 0) continue;
      const prange r = roi_bound(S, k, objs[i].uv[k][0], objs[i].uv[k][1],
                                 S->r_px, S->r_px);
      if (pr_empty(r)) continue;
      for (int row = r.rmin; row < r.rmax; ++row)
        for (int col = r.cmin; col < r.cmax; ++col) {
          const double p = project_int(S, &objs[i], k, row, col);
          if (p == 0.0) continue;
          double *rr = &R[(size_t)row * S->W + col];
          const double cand = O[(size_t)row * S->W + col] - p;
          if (cand < *rr) *rr = cand;
          if (output_ipr && *rr < 0) *rr = 0.0;
        }
    }
  }
}

typedef struct {
  prange roi[OLPT_MAX_CAM];
  double *aug[OLPT_MAX_CAM];
} roiinfo;

static int roi_in(const prange *r, int row, int col) {
  if (r->rmax - r->rmin <= 0 || r->cmax - r->cmin <= 0) return 0;
  return !(row < r->rmin || row >= r->rmax || col < r->cmin || col >= r->cmax);
}
static double *roi_at(roiinfo *R, int k, int row, int col) {
  const prange *r = &R->roi[k];
  return &R->aug[k][(size_t)(row - r->rmin) * (r->cmax - r->cmin) + (col - r->cmin)];
}

/* Shake::buildROIInfo (Shake.cpp:279-328) */
static void build_roi(const shctx *S, const sobj *o, const double *res,
                      roiinfo *R) {
  const size_t HW = (size_t)S->H * S->W;
  for (int k = 0; k < S->n_cam; ++k) {
    prange z = {0, 0, 0, 0};
    R->roi[k] = z;
    R->aug[k] = NULL;
    if (!S->active[k]) continue;
    const double d = S->r_px > 0.0 ? S->r_px : 0.0;
    const prange r = roi_bound(S, k, o->uv[k][0], o->uv[k][1], d * S->ratio,
                               d * S->ratio);
    R->roi[k] = r;
    if (pr_empty(r)) continue;
    R->aug[k] = (double *)malloc(sizeof(double) * (size_t)((r.rmax - r.rmin) * (r.cmax - r.cmin) + 1));
    const prange rp = roi_bound(S, k, o->uv[k][0], o->uv[k][1], d, d);
    for (int row = r.rmin; row < r.rmax; ++row)
      for (int col = r.cmin; col < r.cmax; ++col) {
        double a = res[HW * k + (size_t)row * S->W + col];
        if (row >= rp.rmin && row < rp.rmax && col >= rp.cmin && col < rp.cmax)
          a += project_int(S, o, k, row, col);
        *roi_at(R, k, row, col) = a > 0.0 ? a : 0.0;
      }
  }
}

static void free_roi(const shctx *S, roiinfo *R) {
  for (int k = 0; k < S->n_cam; ++k) free(R->aug[k]);
}

/* TracerShakeStrategy::calShakeResidue (Shake.cpp:968-1045) */
static double shake_residue(const shctx *S, const sobj *o, roiinfo *R) {
  int used = 0;
  double acc = 0.0;
  for (int k = 0; k < S->n_cam; ++k) {
    if (!S->active[k]) continue;
    const prange *r = &R->roi[k];
    const int nrow = r->rmax - r->rmin, ncol = r->cmax - r->cmin;
    if (nrow <= 0 || ncol <= 0) continue;
    double sse = 0.0;
    size_t pix = 0;
    const double u = o->uv[k][0], v = o->uv[k][1];
    const int prmin = (int)floor(v - S->r_px), prmax = (int)ceil(v + S->r_px + 1);
    const int pcmin = (int)floor(u - S->r_px), pcmax = (int)ceil(u + S->r_px + 1);
    for (int row = r->rmin; row < r->rmax; ++row)
      for (int col = r->cmin; col < r->cmax; ++col) {
        const double meas = *roi_at(R, k, row, col);
        double pred = 0.0;
        if (row >= prmin && row < prmax && col >= pcmin && col < pcmax)
          pred = project_int(S, o, k, row, col);
        const double d = meas - pred;
        sse += d * d;
        ++pix;
      }
    if (pix > 0) {
      acc += sse / (double)pix;
      ++used;
    }
  }
  return used > 1 ? acc / (double)used : INFINITY;
}

static int cmp_dbl(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static double median(double *v, int n) {
  if (n == 0) return 0.0;
  qsort(v, (size_t)n, sizeof(double), cmp_dbl);
  const int mid = n / 2;
  double m = v[mid];
  if ((n & 1) == 0) m = 0.5 * (m + v[mid - 1]);
  return m;
}

static int in_roi_d(const prange *r, double u, double v) {
  return r->cmin <= u && u < r->cmax && r->rmin <= v && v < r->rmax;
}

/* Shake::shakeOneObject (Shake.cpp:409-665).  The temporary object keeps,
 * on every axis, the coordinate it was last evaluated at (the reference
 * creates it once before the axis loop and only writes the current axis). */
static void shake_one(const shctx *S, sobj *obj, roiinfo *R, double delta) {
  const int nc = S->n_cam;
  sobj tmp = *obj;
  for (int ax = 0; ax < 3; ++ax) {
    const double x0 = obj->X[ax];
    const double DL = -delta, DR = +delta;
    const double rel[3] = {DL, 0.0, DR};
    double f[3];
    double uvs[3][OLPT_MAX_CAM][2];
    int ok[3][OLPT_MAX_CAM];
    for (int j = 0; j < 3; ++j) {
      tmp.X[ax] = x0 + rel[j];
      sobj_project(S, &tmp);
      f[j] = shake_residue(S, &tmp, R);
      for (int k = 0; k < nc; ++k) {
        uvs[j][k][0] = tmp.uv[k][0];
        uvs[j][k][1] = tmp.uv[k][1];
        ok[j][k] = in_roi_d(&R->roi[k], tmp.uv[k][0], tmp.uv[k][1]);
      }
    }
    const double fL = f[0], f0 = f[1], fR = f[2];
    const double tol = 1e-3 * fmax(1.0, f0);
    const int mono_dec = (fL > f0 + tol) && (f0 > fR + tol);
    const int mono_inc = (fR > f0 + tol) && (f0 > fL + tol);
    const int mono = mono_dec || mono_inc;
    int s = mono_dec ? +1 : (mono_inc ? -1 : 0);
    const double kappa = (fL - 2.0 * f0 + fR);
    const double g = (fR - fL) / fmax(2.0 * delta, 1e-12);
    double cx[5] = {DL, 0.0, DR}, cf[5] = {fL, f0, fR}, cabs[5] = {x0 + DL, x0, x0 + DR};
    int nc_ = 3;
    int skip_expand = 0, allow_xhat = 1;
    if (kappa > 0.0) {
      const double x3 = delta * (fL - fR) / (2.0 * kappa);
      if (fabs(x3) <= 0.95 * delta) {
        const double a = kappa / (delta * delta + 1e-12);
        const double rho = (g * g) / (2.0 * a * fmax(f0, 1.0));
        skip_expand = 1;
        if (rho < 8e-4) allow_xhat = 0;
      }
    }
    if (!skip_expand && mono) {
      double Jv[OLPT_MAX_CAM], mg[OLPT_MAX_CAM];
      int nJ = 0, nm = 0;
      for (int k = 0; k < nc; ++k) {
        if (ok[0][k] && ok[2][k]) {
          const double du = uvs[2][k][0] - uvs[0][k][0];
          const double dv = uvs[2][k][1] - uvs[0][k][1];
          Jv[nJ++] = hypot(du, dv) / fmax(2.0 * delta, 1e-12);
        }
        if (ok[1][k]) {
          const prange *rr = &R->roi[k];
          const double mu = fmin(uvs[1][k][0] - rr->cmin, rr->cmax - 1.0 - uvs[1][k][0]);
          const double mv = fmin(uvs[1][k][1] - rr->rmin, rr->rmax - 1.0 - uvs[1][k][1]);
          mg[nm++] = fmin(mu, mv);
        }
      }
      const double Jx = median(Jv, nJ);
      const double mpx = median(mg, nm);
      const double Droi = (Jx > 0.0 && mpx > 0.0) ? (0.5 * mpx / Jx) : 0.0;
      double Dtry = 0.0;
      if (kappa > 0.0) {
        const double x3 = delta * (fL - fR) / (2.0 * kappa);
        if (fabs(x3) > delta) {
          s = (x3 > 0.0 ? +1 : -1);
          if (Jx > 0.0 && fabs(x3) > (1.0 / Jx))
            Dtry = 1.0 / Jx;
          else
            Dtry = 1.15 * fabs(x3);
        }
      }
      if (Dtry <= 0.0 && Jx > 0.0) Dtry = 1.0 / Jx;
      const double Dp = (Dtry > 0.0 && Droi > 0.0) ? fmin(Dtry, Droi) : 0.0;
      if (Dp > 0.0) {
        const double dxe = (s >= 0 ? +Dp : -Dp);
        tmp.X[ax] = x0 + dxe;
        sobj_project(S, &tmp);
        int out = 0;
        for (int k = 0; k < nc; ++k)
          if (!in_roi_d(&R->roi[k], tmp.uv[k][0], tmp.uv[k][1])) {
            out = 1;
            break;
          }
        if (!out) {
          cx[nc_] = dxe;
          cf[nc_] = shake_residue(S, &tmp, R);
          cabs[nc_] = x0 + dxe;
          nc_++;
        }
      }
    }
    double coef[3];
    polyfit2(cx, cf, nc_, coef);
    if (allow_xhat && coef[2] > 0.0) {
      const double xh = -coef[1] / (2.0 * coef[2]);
      double xmin = cx[0], xmax = cx[0];
      for (int t = 0; t < nc_; ++t) {
        xmin = fmin(xmin, cx[t]);
        xmax = fmax(xmax, cx[t]);
      }
      if (xh >= xmin && xh <= xmax) {
        tmp.X[ax] = x0 + xh;
        sobj_project(S, &tmp);
        int out = 0;
        for (int k = 0; k < nc; ++k)
          if (!in_roi_d(&R->roi[k], tmp.uv[k][0], tmp.uv[k][1])) {
            out = 1;
            break;
          }
        if (!out) {
          cx[nc_] = xh;
          cf[nc_] = shake_residue(S, &tmp, R);
          cabs[nc_] = x0 + xh;
          nc_++;
        }
      }
    }
    int bi = 0;
    double bf = cf[0];
    for (int t = 1; t < nc_; ++t)
      if (cf[t] < bf) {
        bf = cf[t];
        bi = t;
      }
    obj->X[ax] = cabs[bi];
    sobj_project(S, obj);
  }
  sobj_project(S, obj);
}

/* Shake::calObjectScore (Shake.cpp:668-782) */
static double object_score(const shctx *S, const sobj *o, roiinfo *R,
                           const double *res) {
  const int nc = S->n_cam;
  const size_t HW = (size_t)S->H * S->W;
  prange rect[OLPT_MAX_CAM];
  int use[OLPT_MAX_CAM] = {0};
  double peak[OLPT_MAX_CAM] = {0};
  int n_used = 0;
  for (int k = 0; k < nc; ++k) {
    if (!S->active[k]) continue;
    const double d = S->r_px > 0.0 ? S->r_px : 0.0;
    if (d <= 0.0) continue;
    rect[k] = roi_bound(S, k, o->uv[k][0], o->uv[k][1], d, d);
    if (pr_empty(rect[k])) continue;
    use[k] = 1;
    ++n_used;
    for (int row = rect[k].rmin; row < rect[k].rmax; ++row)
      for (int col = rect[k].cmin; col < rect[k].cmax; ++col) {
        const double v = roi_in(&R->roi[k], row, col)
                             ? *roi_at(R, k, row, col)
                             : res[HW * k + (size_t)row * S->W + col];
        if (v > peak[k]) peak[k] = v;
      }
  }
  if (n_used < 2) return 1e-12;
  if (n_used > 2) {
    int mi = -1;
    double mv = -INFINITY;
    for (int k = 0; k < nc; ++k)
      if (use[k] && peak[k] > mv) {
        mv = peak[k];
        mi = k;
      }
    if (mi >= 0) {
      use[mi] = 0;
      --n_used;
    }
  }
  double tn = 0.0, td = 0.0;
  for (int k = 0; k < nc; ++k) {
    if (!use[k]) continue;
    double num = 0.0, den = 0.0;
    for (int row = rect[k].rmin; row < rect[k].rmax; ++row)
      for (int col = rect[k].cmin; col < rect[k].cmax; ++col) {
        double m = roi_in(&R->roi[k], row, col)
                       ? *roi_at(R, k, row, col)
                       : res[HW * k + (size_t)row * S->W + col];
        m = m > 0.0 ? m : 0.0;
        num += m;
        den += project_int(S, o, k, row, col);
      }
    tn += num;
    td += den;
  }
  double ratio = (td > 1e-12) ? (tn / td) : 1.0;
  ratio = fmax(2.0 / 3.0, fmin(1.5, ratio));
  return sqrt(ratio);
}

/* Shake::markRepeatedObj (Shake.cpp:785-823): greedy in index order over a
 * strict radius search (nanoflann RadiusResultSet: dist < r^2). */
static void mark_repeated(const sobj *o, int n, double tol, uint8_t *rep) {
  memset(rep, 0, (size_t)n);
  const double t2 = tol * tol;
  for (int i = 0; i < n - 1; ++i) {
    if (rep[i]) continue;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double d0 = o[i].X[0] - o[j].X[0];
      const double d1 = o[i].X[1] - o[j].X[1];
      const double d2 = o[i].X[2] - o[j].X[2];
      const double d = d0 * d0 + d1 * d1 + d2 * d2;
      if (d < t2) rep[j] = 1;
    }
  }
}

/* Shake::runShake (Shake.cpp:36-163).  objs are updated in place; res is an
 * n_cam*H*W work image. */
static int run_shake(const shctx *S, const olpt_ipr_params *p, sobj *objs,
                     int n, const double *orig, double *res, uint8_t *flags,
                     double *score) {
  for (int i = 0; i < n; ++i) score[i] = 1.0;
  double delta = p->shake_width;
  double dmin = p->shakewidth_min;
  dmin = (dmin > 0) ? dmin : delta / 20;
  for (int i = 0; i < n; ++i) sobj_project(S, &objs[i]);
  for (int it = 0; it < p->n_shake_loop; ++it) {
    residue_image(S, objs, n, NULL, 0, orig, res);
    /* selectShakeCam returns all cameras for tracers (Shake.h:149-153), so
     * the "fewer than 2 shake cameras" branch needs n_cam < 2. */
    if (S->n_cam < 2) {
      for (int i = 0; i < n; ++i) score[i] = 0.0;
    } else {
#pragma omp parallel for schedule(dynamic, 64)
      for (int i = 0; i < n; ++i) {
        roiinfo R;
        build_roi(S, &objs[i], res, &R);
        shake_one(S, &objs[i], &R, delta);
        score[i] *= object_score(S, &objs[i], &R, res);
        free_roi(S, &R);
      }
    }
    delta *= 0.5;
    if (delta < dmin) delta = dmin;
  }
  uint8_t *rep = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
  mark_repeated(objs, n, S->tol3d, rep);
  double sum = 0.0;
  size_t cnt = 0;
  for (int i = 0; i < n; ++i) {
    flags[i] = rep[i] ? OLPT_FLAG_REPEATED : OLPT_FLAG_NONE;
    if (rep[i]) continue;
    sum += score[i];
    ++cnt;
  }
  const double mean = cnt ? sum / cnt : 0.0;
  for (int i = 0; i < n; ++i)
    if (score[i] < mean * p->thred_ghost || score[i] == 0.0) flags[i] |= OLPT_FLAG_GHOST;
  free(rep);
  return OK;
}

static void shctx_init(shctx *S, const olpt_camera *cams, int n_cam,
                       const int32_t *active, int H, int W,
                       const olpt_ipr_params *p, const olpt_otf *otf) {
  memset(S, 0, sizeof(*S));
  S->cams = cams;
  S->n_cam = n_cam;
  for (int k = 0; k < n_cam; ++k) S->active[k] = active ? active[k] != 0 : 1;
  S->H = H;
  S->W = W;
  S->otf = otf;
  S->r_px = p->radius_obj;
  S->ratio = p->ratio_augimg;
  S->tol3d = p->tol_3d_mm;
}

/* ------------------------------------------------------------------------ */
/* IPR (IPR.cpp)                                                            */
/* ------------------------------------------------------------------------ */

/* myMATH::generateCombinations (myMATH.cpp:39-71), lexicographic */
static int next_combination(int *c, int K, int N) {
  int i = K - 1;
  while (i >= 0 && c[i] == N - K + i) --i;
  if (i < 0) return 0;
  c[i]++;
  for (int j = i + 1; j < K; ++j) c[j] = c[j - 1] + 1;
  return 1;
}

/* limit2DObjectNumber (IPR.cpp:19-39): each active camera's list longer
 * than cap is shuffled with one std::default_random_engine(1234) shared
 * across the cameras (in camera order) and cut to cap.  The engine and
 * std::shuffle are restated from GCC 11's libstdc++, the reference's
 * toolchain: minstd_rand0 (x <- 16807 x mod 2^31-1), uniform_int_
 * distribution by downscaling with two divisions, and shuffle's two-
 * positions-per-draw path while n^2 fits the engine's range. */
typedef struct { uint64_t x; } lcg0;

static uint64_t lcg0_next(lcg0 *g) {
  g->x = (g->x * 16807u) % 2147483647u;
  return g->x;
}

static uint64_t lcg0_upto(lcg0 *g, uint64_t b) {
  const uint64_t R = 2147483645u;
  if (b == R) return lcg0_next(g) - 1;
  const uint64_t er = b + 1, sc = R / er, past = er * sc;
  uint64_t r;
  do r = lcg0_next(g) - 1; while (r >= past);
  return r / sc;
}

static void swap_pt(double *xy, int64_t i, int64_t j) {
  const double x = xy[2 * i], y = xy[2 * i + 1];
  xy[2 * i] = xy[2 * j];
  xy[2 * i + 1] = xy[2 * j + 1];
  xy[2 * j] = x;
  xy[2 * j + 1] = y;
}

static void shuffle_pts(double *xy, int64_t n, lcg0 *g) {
  const uint64_t R = 2147483645u;
  if (n <= 0) return;
  if (R / (uint64_t)n >= (uint64_t)n) {
    int64_t i = 1;
    if (n % 2 == 0) swap_pt(xy, i++, (int64_t)lcg0_upto(g, 1));
    for (; i != n; i += 2) {
      const uint64_t m = (uint64_t)i + 2, x = lcg0_upto(g, (m - 1) * m - 1);
      swap_pt(xy, i, (int64_t)(x / m));
      swap_pt(xy, i + 1, (int64_t)(x % m));
    }
    return;
  }
  for (int64_t i = 1; i < n; ++i) swap_pt(xy, i, (int64_t)lcg0_upto(g, (uint64_t)i));
}

/* runSingleIPRIteration (IPR.cpp:70-193).  images (n_cam*H*W) is replaced
 * by the residual; kept objects are appended to (xyz, uv). */
static int ipr_iteration(const olpt_camera *cams, int n_cam,
                         const int32_t *active, double *images, int H, int W,
                         const olpt_ipr_params *p, const olpt_otf *otf,
                         dvec *out_xyz, dvec *out_uv) {
  const size_t HW = (size_t)H * W;
  dvec pk[OLPT_MAX_CAM];
  memset(pk, 0, sizeof(pk));
  int32_t npts[OLPT_MAX_CAM] = {0};
  int rc = OK;
  for (int c = 0; c < n_cam; ++c) {
    if (!active[c]) continue;
    rc = find_peaks(images + HW * c, H, W, (double)p->min_obj_int, &pk[c]);
    if (rc) return rc;
    npts[c] = pk[c].n / 2;
  }
  if (p->n_obj2d_process_max > 0) {
    lcg0 g = {1234};
    for (int c = 0; c < n_cam; ++c)
      if (active[c] && npts[c] > p->n_obj2d_process_max) {
        shuffle_pts(pk[c].v, npts[c], &g);
        npts[c] = p->n_obj2d_process_max;
      }
  }
  int tot = 0;
  for (int c = 0; c < n_cam; ++c) tot += npts[c];
  double *xy = (double *)malloc(sizeof(double) * 2 * (size_t)(tot > 0 ? tot : 1));
  int off = 0;
  for (int c = 0; c < n_cam; ++c) {
    if (npts[c]) memcpy(xy + 2 * off, pk[c].v, sizeof(double) * 2 * npts[c]);
    off += npts[c];
    free(pk[c].v);
  }
  smctx SM;
  rc = sm_init(&SM, cams, n_cam, active, xy, npts, p);
  match_out mo;
  if (!rc) rc = stereo_match(&SM, &mo);
  sm_free(&SM);
  if (rc) {
    free(xy);
    return rc;
  }
  const int n = mo.n_obj;
  if (n > 0) {
    shctx S;
    shctx_init(&S, cams, n_cam, active, H, W, p, otf);
    sobj *objs = (sobj *)calloc((size_t)n, sizeof(sobj));
    uint8_t *fl = (uint8_t *)malloc((size_t)n);
    double *sc = (double *)malloc(sizeof(double) * (size_t)n);
    double *res = (double *)malloc(sizeof(double) * HW * n_cam);
    for (int i = 0; i < n; ++i) memcpy(objs[i].X, &mo.xyz[3 * i], sizeof(objs[i].X));
    run_shake(&S, p, objs, n, images, res, fl, sc);
    /* filterOutInvalid (IPR.cpp:41-63) then calResidualImage (IPR mode) */
    int nk = 0;
    for (int i = 0; i < n; ++i)
      if (!(fl[i] & (OLPT_FLAG_GHOST | OLPT_FLAG_REPEATED))) objs[nk++] = objs[i];
    residue_image(&S, objs, nk, NULL, 1, images, res);
    memcpy(images, res, sizeof(double) * HW * n_cam);
    for (int i = 0; i < nk; ++i) {
      for (int k = 0; k < 3; ++k) dvec_push(out_xyz, objs[i].X[k]);
      for (int c = 0; c < n_cam; ++c) {
        dvec_push(out_uv, objs[i].uv[c][0]);
        dvec_push(out_uv, objs[i].uv[c][1]);
      }
    }
    free(objs); free(fl); free(sc); free(res);
  }
  match_free(&mo);
  free(xy);
  return OK;
}

/* ------------------------------------------------------------------------ */
/* extern "C" entry points                                                  */
/* ------------------------------------------------------------------------ */
int orc_project(const olpt_camera *cam, const double *xyz, int n, double *uv) {
  for (int i = 0; i < n; ++i) project(cam, &xyz[3 * i], &uv[2 * i]);
  return OK;
}

int orc_line_of_sight(const olpt_camera *cam, const double *uv, int n,
                      double *out) {
  for (int i = 0; i < n; ++i) {
    line3 L;
    line_of_sight(cam, uv[2 * i], uv[2 * i + 1], &L);
    memcpy(&out[6 * i], L.pt, sizeof(L.pt));
    memcpy(&out[6 * i + 3], L.u, sizeof(L.u));
  }
  return OK;
}

int orc_triangulate(const double *lines, int n_line, double *xyz, double *err) {
  if (n_line < 2 || n_line > OLPT_MAX_CAM) return ERR_ARG;
  line3 L[OLPT_MAX_CAM];
  const line3 *P[OLPT_MAX_CAM];
  for (int i = 0; i < n_line; ++i) {
    memcpy(L[i].pt, &lines[6 * i], sizeof(L[i].pt));
    memcpy(L[i].u, &lines[6 * i + 3], sizeof(L[i].u));
    P[i] = &L[i];
  }
  return triangulate(P, n_line, xyz, err) == 0 ? OK : -1;
}

int orc_polyfit(const double *x, const double *y, int n, int order,
                double *coeff) {
  if (order < 1) return ERR_ARG;
  if (order == 2)
    polyfit2(x, y, n, coeff);
  else
    polyfit_n(x, y, n, order, coeff);
  return OK;
}

int orc_trilinear(const double *lim, const double *v, const double *p, double *out) {
  *out = trilinear(lim, v, p);
  return OK;
}

int orc_is_local_max(const double *img, int H, int W, int row, int col) {
  return is_local_max(img, H, W, row, col);
}

int orc_otf_param(const olpt_otf *otf, int cam_id, const double *xyz, int n,
                  double *out) {
  for (int i = 0; i < n; ++i) otf_param(otf, cam_id, &xyz[3 * i], &out[5 * i]);
  return OK;
}

int orc_find_peaks(const double *img, int H, int W, const olpt_ipr_params *p,
                   int cap, double *xy, int32_t *n_out) {
  dvec v = {0};
  int rc = find_peaks(img, H, W, (double)p->min_obj_int, &v);
  if (rc) return rc;
  *n_out = v.n / 2;
  if (v.n / 2 > cap) {
    free(v.v);
    return ERR_CAP;
  }
  if (v.n) memcpy(xy, v.v, sizeof(double) * v.n);
  free(v.v);
  return OK;
}

int orc_stereo_match(const olpt_camera *cams, int n_cam, const int32_t *active,
                     const double *xy, const int32_t *n_pts,
                     const olpt_ipr_params *p, int cap_cand, int32_t *cand_ids,
                     double *cand_e, int32_t *sel_idx, int cap_obj,
                     double *obj_xyz, int32_t *obj_ids, int32_t *counts) {
  smctx S;
  int rc = sm_init(&S, cams, n_cam, active, xy, n_pts, p);
  match_out mo;
  memset(&mo, 0, sizeof(mo));
  if (!rc) rc = stereo_match(&S, &mo);
  sm_free(&S);
  if (rc) {
    match_free(&mo);
    return rc;
  }
  int ns = 0;
  for (int j = 0; j < mo.M; ++j) ns += mo.sel[j];
  counts[0] = mo.M;
  counts[1] = ns;
  counts[2] = mo.n_obj;
  if (mo.M > cap_cand || mo.n_obj > cap_obj) {
    match_free(&mo);
    return ERR_CAP;
  }
  memcpy(cand_ids, mo.ids, sizeof(int32_t) * (size_t)mo.M * n_cam);
  memcpy(cand_e, mo.e, sizeof(double) * (size_t)mo.M);
  int s = 0;
  for (int j = 0; j < mo.M; ++j)
    if (mo.sel[j]) sel_idx[s++] = j;
  memcpy(obj_xyz, mo.xyz, sizeof(double) * 3 * (size_t)mo.n_obj);
  memcpy(obj_ids, mo.oids, sizeof(int32_t) * (size_t)mo.n_obj * n_cam);
  match_free(&mo);
  return OK;
}

/* StereoMatch::pruneMatch alone (StereoMatch.cpp:631-980) on given
 * candidates: ids [M][n_cam] (-1 = camera unused), e [M]; n_pts[c] 2D
 * points per camera; sel[j] = 1 for the selected candidates. */
int orc_prune(int n_cam, const int32_t *active, const int32_t *n_pts, const int32_t *ids,
              const double *e, int M, uint8_t *sel) {
  smctx S;
  memset(&S, 0, sizeof(S));
  if (n_cam < 1 || n_cam > OLPT_MAX_CAM) return ERR_ARG;
  S.n_cam = n_cam;
  for (int c = 0; c < n_cam; ++c) {
    S.active[c] = active ? active[c] != 0 : 1;
    S.npts[c] = n_pts[c];
  }
  return prune_match(&S, ids, e, M, sel);
}

int orc_shake(const olpt_camera *cams, int n_cam, const int32_t *active,
              const double *imgs, int H, int W, const olpt_ipr_params *p,
              const olpt_otf *otf, const double *xyz_in, int n_obj,
              double *xyz_out, uint8_t *flags, double *score, double *res_out) {
  if (n_cam < 1 || n_cam > OLPT_MAX_CAM) return ERR_ARG;
  shctx S;
  shctx_init(&S, cams, n_cam, active, H, W, p, otf);
  sobj *objs = (sobj *)calloc((size_t)(n_obj > 0 ? n_obj : 1), sizeof(sobj));
  double *res = (double *)malloc(sizeof(double) * (size_t)H * W * n_cam);
  for (int i = 0; i < n_obj; ++i) memcpy(objs[i].X, &xyz_in[3 * i], sizeof(objs[i].X));
  run_shake(&S, p, objs, n_obj, imgs, res, flags, score);
  for (int i = 0; i < n_obj; ++i) memcpy(&xyz_out[3 * i], objs[i].X, sizeof(objs[i].X));
  if (res_out) {
    int nk = 0;
    for (int i = 0; i < n_obj; ++i)
      if (flags[i] == OLPT_FLAG_NONE) objs[nk++] = objs[i];
    residue_image(&S, objs, nk, NULL, 1, imgs, res_out);
  }
  free(objs);
  free(res);
  return OK;
}

int orc_residual(const olpt_camera *cams, int n_cam, const double *imgs, int H,
                 int W, const olpt_ipr_params *p, const olpt_otf *otf,
                 const double *xyz, int n_obj, double *res_out) {
  shctx S;
  shctx_init(&S, cams, n_cam, NULL, H, W, p, otf);
  sobj *objs = (sobj *)calloc((size_t)(n_obj > 0 ? n_obj : 1), sizeof(sobj));
  for (int i = 0; i < n_obj; ++i) {
    memcpy(objs[i].X, &xyz[3 * i], sizeof(objs[i].X));
    sobj_project(&S, &objs[i]);
  }
  residue_image(&S, objs, n_obj, NULL, 1, imgs, res_out);
  free(objs);
  return OK;
}

int orc_run_ipr(const olpt_camera *cams, int n_cam, const double *imgs, int H,
                int W, const olpt_ipr_params *p, const olpt_otf *otf, int cap,
                double *obj_xyz, double *obj_uv, int cap_loop,
                int32_t *loop_total, int32_t *counts) {
  counts[0] = counts[1] = 0;
  if (n_cam < 2) return OK; /* IPR.cpp:200-204 */
  if (n_cam > OLPT_MAX_CAM) return ERR_ARG;
  const size_t HW = (size_t)H * W;
  double *images = (double *)malloc(sizeof(double) * HW * n_cam);
  if (!images) return ERR_ALLOC;
  memcpy(images, imgs, sizeof(double) * HW * n_cam);
  olpt_ipr_params q = *p;
  dvec xyz = {0}, uv = {0};
  int n_loop = 0, rc = OK;
  int max_drop = n_cam - 2 > 0 ? n_cam - 2 : 0;
  if (p->n_cam_reduced < max_drop) max_drop = p->n_cam_reduced;
  const int orig_mcc = p->match_cam_count;
  for (int drop = 0; drop <= max_drop && !rc; ++drop) {
    const int K = n_cam - drop;
    const int loops = drop == 0 ? p->n_loop_ipr : p->n_loop_ipr_reduced;
    q.match_cam_count = orig_mcc < K ? orig_mcc : K;
    int comb[OLPT_MAX_CAM];
    for (int i = 0; i < K; ++i) comb[i] = i;
    do {
      int32_t act[OLPT_MAX_CAM] = {0};
      for (int i = 0; i < K; ++i) act[comb[i]] = 1;
      const double tol_orig = q.tol_2d_px;
      for (int loop = 0; loop < loops && !rc; ++loop) {
        rc = ipr_iteration(cams, n_cam, act, images, H, W, &q, otf, &xyz, &uv);
        if (n_loop < cap_loop) loop_total[n_loop] = xyz.n / 3;
        ++n_loop;
        q.tol_2d_px = q.tol_2d_px * 1.5;
      }
      q.tol_2d_px = tol_orig;
    } while (!rc && next_combination(comb, K, n_cam));
  }
  free(images);
  counts[0] = xyz.n / 3;
  counts[1] = n_loop;
  if (!rc && xyz.n / 3 > cap) rc = ERR_CAP;
  if (!rc) {
    memcpy(obj_xyz, xyz.v, sizeof(double) * xyz.n);
    memcpy(obj_uv, uv.v, sizeof(double) * uv.n);
  }
  free(xyz.v);
  free(uv.v);
  return rc;
}


/* ======================================================================
 * Tracking (STB) pieces
 * ==================================================================== */

/* STB::predictNext (STB.cpp:233-316) for tracks of len[i] points stored at
 * pos[i][0..len[i]) (row stride `stride` points): normalised LMS fit of the
 * last sample from the previous `order` samples, iterated while
 * |error| > 1e-8 (SMALLNUMBER) for at most 5000 steps (WIENER_MAX_ITER),
 * then the one-step-ahead prediction.  ok[i] = 0 for tracks under 3 points
 * (the reference returns null). */
int orc_stb_predict(const double *pos, const int32_t *len, int n_tr, int stride, double *out,
                    int32_t *ok) {
  for (int i = 0; i < n_tr; ++i) {
    const int n = len[i];
    if (n < 3) {
      ok[i] = 0;
      out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0.0;
      continue;
    }
    ok[i] = 1;
    const int order = n < 6 ? n - 1 : 5;
    const int start = n - 1 - order;
    for (int axis = 0; axis < 3; ++axis) {
      double series[6], filter[5];
      for (int j = 0; j <= order; ++j)
        series[j] = pos[((size_t)i * stride + start + j) * 3 + axis];
      int shifted = 0;
      if (fabs(series[order]) < 1.0) {
        shifted = 1;
        for (int j = 0; j <= order; ++j) series[j] += 10.0;
      }
      double denom = 0.0;
      for (int j = 0; j < order; ++j) denom += series[j] * series[j];
      const double step = (denom > 1e-12) ? (1.0 / denom) : 0.0;
      for (int j = 0; j < order; ++j) filter[j] = 0.0;
      double prediction = 0.0;
      double error = series[order] - prediction;
      int iter = 0;
      while (fabs(error) > 1e-8 && iter < 5000 && step > 0.0) {
        prediction = 0.0;
        for (int j = 0; j < order; ++j) {
          filter[j] += step * series[j] * error;
          prediction += filter[j] * series[j];
        }
        error = series[order] - prediction;
        ++iter;
      }
      if (order >= 1 && step == 0.0) {
        prediction = series[order] + (series[order] - series[order - 1]);
      } else {
        prediction = 0.0;
        for (int j = 0; j < order; ++j) prediction += filter[j] * series[j + 1];
      }
      if (shifted) prediction -= 10.0;
      out[3 * i + axis] = prediction;
    }
  }
  return 0;
}

/* STB::checkLinearFit (STB.cpp:874-924): a straight-line least-squares fit
 * of the last min(n_init, 4) points per axis through the Matrix products
 * inverse(X'X) X' y (Matrix.hpp:392-408 summation order); the track passes
 * when the fitted last point lies within 5e-2 mm (MAX_ERR_LINEARFIT) of the
 * measured one. */
int orc_stb_linear_fit(int n_init, const double *pos, const int32_t *len, int n_tr, int stride,
                       int32_t *ok) {
  const int n_pts = n_init > 4 ? 4 : n_init;
  double X[8], Xt[8], XtX[4] = {0, 0, 0, 0}, inv[4], K[8];
  for (int i = 0; i < n_pts; ++i) {
    X[2 * i] = 1;
    X[2 * i + 1] = i;
    Xt[i] = 1;
    Xt[n_pts + i] = i;
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < n_pts; ++k) XtX[2 * a + b] += Xt[a * n_pts + k] * X[2 * k + b];
  gauss_inverse_n(XtX, 2, inv);
  for (int a = 0; a < 2; ++a)
    for (int j = 0; j < n_pts; ++j) {
      K[a * n_pts + j] = 0.0;
      for (int k = 0; k < 2; ++k) K[a * n_pts + j] += inv[2 * a + k] * Xt[k * n_pts + j];
    }
  for (int t = 0; t < n_tr; ++t) {
    const int L = len[t];
    if (L < n_pts) {
      ok[t] = 0;
      continue;
    }
    double est[3];
    for (int a = 0; a < 3; ++a) {
      double c0 = 0.0, c1 = 0.0;
      for (int j = 0; j < n_pts; ++j)
        c0 += K[j] * pos[((size_t)t * stride + L - n_pts + j) * 3 + a];
      for (int j = 0; j < n_pts; ++j)
        c1 += K[n_pts + j] * pos[((size_t)t * stride + L - n_pts + j) * 3 + a];
      est[a] = c0 + (n_pts - 1) * c1;
    }
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a)
      d2 += pow(est[a] - pos[((size_t)t * stride + L - 1) * 3 + a], 2);
    ok[t] = sqrt(d2) > 5e-2 ? 0 : 1;
  }
  return 0;
}

/* PredField::fitPeakBinLocGauss (PredField.cpp:345-371) */
static double pf_fit_gauss(double y1, double v1, double y2, double v2, double y3, double v3) {
  const double l1 = v1 < LOGSMALLNUMBER ? log(LOGSMALLNUMBER) : log(v1);
  const double l2 = v2 < LOGSMALLNUMBER ? log(LOGSMALLNUMBER) : log(v2);
  const double l3 = v3 < LOGSMALLNUMBER ? log(LOGSMALLNUMBER) : log(v3);
  double yc = -0.5 * ((l1 * (double)((y2 * y2) - (y3 * y3))) - (l2 * (double)((y1 * y1) - (y3 * y3))) +
                      (l3 * (double)((y1 * y1) - (y2 * y2)))) /
              ((l1 * (double)(y3 - y2)) - (l3 * (double)(y1 - y2)) + (l2 * (double)(y1 - y3)));
  if (!isfinite(yc)) yc = y2;
  return yc;
}

/* PredField::calPredField (PredField.cpp:21-86) + smoothDispField
 * (:374-485), then getDisp (:91-165) at n_q points.  lim = view volume,
 * grid nx x ny x nz (node id x*ny*nz + y*nz + z), half-width r, nbin bins
 * per axis; field [nx*ny*nz][3]. */
int orc_pred_field(const double *lim, int nx, int ny, int nz, double r, const int32_t *nbin,
                   int smooth, const double *sigma, const double *prev, int n_prev,
                   const double *curr, int n_curr, double *field, const double *q, int n_q,
                   double *disp) {
  const int ng = nx * ny * nz;
  const int nb = nbin[0] * nbin[1] * nbin[2];
  const int nbv[3] = {nbin[0], nbin[1], nbin[2]};
  const int nn[3] = {nx, ny, nz};
  double m[3], c[3];
  for (int a = 0; a < 3; ++a) {
    m[a] = (nbv[a] - 1) / (4 * r);
    c[a] = (double)((nbv[a] - 1) / 2);
  }
  double *pdf = (double *)malloc(sizeof(double) * (size_t)nb);
  int *lp = (int *)malloc(sizeof(int) * (size_t)(n_prev > 0 ? n_prev : 1));
  int *lc = (int *)malloc(sizeof(int) * (size_t)(n_curr > 0 ? n_curr : 1));
  if (!pdf || !lp || !lc) return 1;
  for (int xi = 0; xi < nx; ++xi)
    for (int yi = 0; yi < ny; ++yi)
      for (int zi = 0; zi < nz; ++zi) {
        const int g = xi * ny * nz + yi * nz + zi;
        const double node[3] = {linspace_at(lim[0], lim[1], nx, xi),
                                linspace_at(lim[2], lim[3], ny, yi),
                                linspace_at(lim[4], lim[5], nz, zi)};
        int np = 0, nc = 0;
        for (int i = 0; i < n_prev; ++i) {
          const double dx = prev[3 * i] - node[0], dy = prev[3 * i + 1] - node[1],
                       dz = prev[3 * i + 2] - node[2];
          if (dx < r && dx > -r && dy < r && dy > -r && dz < r && dz > -r) lp[np++] = i;
        }
        for (int i = 0; i < n_curr; ++i) {
          const double dx = curr[3 * i] - node[0], dy = curr[3 * i + 1] - node[1],
                       dz = curr[3 * i + 2] - node[2];
          if (dx < r && dx > -r && dy < r && dy > -r && dz < r && dz > -r) lc[nc++] = i;
        }
        if (nc == 0 || np == 0) {
          field[3 * g] = field[3 * g + 1] = field[3 * g + 2] = 0;
          continue;
        }
        for (int b = 0; b < nb; ++b) pdf[b] = 0;
        for (int j = 0; j < nc; ++j)
          for (int k = 0; k < np; ++k) {
            double d[3], bin[3];
            int br[3];
            for (int a = 0; a < 3; ++a) {
              d[a] = curr[3 * lc[j] + a] - prev[3 * lp[k] + a];
              bin[a] = m[a] * d[a] + c[a];
              br[a] = (int)round(bin[a]);
            }
            for (int i = br[0] - 1 < 0 ? 0 : br[0] - 1; i <= (nbv[0] - 1 < br[0] + 1 ? nbv[0] - 1 : br[0] + 1); ++i) {
              const double xx = i - bin[0];
              for (int jj = br[1] - 1 < 0 ? 0 : br[1] - 1; jj <= (nbv[1] - 1 < br[1] + 1 ? nbv[1] - 1 : br[1] + 1); ++jj) {
                const double yy = jj - bin[1];
                for (int kk = br[2] - 1 < 0 ? 0 : br[2] - 1; kk <= (nbv[2] - 1 < br[2] + 1 ? nbv[2] - 1 : br[2] + 1); ++kk) {
                  const double zz = kk - bin[2];
                  pdf[(i * nbv[1] + jj) * nbv[2] + kk] += exp(-(xx * xx + yy * yy + zz * zz));
                }
              }
            }
          }
        int best = 0;
        for (int b = 1; b < nb; ++b)
          if (pdf[b] > pdf[best]) best = b;
        int pk[3] = {-1, -1, -1};
        if (pdf[best] > 1e-1) {
          pk[0] = best / (nbv[1] * nbv[2]);
          pk[1] = (best - pk[0] * nbv[1] * nbv[2]) / nbv[2];
          pk[2] = best % nbv[2];
        }
        for (int a = 0; a < 3; ++a) {
          double v;
          if (pk[a] < 0) v = 0.0;
          else if (pk[a] == 0) v = -2 * r;
          else if (pk[a] == nbv[a] - 1) v = 2 * r;
          else {
            int lo[3] = {pk[0], pk[1], pk[2]}, hi[3] = {pk[0], pk[1], pk[2]};
            lo[a] -= 1;
            hi[a] += 1;
            const double loc = pf_fit_gauss(pk[a] - 1, pdf[(lo[0] * nbv[1] + lo[1]) * nbv[2] + lo[2]],
                                            pk[a], pdf[(pk[0] * nbv[1] + pk[1]) * nbv[2] + pk[2]],
                                            pk[a] + 1, pdf[(hi[0] * nbv[1] + hi[1]) * nbv[2] + hi[2]]);
            v = (loc - c[a]) / m[a];
          }
          field[3 * g + a] = v;
        }
      }
  free(pdf);
  free(lp);
  free(lc);
  if (smooth) {
    double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)ng);
    if (!tmp) return 1;
    for (int a = 0; a < 3; ++a) {
      const int rad = (int)round(3 * sigma[a]);
      double kern[64], sum = 0.0;
      if (2 * rad + 1 > 64) {
        free(tmp);
        return 1;
      }
      for (int i = 0; i < 2 * rad + 1; ++i) {
        const int x = i - rad;
        kern[i] = exp(-(x * x) / (2 * sigma[a] * sigma[a]));
        sum += kern[i];
      }
      for (int i = 0; i < 2 * rad + 1; ++i) kern[i] /= sum;
      for (int dir = 0; dir < 3; ++dir) {
        memcpy(tmp, field, sizeof(double) * 3 * (size_t)ng);
        for (int x = 0; x < nx; ++x)
          for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
              const int p3[3] = {x, y, z};
              double acc = 0.0;
              for (int k = -rad; k <= rad; ++k) {
                const int idx = p3[dir] + k;
                if (idx >= 0 && idx < nn[dir]) {
                  const int sx = dir == 0 ? idx : x, sy = dir == 1 ? idx : y, sz = dir == 2 ? idx : z;
                  acc += tmp[3 * (sx * ny * nz + sy * nz + sz) + a] * kern[k + rad];
                }
              }
              field[3 * (x * ny * nz + y * nz + z) + a] = acc;
            }
      }
    }
    free(tmp);
  }
  const double dx = (lim[1] - lim[0]) / (nx - 1), dy = (lim[3] - lim[2]) / (ny - 1),
               dz = (lim[5] - lim[4]) / (nz - 1);
  for (int i = 0; i < n_q; ++i) {
    const double px = fmin(fmax(q[3 * i], lim[0]), lim[1]);
    const double py = fmin(fmax(q[3 * i + 1], lim[2]), lim[3]);
    const double pz = fmin(fmax(q[3 * i + 2], lim[4]), lim[5]);
    int xi = (int)floor((px - lim[0]) / dx), yi = (int)floor((py - lim[2]) / dy),
        zi = (int)floor((pz - lim[4]) / dz);
    xi = xi < 0 ? 0 : xi;
    yi = yi < 0 ? 0 : yi;
    zi = zi < 0 ? 0 : zi;
    xi = xi > nx - 2 ? nx - 2 : xi;
    yi = yi > ny - 2 ? ny - 2 : yi;
    zi = zi > nz - 2 ? nz - 2 : zi;
    const double cell[6] = {linspace_at(lim[0], lim[1], nx, xi), linspace_at(lim[0], lim[1], nx, xi + 1),
                            linspace_at(lim[2], lim[3], ny, yi), linspace_at(lim[2], lim[3], ny, yi + 1),
                            linspace_at(lim[4], lim[5], nz, zi), linspace_at(lim[4], lim[5], nz, zi + 1)};
#define G3(a, b, cc) ((a) * ny * nz + (b) * nz + (cc))
    const int ids[8] = {G3(xi, yi, zi),         G3(xi + 1, yi, zi),     G3(xi + 1, yi, zi + 1),
                        G3(xi, yi, zi + 1),     G3(xi, yi + 1, zi),     G3(xi + 1, yi + 1, zi),
                        G3(xi + 1, yi + 1, zi + 1), G3(xi, yi + 1, zi + 1)};
#undef G3
    const double pt[3] = {px, py, pz};
    for (int a = 0; a < 3; ++a) {
      double v[8];
      for (int k = 0; k < 8; ++k) v[k] = field[3 * ids[k] + a];
      disp[3 * i + a] = trilinear(cell, v, pt);
    }
  }
  return 0;
}
