// olpt_gi3.hpp - 3x3 inverse of myMATH::gaussInverse (myMATH.h:183-289):
// LU with partial pivoting scaled by each row's signed maximum, then forward
// and back substitution per column.  Rows are swapped in registers instead
// of through a permutation index, which performs the same floating-point
// operations in the same order and keeps the matrix out of scratch memory.
#pragma once

#ifdef __HIPCC__
#define OLPT_HD __host__ __device__ __forceinline__
#else
#define OLPT_HD inline
#include <cmath>
using std::fabs;
#endif

namespace olpt {

OLPT_HD void gi3_swap(double &a, double &b, bool s) {
  const double x = s ? b : a, y = s ? a : b;
  a = x;
  b = y;
}

OLPT_HD void gauss_inverse3(const double A[9], double out[9]) {
  double u[3][3], mr[3], bm[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double m = A[3 * i];
    if (A[3 * i + 1] > m) m = A[3 * i + 1];
    if (A[3 * i + 2] > m) m = A[3 * i + 2];
    mr[i] = m;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      u[i][j] = A[3 * i + j];
      bm[i][j] = i == j ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int y = 0; y < 2; ++y) {
    double mp = 0;
    int t = 0;
#pragma unroll
    for (int x = y; x < 3; ++x) {
      const double pv = fabs(u[x][y] / mr[x]);
      if (pv > mp) {
        mp = pv;
        t = x;
      }
    }
    // t stays 0 when no scaled pivot is positive (a singular column), so at
    // y = 1 the reference's index swap can pull back the first pivot row
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      if (x == y) continue;
      const bool s = t == x;
      gi3_swap(mr[y], mr[x], s);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        gi3_swap(u[y][j], u[x][j], s);
        gi3_swap(bm[y][j], bm[x][j], s);
      }
    }
#pragma unroll
    for (int x = y + 1; x < 3; ++x) {
      const double m = u[x][y] / u[y][y];
      u[x][y] = m;
#pragma unroll
      for (int it = y + 1; it < 3; ++it) u[x][it] -= m * u[y][it];
    }
  }
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    double bc[3] = {bm[0][col], bm[1][col], bm[2][col]};
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = y + 1; x < 3; ++x) bc[x] -= u[x][y] * bc[y];
    out[6 + col] = bc[2] / u[2][2];
#pragma unroll
    for (int x = 1; x > -1; --x) {
      double z = bc[x];
#pragma unroll
      for (int y = x + 1; y < 3; ++y) z -= u[x][y] * out[3 * y + col];
      out[3 * x + col] = z / u[x][x];
    }
  }
}

} // namespace olpt
