// tri_decide (olpt_device.hpp) against triangulate_n on the CPU: for random
// line sets aimed near a point of the 40 mm cube from cameras ~400 mm away at
// +-30 deg (the rig of the IPR workloads), with the tolerance placed within
// 1e-5 relative of the exact error (some exactly at it, some one ulp below),
// every decision tri_decide takes must equal triangulate_n(...) && e <= tq.
// Prints one line per line count: trials, undecided, wrong.
#include "olpt_device.hpp"

#include <cstdio>
#include <cstdlib>
#include <random>

using namespace olpt;

template <int N> void run(std::mt19937_64 &g, long trials, long &undec, long &wrong) {
  std::normal_distribution<double> nd(0, 1);
  std::uniform_real_distribution<double> ud(-1, 1);
  for (long t = 0; t < trials; ++t) {
    const double X0[3] = {20 * ud(g), 20 * ud(g), 20 * ud(g)};
    double pt[N][3], u[N][3];
    for (int i = 0; i < N; ++i) {
      const double az = (i & 1 ? 30 : -30) * M_PI / 180, el = (i & 2 ? 30 : -30) * M_PI / 180;
      const double c[3] = {400 * cos(el) * sin(az), 400 * sin(el), 400 * cos(el) * cos(az)};
      const double sc = 0.05 * fabs(nd(g));
      double d[3];
      for (int k = 0; k < 3; ++k) {
        pt[i][k] = c[k];
        d[k] = X0[k] + sc * nd(g) - c[k];
      }
      const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int k = 0; k < 3; ++k) u[i][k] = d[k] / n;
    }
    double X[3], e;
    const bool ok = triangulate_n<N>(pt, u, X, e);
    double mx = -1;
    for (int i = 0; i < N; ++i)
      for (int j = i + 1; j < N; ++j) {
        const double dot = u[i][0] * u[j][0] + u[i][1] * u[j][1] + u[i][2] * u[j][2];
        mx = fmax(mx, fmin(1.0, fmax(-1.0, dot)));
      }
    double tq = e * (1.0 + 1e-5 * ud(g));
    if (t % 7 == 0) tq = e;
    if (t % 11 == 0) tq = std::nextafter(e, 0.0);
    if (t % 13 == 0) tq = e * (1 + 0.3 * ud(g));
    const int want = ok && e <= tq;
    const int d = tri_decide<N>(pt, u, tq, mx);
    if (d < 0)
      ++undec;
    else if (d != want)
      ++wrong;
  }
}

int main(int argc, char **argv) {
  const long trials = argc > 1 ? atol(argv[1]) : 200000;
  std::mt19937_64 g(12345);
  for (int n = 2; n <= 4; ++n) {
    long und = 0, bad = 0;
    if (n == 2) run<2>(g, trials, und, bad);
    if (n == 3) run<3>(g, trials, und, bad);
    if (n == 4) run<4>(g, trials, und, bad);
    printf("N=%d trials %ld undecided %ld wrong %ld\n", n, trials, und, bad);
  }
  return 0;
}
