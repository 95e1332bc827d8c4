// Host build of the eigen routines the gfx950 kernels use (qex_amd/csrc/band.h): Sturm-count
// bisection and LDL^T inverse iteration for the lowest k eigenpairs of T + diag(v).
// stdin: N k dx, then N potential values.  stdout: per eigenpair the eigenvalue (Rayleigh
// quotient) and the 2-normalised eigenvector.
#include <math.h>
#include <stdio.h>

#include <vector>

#include "band.h"

using namespace qexband;

int main() {
  int N, K;
  double dx;
  if (scanf("%d %d %lf", &N, &K, &dx) != 3) return 1;
  std::vector<double> a(N), y(N), F(4 * N);
  const double s = -0.5 / (dx * dx), Tb = s * (4.0 / 3.0), Tc = s * (-1.0 / 12.0), Td = s * -2.5;
  double amin = 1e300, amax = -1e300, aabs = 0.0;
  for (int i = 0; i < N; ++i) {
    double v;
    if (scanf("%lf", &v) != 1) return 2;
    a[i] = Td + v;
    amin = fmin(amin, a[i]);
    amax = fmax(amax, a[i]);
    aabs = fmax(aabs, fabs(a[i]));
  }
  const double off = 2.0 * (fabs(Tb) + fabs(Tc)), hn = aabs + off;
  for (int j = 0; j < K; ++j) {
    double lo = amin - off, hi = amax + off;
    for (int it = 0; it < 200 && hi - lo > 1e-9 * fmax(1.0, fabs(0.5 * (lo + hi))) + 64 * 2.2e-16 * hn; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (sturm_count(a.data(), N, Tb, Tc, mid) > j)
        hi = mid;
      else
        lo = mid;
    }
    const double sigma = 0.5 * (lo + hi);
    for (int i = 0; i < N; ++i) y[i] = sin(1.3 * i + 0.7 * j + 0.1);
    for (int it = 0; it < 2; ++it) {
      if (it == 0)
        ldlt_factor_forward(a.data(), y.data(), N, Tb, Tc, sigma, 2.2e-16 * hn, F.data());
      else
        ldlt_forward(y.data(), N, F.data());
      ldlt_backward(F.data(), N, y.data());
      double nr = 0.0;
      for (int i = 0; i < N; ++i) nr += y[i] * y[i];
      nr = 1.0 / sqrt(nr);
      for (int i = 0; i < N; ++i) y[i] *= nr;
    }
    double rq = 0.0;
    for (int i = 0; i < N; ++i) {
      double hx = a[i] * y[i];
      if (i >= 1) hx += Tb * y[i - 1];
      if (i + 1 < N) hx += Tb * y[i + 1];
      if (i >= 2) hx += Tc * y[i - 2];
      if (i + 2 < N) hx += Tc * y[i + 2];
      rq += y[i] * hx;
    }
    printf("%.17g\n", rq);
    for (int i = 0; i < N; ++i) printf("%.17g ", y[i]);
    printf("\n");
  }
  return 0;
}
