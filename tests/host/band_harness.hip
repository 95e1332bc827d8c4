// Host build of the eigen routines the gfx950 kernels use (qex_amd/csrc/band.h): Sturm-count
// bisection and LDL^T inverse iteration for the lowest k eigenpairs of T + diag(v).
// stdin: N k dx, then N potential values.  stdout: per eigenpair the eigenvalue (Rayleigh
// quotient) and the 2-normalised eigenvector.  With the argument "2" the two-sided routines (top
// and bottom halves meeting at the middle pair) are used, and a last line gives the number of
// shifts, of 400 across the spectrum, whose two-sided count differs from the one-sided count.
#include <math.h>
#include <stdio.h>

#include <vector>

#include "band.h"

using namespace qexband;

int main(int argc, char** argv) {
  const bool two = argc > 1 && argv[1][0] == '2';
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
  const int k = (N - 2) / 2, kb = N - 2 - k;  // top rows 0..k-1, middle k, k+1, bottom kb rows
  auto count = [&](double x) {
    if (!two) return sturm_count(a.data(), N, Tb, Tc, x);
    const Side t = sturm_side(a.data(), 1, k, Tb, Tc, x), u = sturm_side(a.data() + N - 1, -1, kb, Tb, Tc, x);
    return sturm_join(t, u, a[k], a[k + 1], Tb, Tc, x);
  };
  for (int j = 0; j < K; ++j) {
    double lo = amin - off, hi = amax + off;
    for (int it = 0; it < 200 && hi - lo > 1e-9 * fmax(1.0, fabs(0.5 * (lo + hi))) + 64 * 2.2e-16 * hn; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (count(mid) > j)
        hi = mid;
      else
        lo = mid;
    }
    const double sigma = 0.5 * (lo + hi);
    for (int i = 0; i < N; ++i) y[i] = sin(1.3 * i + 0.7 * j + 0.1);
    Side st{}, sb{};
    for (int it = 0; it < 2; ++it) {
      if (two) {
        const double tiny = 2.2e-16 * hn;
        double* Fb = F.data() + 4 * (N - 1);
        SideZ zt, zb;
        if (it == 0) {
          zt = ldlt_side_factor_forward(a.data(), y.data(), 1, k, Tb, Tc, sigma, tiny, F.data());
          zb = ldlt_side_factor_forward(a.data() + N - 1, y.data() + N - 1, -1, kb, Tb, Tc, sigma, tiny, Fb);
          st = zt.s;
          sb = zb.s;
        } else {
          zt = ldlt_side_forward(y.data(), 1, k, F.data());
          zb = ldlt_side_forward(y.data() + N - 1, -1, kb, Fb);
        }
        double xk, xk1;
        middle_solve(st, sb, zt, zb, a[k], a[k + 1], y[k], y[k + 1], Tb, Tc, sigma, tiny, xk, xk1);
        ldlt_side_backward(F.data(), 1, k, xk, xk1, y.data());
        ldlt_side_backward(Fb, -1, kb, xk1, xk, y.data() + N - 1);
        y[k] = xk;
        y[k + 1] = xk1;
      } else {
        if (it == 0)
          ldlt_factor_forward(a.data(), y.data(), N, Tb, Tc, sigma, 2.2e-16 * hn, F.data());
        else
          ldlt_forward(y.data(), N, F.data());
        ldlt_backward(F.data(), N, y.data());
      }
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
  if (two) {
    int bad = 0;
    for (int q = 0; q < 400; ++q) {
      const double x = (amin - off) + (amax - amin + 2.0 * off) * (q + 0.5) / 400.0;
      bad += count(x) != sturm_count(a.data(), N, Tb, Tc, x);
    }
    printf("%d\n", bad);
  }
  return 0;
}
