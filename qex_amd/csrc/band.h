// band.h -- symmetric pentadiagonal eigen-kernels shared by the gfx950 eigensolver.
//
// A = diag(a) + b (first off-diagonals) + c (second off-diagonals), the 5-point-stencil
// Kohn-Sham Hamiltonian of the 1D grid.  The functions are __host__ __device__ so the same
// code can be exercised on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define QEX_HD __host__ __device__

namespace qexband {


// 1/d to double precision: the hardware reciprocal (relative error <= 5e-8 on gfx950) and one
// Newton step (relative error ~2e-15) on the device.
QEX_HD inline double rcp64(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
#else
  return 1.0 / d;
#endif
}

// 1 if the sign bit of x is set (also for -0.0), else 0 -- without a compare/VCC round trip.
QEX_HD inline int sign_bit(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (int)(((unsigned)__double2hiint(x)) >> 31);
#else
  return signbit(x) ? 1 : 0;
#endif
}

// 1/d clamped to [-RMAX, RMAX]: a pivot smaller than 1/RMAX in magnitude acts as +-1/RMAX
// (an O(1e-60) perturbation of one diagonal entry) and no inf/NaN can enter a recurrence.
constexpr double RMAX = 1e60;
QEX_HD inline double rcp_clamped(double d) { return fmin(fmax(rcp64(d), -RMAX), RMAX); }

// Number of eigenvalues of the symmetric pentadiagonal matrix (diag a, off b, c) below sigma:
// negative pivots of the LDL^T factorisation of A - sigma I.  With u_i = l_i d_i:
//   d_i = (a_i - sigma) - u_{i-1}^2 / d_{i-1} - c^2 / d_{i-2},   u_i = b - c u_{i-1} / d_{i-1}.
// The diagonal is read 8 entries at a time so the LDS latency is paid once per chunk.
QEX_HD inline int sturm_count(const double* __restrict__ a, int N, double b, double c, double sigma) {
  const double nc2 = -c * c, nc = -c;
  double rm1 = 0.0, rm2 = 0.0, um1 = 0.0;
  int cnt = 0;
  auto step = [&](double ai) {
    const double p = fma(nc2, rm2, ai - sigma);
    const double d = fma(-um1 * um1, rm1, p);
    cnt += sign_bit(d);
    const double u = fma(nc * rm1, um1, b);
    rm2 = rm1;
    rm1 = rcp_clamped(d);
    um1 = u;
  };
  int i = 0;
  for (; i + 8 <= N; i += 8) {
    double av[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) av[q] = a[i + q];
#pragma unroll
    for (int q = 0; q < 8; ++q) step(av[q]);
  }
  for (; i < N; ++i) step(a[i]);
  return cnt;
}

// Inverse iteration without pivoting.  A - sigma I = L D L^T with L unit lower triangular
// (L_{i+1,i} = l_i, L_{i+2,i} = m_i); 1/d is clamped to +-1/tiny (a pivot below `tiny` in
// magnitude acts as +-tiny).  F holds 4 doubles per row: 1/d_i, l_i, m_i and the forward-solve value z_i.
//
// ldlt_factor_forward: factorisation fused with the forward solve L z = y.
QEX_HD inline void ldlt_factor_forward(const double* __restrict__ a, const double* __restrict__ y, int N, double b,
                                       double c, double sigma, double tiny, double* __restrict__ F) {
  const double c2 = c * c, rtiny = 1.0 / tiny;
  double rm1 = 0.0, rm2 = 0.0, um1 = 0.0, lm1 = 0.0, mm1 = 0.0, mm2 = 0.0, z1 = 0.0, z2 = 0.0;
  auto step = [&](int i, double ai, double yi) {
    const double d = fma(-um1 * um1, rm1, fma(-c2, rm2, ai - sigma));
    const double r = fmin(fmax(rcp64(d), -rtiny), rtiny);
    const double u = fma(-c * rm1, um1, b);
    const double l = u * r, m = c * r;
    const double z = fma(-mm2, z2, fma(-lm1, z1, yi));
    F[4 * i + 0] = r;
    F[4 * i + 1] = l;
    F[4 * i + 2] = m;
    F[4 * i + 3] = z;
    z2 = z1;
    z1 = z;
    mm2 = mm1;
    mm1 = m;
    lm1 = l;
    rm2 = rm1;
    rm1 = r;
    um1 = u;
  };
  int i = 0;
  for (; i + 8 <= N; i += 8) {
    double av[8], yv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      av[q] = a[i + q];
      yv[q] = y[i + q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) step(i + q, av[q], yv[q]);
  }
  for (; i < N; ++i) step(i, a[i], y[i]);
}

// Forward solve L z = y with an existing factorisation (z stored in F[4i+3]).
QEX_HD inline void ldlt_forward(const double* __restrict__ y, int N, double* __restrict__ F) {
  double z1 = 0.0, z2 = 0.0, lm1 = 0.0, mm1 = 0.0, mm2 = 0.0;
  auto step = [&](int i, double yi, double l, double m) {
    const double z = fma(-mm2, z2, fma(-lm1, z1, yi));
    F[4 * i + 3] = z;
    z2 = z1;
    z1 = z;
    mm2 = mm1;
    mm1 = m;
    lm1 = l;
  };
  int i = 0;
  for (; i + 8 <= N; i += 8) {
    double yv[8], lv[8], mv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      yv[q] = y[i + q];
      lv[q] = F[4 * (i + q) + 1];
      mv[q] = F[4 * (i + q) + 2];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) step(i + q, yv[q], lv[q], mv[q]);
  }
  for (; i < N; ++i) step(i, y[i], F[4 * i + 1], F[4 * i + 2]);
}

// Back substitution L^T x = D^{-1} z:  x_i = z_i / d_i - l_i x_{i+1} - m_i x_{i+2}.
QEX_HD inline void ldlt_backward(const double* __restrict__ F, int N, double* __restrict__ x) {
  double x1 = 0.0, x2 = 0.0;
  auto step = [&](int i, double r, double l, double m, double z) {
    const double xi = fma(-m, x2, fma(-l, x1, z * r));
    x[i] = xi;
    x2 = x1;
    x1 = xi;
  };
  int i = N - 1;
  for (; i >= 7; i -= 8) {
    double rv[8], lv[8], mv[8], zv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = i - q;
      rv[q] = F[4 * k + 0];
      lv[q] = F[4 * k + 1];
      mv[q] = F[4 * k + 2];
      zv[q] = F[4 * k + 3];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) step(i - q, rv[q], lv[q], mv[q], zv[q]);
  }
  for (; i >= 0; --i) step(i, F[4 * i], F[4 * i + 1], F[4 * i + 2], F[4 * i + 3]);
}

// ---- two-sided elimination ------------------------------------------------------------------
// The rows of A - sigma are eliminated from the top (rows 0 .. k-1) and from the bottom (rows
// N-1 .. k+2) at the same time.  Neither side touches the other's rows: a top row couples to rows
// <= k+1 and a bottom row to rows >= k, so both only update the 2x2 block of the middle pair
// (k, k+1), and A - sigma = P diag(D_top, M, D_bottom) P^T.  Each side runs the one-sided
// recurrence over its own rows (half the chain), read with stride st = +1 from row 0 or st = -1
// from row N-1; the inertia of A - sigma is that of the pivots plus that of M.
struct Side {
  double r1, r2, u1;  // 1/d of the side's last two pivots, Schur off-diagonal u of its last row
  int neg;            // negative pivots
};

// Sturm count of one side (rows >= 0 rows, a = the side's first row).
QEX_HD inline Side sturm_side(const double* __restrict__ a, int st, int rows, double b, double c, double sigma) {
  const double nc2 = -c * c, nc = -c;
  double rm1 = 0.0, rm2 = 0.0, um1 = 0.0;
  int cnt = 0;
  auto step = [&](double ai) {
    const double p = fma(nc2, rm2, ai - sigma);
    const double d = fma(-um1 * um1, rm1, p);
    cnt += sign_bit(d);
    const double u = fma(nc * rm1, um1, b);
    rm2 = rm1;
    rm1 = rcp_clamped(d);
    um1 = u;
  };
  int i = 0;
  for (; i + 8 <= rows; i += 8) {
    double av[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) av[q] = a[st * (i + q)];
#pragma unroll
    for (int q = 0; q < 8; ++q) step(av[q]);
  }
  for (; i < rows; ++i) step(a[st * i]);
  return Side{rm1, rm2, um1, cnt};
}

// The middle block M of the pair (k, k+1) from the two sides; ak, ak1 = diagonal of rows k, k+1.
QEX_HD inline void middle_block(const Side& t, const Side& u, double ak, double ak1, double b, double c, double sigma,
                                double& m00, double& m01, double& m11) {
  const double c2 = c * c;
  m00 = (ak - sigma) - t.u1 * t.u1 * t.r1 - c2 * t.r2 - c2 * u.r1;
  m11 = (ak1 - sigma) - c2 * t.r1 - u.u1 * u.u1 * u.r1 - c2 * u.r2;
  m01 = b - c * t.u1 * t.r1 - c * u.u1 * u.r1;
}

// Number of eigenvalues below sigma from the two sides and the inertia of M.
QEX_HD inline int sturm_join(const Side& t, const Side& u, double ak, double ak1, double b, double c, double sigma) {
  double m00, m01, m11;
  middle_block(t, u, ak, ak1, b, c, sigma, m00, m01, m11);
  const double det = m00 * m11 - m01 * m01;
  const int neg = det < 0.0 ? 1 : (det > 0.0 ? (m00 < 0.0 ? 2 : 0) : (m00 + m11 < 0.0 ? 1 : 0));
  return t.neg + u.neg + neg;
}

// Forward state of one side of a solve: the Side of the factorisation plus the multipliers and
// forward-solve values of its last two rows, which enter the middle right-hand side.
struct SideZ {
  Side s;
  double l1, m1, m2, z1, z2;  // l, m of the last row, m of the row before; z of the last two rows
};

// Factorisation of one side fused with its forward solve L z = y (a, y = the side's first row).
// F rows are written at F + 4 st i (F = the side's first row), 4 doubles per row as in
// ldlt_factor_forward (1/d, l, m, z).
QEX_HD inline SideZ ldlt_side_factor_forward(const double* __restrict__ a, const double* __restrict__ y, int st,
                                             int rows, double b, double c, double sigma, double tiny,
                                             double* __restrict__ F) {
  const double c2 = c * c, rtiny = 1.0 / tiny;
  double rm1 = 0.0, rm2 = 0.0, um1 = 0.0, lm1 = 0.0, mm1 = 0.0, mm2 = 0.0, z1 = 0.0, z2 = 0.0;
  auto step = [&](int i, double ai, double yi) {
    const double d = fma(-um1 * um1, rm1, fma(-c2, rm2, ai - sigma));
    const double r = fmin(fmax(rcp64(d), -rtiny), rtiny);
    const double u = fma(-c * rm1, um1, b);
    const double l = u * r, m = c * r;
    const double z = fma(-mm2, z2, fma(-lm1, z1, yi));
    double* f = F + 4 * st * i;
    f[0] = r;
    f[1] = l;
    f[2] = m;
    f[3] = z;
    z2 = z1;
    z1 = z;
    mm2 = mm1;
    mm1 = m;
    lm1 = l;
    rm2 = rm1;
    rm1 = r;
    um1 = u;
  };
  int i = 0;
  for (; i + 8 <= rows; i += 8) {
    double av[8], yv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      av[q] = a[st * (i + q)];
      yv[q] = y[st * (i + q)];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) step(i + q, av[q], yv[q]);
  }
  for (; i < rows; ++i) step(i, a[st * i], y[st * i]);
  return SideZ{Side{rm1, rm2, um1, 0}, lm1, mm1, mm2, z1, z2};
}

// Forward solve of one side with its existing factorisation (z stored in F); the returned Side
// part is that of the factorisation's last rows.
QEX_HD inline SideZ ldlt_side_forward(const double* __restrict__ y, int st, int rows, double* __restrict__ F) {
  double z1 = 0.0, z2 = 0.0, lm1 = 0.0, mm1 = 0.0, mm2 = 0.0;
  auto step = [&](int i, double yi, double l, double m) {
    const double z = fma(-mm2, z2, fma(-lm1, z1, yi));
    F[4 * st * i + 3] = z;
    z2 = z1;
    z1 = z;
    mm2 = mm1;
    mm1 = m;
    lm1 = l;
  };
  int i = 0;
  for (; i + 8 <= rows; i += 8) {
    double yv[8], lv[8], mv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      yv[q] = y[st * (i + q)];
      lv[q] = F[4 * st * (i + q) + 1];
      mv[q] = F[4 * st * (i + q) + 2];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) step(i + q, yv[q], lv[q], mv[q]);
  }
  for (; i < rows; ++i) step(i, y[st * i], F[4 * st * i + 1], F[4 * st * i + 2]);
  return SideZ{Side{0.0, 0.0, 0.0, 0}, lm1, mm1, mm2, z1, z2};
}

// Solve of the middle pair: M [x_k, x_k+1] = [y_k, y_k+1] minus both sides' forward terms, M by
// a 2x2 LDL^T with the pivots clamped to +-tiny like the sides' pivots.
QEX_HD inline void middle_solve(const Side& t, const Side& u, const SideZ& zt, const SideZ& zu, double ak, double ak1,
                                double yk, double yk1, double b, double c, double sigma, double tiny, double& xk,
                                double& xk1) {
  double m00, m01, m11;
  middle_block(t, u, ak, ak1, b, c, sigma, m00, m01, m11);
  const double r0 = yk - (zt.l1 * zt.z1 + zt.m2 * zt.z2) - zu.m1 * zu.z1;
  const double r1 = yk1 - zt.m1 * zt.z1 - (zu.l1 * zu.z1 + zu.m2 * zu.z2);
  double p0 = m00;
  if (fabs(p0) < tiny) p0 = p0 < 0.0 ? -tiny : tiny;
  const double l = m01 / p0;
  double p1 = m11 - l * m01;
  if (fabs(p1) < tiny) p1 = p1 < 0.0 ? -tiny : tiny;
  xk1 = (r1 - l * r0) / p1;
  xk = r0 / p0 - l * xk1;
}

// Back substitution of one side toward its outer end, x_i = z_i / d_i - l_i x_(i-1 side) -
// m_i x_(i-2 side); x1, x2 = the middle values nearest and next (top: x_k, x_k+1; bottom:
// x_k+1, x_k).  x = the side's first row.
QEX_HD inline void ldlt_side_backward(const double* __restrict__ F, int st, int rows, double x1, double x2,
                                      double* __restrict__ x) {
  auto step = [&](int i, double r, double l, double m, double z) {
    const double xi = fma(-m, x2, fma(-l, x1, z * r));
    x[st * i] = xi;
    x2 = x1;
    x1 = xi;
  };
  int i = rows - 1;
  for (; i >= 7; i -= 8) {
    double rv[8], lv[8], mv[8], zv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double* f = F + 4 * st * (i - q);
      rv[q] = f[0];
      lv[q] = f[1];
      mv[q] = f[2];
      zv[q] = f[3];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) step(i - q, rv[q], lv[q], mv[q], zv[q]);
  }
  for (; i >= 0; --i) {
    const double* f = F + 4 * st * i;
    step(i, f[0], f[1], f[2], f[3]);
  }
}

}  // namespace qexband
