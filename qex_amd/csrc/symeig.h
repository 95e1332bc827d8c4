// symeig.h -- dense symmetric eigensolver for one workgroup (n <= 96, FP64, matrix in LDS).
//
//   1. Householder tridiagonalisation: for k = 0..n-3 the reflector H_k = I - 2 v v^T (|v| = 1)
//      annihilates column k below the subdiagonal; the trailing block is updated as
//      A <- A - v w^T - w v^T with p = A v, beta = v^T p, w = 2 p - 2 beta v.  v is kept in column k
//      of A (below the diagonal), d / e collect the tridiagonal.
//   2. T splits into unreduced blocks where |e_k| <= eps (|d_k| + |d_k+1|).
//   3. Eigenvalues: multisection on Sturm counts (negative pivots of the LDL^T of T - x), G threads
//      per eigenvalue, a fixed number of rounds (interval / (G+1) per round).
//   4. Eigenvectors of T: inverse iteration (3 steps with the LDL^T pivots of T - lam, pivots
//      smaller than eps |T| replaced by +-eps |T|) from a fixed pseudo-random start; eigenvalues of one block
//      closer than 1e-8 |T| form a cluster whose vectors one thread computes in order, removing
//      the earlier members' components after every solve.
//   5. Ascending sort and back-transformation Z <- H_0 ... H_{n-3} Y.
// All reductions run in a fixed order: the result does not depend on timing.
#pragma once
#include <hip/hip_runtime.h>

namespace qexeig {

constexpr double EPS = 2.220446049250313e-16;
constexpr int MAXN = 96;

// one DPP move of a double (both halves)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}

// Sum over the 64 lanes (all active), the same value in every lane: DPP within each row of 16
// lanes (pairs, quads, half-row mirror, row mirror), then the four row sums in a fixed order.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

// Shared state of the solver (static LDS besides the two n x S matrices).
struct Smem {
  double d[MAXN], e[MAXN], e2[MAXN], lam[MAXN], lo[MAXN], rq[MAXN];
  double sg, alpha, iv, v0;  // reflector scalars of the current Householder step
  int bs[MAXN], bt[MAXN], loc[MAXN], perm[MAXN];
  unsigned char cnt[2][1024];
  unsigned char need[MAXN];  // slot's vector wanted (nvec < n: its eigenvalue ranks below nvec)
  double norm;
};

// 1/d: hardware reciprocal refined by one Newton step (relative error ~1e-15)
__device__ __forceinline__ double rcp_nr(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// number of eigenvalues of the block below x (negative pivots of the LDL^T of T - x); e2 = e^2
// The reciprocal of a pivot is clamped to +-rmax = 1/pivmin instead of testing the pivot; the
// sign bit of each pivot is counted with integer arithmetic.
// The diagonal and e^2 are read from LDS 8 entries at a time, so the LDS latency is paid once
// per 8 steps of the pivot recurrence.
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int s, int t, double x, double rmax) {
  double q = d[s] - x;
  unsigned c = (unsigned)((unsigned long long)__double_as_longlong(q) >> 63);
  auto step = [&](double dj, double ej) {
    const double r = fmin(fmax(rcp_nr(q), -rmax), rmax);
    q = fma(-ej, r, dj - x);
    c += (unsigned)((unsigned long long)__double_as_longlong(q) >> 63);
  };
  int j = s + 1;
  for (; j + 8 <= t; j += 8) {
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[j + u];
      ev[u] = e2[j + u - 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) step(dv[u], ev[u]);
  }
  for (; j < t; ++j) step(d[j], e2[j - 1]);
  return (int)c;
}

// deterministic start vector entry in [-1, 1)
__device__ __forceinline__ double start_entry(int slot, int i) {
  unsigned x = 2654435761u * (unsigned)(slot + 1) ^ (40503u * (unsigned)(i + 7));
  x ^= x >> 13;
  x *= 0x5bd1e995u;
  x ^= x >> 15;
  return (double)(x & 0xFFFFFF) / 8388608.0 - 1.0;
}

// Inverse pivots 1/D_i of the LDL^T factorisation of T_block - lam (|D| >= tiny), stored in column
// `col` of Dm (row stride S).  d and e^2 are read 4 rows per LDS wait (one thread runs the whole
// recurrence, so only the arithmetic stays on its dependency chain; 8 rows spill registers).
__device__ __forceinline__ void ldl_pivots(const double* d, const double* e2, int s, int t, double lam, double tiny, double* Dm,
                           int S, int col) {
  double q = d[s] - lam;
  if (fabs(q) < tiny) q = q < 0.0 ? -tiny : tiny;
  double qi = 1.0 / q;
  Dm[s * S + col] = qi;
  auto step = [&](int i, double di, double ei) {
    q = (di - lam) - ei * qi;
    if (fabs(q) < tiny) q = q < 0.0 ? -tiny : tiny;
    qi = 1.0 / q;
    Dm[i * S + col] = qi;
  };
  int i = s + 1;
  for (; i + 4 <= t; i += 4) {
    double dv[4], ev[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      dv[u] = d[i + u];
      ev[u] = e2[i + u - 1];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) step(i + u, dv[u], ev[u]);
  }
  for (; i < t; ++i) step(i, d[i], e2[i - 1]);
}

// (T_block - lam) x = b in place in column `col` of Y, with the inverse pivots from ldl_pivots:
// forward y_i = b_i - e_i-1 (1/D_i-1) y_i-1, then x_i = (y_i - e_i x_i+1) / D_i.  The operands of
// 4 rows are loaded per LDS wait; the chain carries one FMA (and one multiply) per row.
__device__ __forceinline__ void ldl_solve(const double* e, int s, int t, const double* Dm, double* Y, int S, int col) {
  double prev = Y[s * S + col];
  int i = s + 1;
  for (; i + 4 <= t; i += 4) {
    double yv[4], cv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      yv[u] = Y[(i + u) * S + col];
      cv[u] = e[i + u - 1] * Dm[(i + u - 1) * S + col];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      prev = yv[u] - cv[u] * prev;
      Y[(i + u) * S + col] = prev;
    }
  }
  for (; i < t; ++i) {
    const double y = Y[i * S + col] - e[i - 1] * Dm[(i - 1) * S + col] * prev;
    Y[i * S + col] = y;
    prev = y;
  }
  double next = Y[(t - 1) * S + col] * Dm[(t - 1) * S + col];
  Y[(t - 1) * S + col] = next;
  i = t - 2;
  for (; i - 3 >= s; i -= 4) {
    double yv[4], ev[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      yv[u] = Y[(i - u) * S + col];
      ev[u] = e[i - u];
      dv[u] = Dm[(i - u) * S + col];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      next = (yv[u] - ev[u] * next) * dv[u];
      Y[(i - u) * S + col] = next;
    }
  }
  for (; i >= s; --i) {
    next = (Y[i * S + col] - e[i] * next) * Dm[i * S + col];
    Y[i * S + col] = next;
  }
}

// sum over the 8 lanes of an aligned lane group by DPP moves (swap within pairs, within quads,
// then the half-row mirror pairs each quad with the other one); every lane receives the same value
__device__ __forceinline__ double group8_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  return x;
}

// sum_{i<m} a[i*sa] * b[i*sb] with 8 independent partial sums (shorter dependency chains on LDS
// loads), combined in a fixed order
__device__ __forceinline__ double dot8(const double* a, int sa, const double* b, int sb, int m) {
  double c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int i = 0;
  for (; i + 7 < m; i += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = fma(a[(i + u) * sa], b[(i + u) * sb], c[u]);
  }
  for (; i < m; ++i) c[0] = fma(a[i * sa], b[i * sb], c[0]);
  return ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
}

// sum_i a[i * S] b[i * S] over rows s..t-1 of two columns: 4 rows per LDS wait, 4 partial sums
// combined in a fixed order
__device__ __forceinline__ double col_dot4(const double* a, const double* b, int S, int s, int t) {
  double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
  int i = s;
  for (; i + 4 <= t; i += 4) {
    const double a0 = a[i * S], a1 = a[(i + 1) * S], a2 = a[(i + 2) * S], a3 = a[(i + 3) * S];
    const double b0 = b[i * S], b1 = b[(i + 1) * S], b2 = b[(i + 2) * S], b3 = b[(i + 3) * S];
    c0 = fma(a0, b0, c0);
    c1 = fma(a1, b1, c1);
    c2 = fma(a2, b2, c2);
    c3 = fma(a3, b3, c3);
  }
  for (; i < t; ++i) c0 = fma(a[i * S], b[i * S], c0);
  return (c0 + c1) + (c2 + c3);
}

__device__ __forceinline__ double col_norm2(const double* Y, int S, int s, int t, int col) {
  return col_dot4(Y + col, Y + col, S, s, t);
}

// x[i * S] *= f over rows s..t-1, 4 rows per LDS wait
__device__ __forceinline__ void col_scale4(double* x, int S, int s, int t, double f) {
  int i = s;
  for (; i + 4 <= t; i += 4) {
    const double x0 = x[i * S], x1 = x[(i + 1) * S], x2 = x[(i + 2) * S], x3 = x[(i + 3) * S];
    x[i * S] = x0 * f;
    x[(i + 1) * S] = x1 * f;
    x[(i + 2) * S] = x2 * f;
    x[(i + 3) * S] = x3 * f;
  }
  for (; i < t; ++i) x[i * S] *= f;
}

// A: n x n symmetric in LDS (row stride S, destroyed); Z: n x n LDS output, column j = eigenvector
// of the j-th smallest eigenvalue; w: ascending eigenvalues (any memory).  f: global scratch of
// 2 n^2 doubles (3 n^2 are reserved by the callers).  Must be called by every thread of the workgroup (blockDim.x a multiple of 64,
// blockDim.x >= n).
// nvec < n: only the eigenvectors of the nvec smallest eigenvalues (ranked by their multisection
// values, whole clusters) are computed and back-transformed, into columns 0..nvec-1 of Z, exactly as
// the full solve computes them; the other eigenvalues are the multisection values (not refined).
__device__ __forceinline__ void symeig(int n, double* A, double* Z, int S, double* w, double* f, Smem& sh,
                                       int nvec = 1 << 30) {
  if (nvec > n) nvec = n;
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  if (n == 1) {
    if (tid == 0) {
      w[0] = A[0];
      Z[0] = 1.0;
    }
    __syncthreads();
    return;
  }
  // ---- 1. tridiagonalisation ----
  const int grp = tid >> 3, sub = tid & 7;  // 8-lane groups
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    const double* colk = A + (size_t)(k + 1) * S + k;  // x_i = colk[i * S]
    // (a) u_i = sum_{j>=1} A22[i][j] x_j (one 8-lane group per row); |x_1..|^2 by the last group
    if (grp < m) {
      const double* row = A + (size_t)(k + 1 + grp) * S + k + 1;
      double acc = 0.0;
      for (int j = 1 + sub; j < m; j += 8) acc = fma(row[j], colk[j * S], acc);
      acc = group8_sum(acc);
      if (sub == 0) sh.lo[grp] = acc;
    } else if (grp == (nt >> 3) - 1) {
      // the last group also forms the reflector's scalars (square roots and a division done once
      // instead of in every wave); the other threads read them after the barrier
      double acc = 0.0;
      for (int i = 1 + sub; i < m; i += 8) acc = fma(colk[i * S], colk[i * S], acc);
      acc = group8_sum(acc);
      if (sub == 0) {
        const double x0 = colk[0];
        const double nx = sqrt(x0 * x0 + acc);
        const double alpha = x0 >= 0.0 ? -nx : nx;
        const double iv = !(acc > 0.0) ? 0.0 : 1.0 / sqrt(acc + (x0 - alpha) * (x0 - alpha));
        sh.sg = acc;
        sh.alpha = alpha;
        sh.iv = iv;
        sh.v0 = (x0 - alpha) * iv;
      }
    }
    __syncthreads();
    const double sg = sh.sg, x0 = colk[0], alpha = sh.alpha, iv = sh.iv, v0 = sh.v0;
    const bool skip = !(sg > 0.0);
    auto vget = [&](int i) -> double { return i == 0 ? v0 : colk[i * S] * iv; };
    // p_i = A22[i][0] v0 + iv u_i
    auto pget = [&](int i) -> double { return fma(A[(size_t)(k + 1 + i) * S + k + 1], v0, iv * sh.lo[i]); };
    if (tid == 0) {
      sh.d[k] = A[k * S + k];
      sh.e[k] = skip ? x0 : alpha;
    }
    if (!skip) {
      // (b) beta = v . p, reduced inside every wave (identical in all lanes and waves)
      double bt = 0.0;
      for (int j = lane; j < m; j += 64) bt = fma(vget(j), pget(j), bt);
      bt = wave_sum(bt);
      // (c) A22 -= v w^T + w v^T, w = 2 p - 2 beta v: rows by wave, columns by lane
      const int j0 = lane, j1 = lane + 64;
      const double vj0 = j0 < m ? vget(j0) : 0.0, wj0 = j0 < m ? 2.0 * (pget(j0) - bt * vj0) : 0.0;
      const double vj1 = j1 < m ? vget(j1) : 0.0, wj1 = j1 < m ? 2.0 * (pget(j1) - bt * vj1) : 0.0;
      __syncthreads();  // every p_j is read before row 0 of A22 (which holds A22[i][0]) changes
      for (int i = wv; i < m; i += nw) {
        const double vi = vget(i), wi = 2.0 * (pget(i) - bt * vi);
        double* row = A + (size_t)(k + 1 + i) * S + k + 1;
        if (j0 < m) row[j0] -= vi * wj0 + wi * vj0;
        if (j1 < m) row[j1] -= vi * wj1 + wi * vj1;
      }
    }
    __syncthreads();
    // keep v in column k (read above by every thread, so store after the barrier)
    for (int i = tid; i < m; i += nt) A[(k + 1 + i) * S + k] = skip ? 0.0 : vget(i);
    // the next step reads only rows / columns > k + 1 of the trailing block and column k + 1
  }
  __syncthreads();
  if (tid == 0) {
    sh.d[n - 2] = A[(n - 2) * S + n - 2];
    sh.e[n - 2] = A[(n - 1) * S + n - 2];
    sh.d[n - 1] = A[(n - 1) * S + n - 1];
    // ---- 2. split into unreduced blocks; slot -> (block, local index) ----
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      const double r = fabs(sh.d[i]) + (i > 0 ? fabs(sh.e[i - 1]) : 0.0) + (i < n - 1 ? fabs(sh.e[i]) : 0.0);
      nrm = fmax(nrm, r);
    }
    sh.norm = nrm > 0.0 ? nrm : 1e-300;
    int s = 0;
    for (int k = 0; k < n; ++k) {
      const bool cut = k == n - 1 || fabs(sh.e[k]) <= EPS * (fabs(sh.d[k]) + fabs(sh.d[k + 1]));
      if (k < n - 1 && cut) sh.e[k] = 0.0;
      if (k < n - 1) sh.e2[k] = sh.e[k] * sh.e[k];
      if (cut) {
        for (int q = s; q <= k; ++q) {
          sh.bs[q] = s;
          sh.bt[q] = k + 1;
          sh.loc[q] = q - s;
        }
        s = k + 1;
      }
    }
  }
  // reflectors (strict lower part of A) to f, so that A can hold the LDL^T pivots of phase 4
  for (int idx = tid; idx < n * n; idx += nt)
    if (idx / n > idx % n) f[idx] = A[(idx / n) * S + idx % n];
  __syncthreads();
  // ---- 3. eigenvalues by multisection ----
  {
    // G shifts per eigenvalue (>= 10 for n <= 96 and 1024 threads), at most 32: every thread scans
    // its eigenvalue's G counts after each round, so small n gains nothing from more
    const int G = nt / n < 32 ? nt / n : 32;
    const int slot = tid / G, g = tid % G;
    const bool act = slot < n;
    double lo = 0.0, hi = 0.0, rmax = 0.0;
    int s = 0, t = 0, li = 0;
    if (act) {
      s = sh.bs[slot];
      t = sh.bt[slot];
      li = sh.loc[slot];
      double emax = 0.0;
      lo = 1e300;
      hi = -1e300;
      for (int j = s; j < t; ++j) {
        const double r = (j > s ? fabs(sh.e[j - 1]) : 0.0) + (j < t - 1 ? fabs(sh.e[j]) : 0.0);
        lo = fmin(lo, sh.d[j] - r);
        hi = fmax(hi, sh.d[j] + r);
        if (j < t - 1) emax = fmax(emax, sh.e2[j]);
      }
      const double span = fmax(hi - lo, EPS * sh.norm);
      lo -= 1e-12 * span + 1e-300;
      hi += 1e-12 * span + 1e-300;
      rmax = 1.0 / fmax(1e-300, 1e-290 * emax);  // bound on |1 / pivot|
    }
    // the interval shrinks (G+1)-fold per round; 2^-35 of its width (ln = 24.3) is enough because
    // inverse iteration refines the vector and the Rayleigh quotient then gives the eigenvalue
    const int rounds = (int)ceil(24.3 / log((double)(G + 1))) + 1;
    const double ig = 1.0 / (double)(G + 1);
    for (int r = 0; r < rounds; ++r) {
      unsigned char* cb = sh.cnt[r & 1];
      const double wd = (hi - lo) * ig;
      if (act) cb[slot * G + g] = (unsigned char)sturm_count(sh.d, sh.e2, s, t, fma(wd, (double)(g + 1), lo), rmax);
      __syncthreads();
      if (act) {
        double nlo = lo, nhi = hi;
        for (int q = 0; q < G; ++q) {
          const double x = fma(wd, (double)(q + 1), lo);
          if (cb[slot * G + q] <= li)
            nlo = fmax(nlo, x);
          else
            nhi = fmin(nhi, x);
        }
        lo = nlo;
        hi = nhi;
      }
    }
    if (act && g == 0) sh.lam[slot] = 0.5 * (lo + hi);
  }
  __syncthreads();
  // wanted vectors: every slot when nvec = n; else the slots whose multisection value ranks below
  // nvec (ties by slot)
  for (int i = tid; i < n; i += nt) {
    int r = 0;
    if (nvec < n)
      for (int j = 0; j < n; ++j) r += (sh.lam[j] < sh.lam[i]) || (sh.lam[j] == sh.lam[i] && j < i);
    sh.need[i] = r < nvec;
  }
  // ---- 4. eigenvectors of T (column `slot` of Z holds the block's rows) ----
  for (int idx = tid; idx < n * n; idx += nt) Z[(idx / n) * S + idx % n] = 0.0;
  __syncthreads();
  {
    const double tiny = EPS * sh.norm, ctol = 1e-8 * sh.norm;
    const int slot = tid;
    if (slot < n) {
      const int s = sh.bs[slot], t = sh.bt[slot];
      const bool leader = slot == s || sh.lam[slot] - sh.lam[slot - 1] >= ctol;
      int last = slot;
      bool want = false;
      if (leader) {
        while (last + 1 < t && sh.lam[last + 1] - sh.lam[last] < ctol) ++last;
        for (int a = slot; a <= last; ++a) want = want || sh.need[a];
        if (!want)
          for (int a = slot; a <= last; ++a) sh.rq[a] = sh.lam[a];  // not refined: no vector
      }
      if (leader && want) {
        for (int a = slot; a <= last; ++a) {
          if (t - s == 1) {
            Z[s * S + a] = 1.0;
            sh.rq[a] = sh.d[s];
            continue;
          }
          ldl_pivots(sh.d, sh.e2, s, t, sh.lam[a], tiny, A, S, a);
          for (int i = s; i < t; ++i) Z[i * S + a] = start_entry(a, i);
          for (int it = 0; it < 3; ++it) {
            col_scale4(Z + a, S, s, t, 1.0 / sqrt(col_norm2(Z, S, s, t, a)));
            ldl_solve(sh.e, s, t, A, Z, S, a);
            for (int b = slot; b < a; ++b) {
              const double dp = col_dot4(Z + b, Z + a, S, s, t);
              for (int i = s; i < t; ++i) Z[i * S + a] = fma(-dp, Z[i * S + b], Z[i * S + a]);
            }
          }
          col_scale4(Z + a, S, s, t, 1.0 / sqrt(col_norm2(Z, S, s, t, a)));
          // Rayleigh quotient z^T T z = sum_i d_i z_i^2 + 2 sum_i e_i z_i z_i+1 (the eigenvalue to
          // ~eps |T|), 4 partial sums
          double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
          {
            const double* z = Z + a;
            int i = s;
            for (; i + 4 <= t; i += 4) {
              const double z0 = z[i * S], z1 = z[(i + 1) * S], z2 = z[(i + 2) * S], z3 = z[(i + 3) * S];
              const double z4 = i + 4 < t ? z[(i + 4) * S] : 0.0;
              q0 = fma(z0, fma(sh.d[i], z0, 2.0 * sh.e[i] * z1), q0);
              q1 = fma(z1, fma(sh.d[i + 1], z1, 2.0 * sh.e[i + 1] * z2), q1);
              q2 = fma(z2, fma(sh.d[i + 2], z2, 2.0 * sh.e[i + 2] * z3), q2);
              q3 = fma(z3, fma(sh.d[i + 3], z3, i + 4 < t ? 2.0 * sh.e[i + 3] * z4 : 0.0), q3);
            }
            for (; i < t; ++i) {
              const double zi = z[i * S];
              q0 = fma(zi, fma(sh.d[i], zi, i + 1 < t ? 2.0 * sh.e[i] * z[(i + 1) * S] : 0.0), q0);
            }
          }
          sh.rq[a] = (q0 + q1) + (q2 + q3);
        }
      }
    }
  }
  __syncthreads();
  // ---- 5. sort ascending (rank; ties by slot), permute columns through A, back-transform ----
  for (int i = tid; i < n; i += nt) {
    int r = 0;
    for (int j = 0; j < n; ++j) r += (sh.rq[j] < sh.rq[i]) || (sh.rq[j] == sh.rq[i] && j < i);
    sh.perm[r] = i;
  }
  __syncthreads();
  // permuted copy through global scratch (f + n^2); reflectors back into A's strict lower part
  double* g = f + (size_t)n * n;
  for (int idx = tid; idx < n * n; idx += nt) {
    const int i = idx / n, c = idx % n;
    g[idx] = Z[i * S + sh.perm[c]];
    if (i > c) A[i * S + c] = f[idx];
  }
  for (int i = tid; i < n; i += nt) w[i] = sh.rq[sh.perm[i]];
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += nt) Z[(idx / n) * S + idx % n] = g[idx];
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const int m = n - k - 1;
    const double* vk = A + (size_t)(k + 1) * S + k;  // v_i = vk[i * S]
    // sh.lo[c] = v^T Z[k+1:, c], one thread per column
    if (grp < nvec) {
      double acc = 0.0;
      for (int i = sub; i < m; i += 8) acc = fma(vk[i * S], Z[(size_t)(k + 1 + i) * S + grp], acc);
      acc = group8_sum(acc);
      if (sub == 0) sh.lo[grp] = acc;
    }
    __syncthreads();
    const int c0 = lane, c1 = lane + 64;
    const double s0 = c0 < nvec ? 2.0 * sh.lo[c0] : 0.0, s1 = c1 < nvec ? 2.0 * sh.lo[c1] : 0.0;
    for (int i = wv; i < m; i += nw) {
      const double vi = vk[i * S];
      double* row = Z + (size_t)(k + 1 + i) * S;
      if (c0 < nvec) row[c0] -= vi * s0;
      if (c1 < nvec) row[c1] -= vi * s1;
    }
    __syncthreads();
  }
}

}  // namespace qexeig
