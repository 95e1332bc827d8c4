// qex3d.hip -- 3D restricted Kohn-Sham / Hartree-Fock cycle for gfx950 (MI355X).
//
// Per molecule (Cartesian s/p Gaussian basis, n_ao <= QEX3D_MAX_NAO, atom-centred grid):
//   k_int1e / k_eri       overlap, kinetic, nuclear attraction and (ij|kl) by McMurchie-Davidson
//                         (ERIs stored as the [npair][npair] matrix of i>=j, k>=l pairs)
//   k_grid                Becke-partitioned atom-centred grid from radial x angular base tables
//   k_eval_ao             phi[g][mu]
//   k_rho                 rho_g = phi_g^T D phi_g, D in LDS (broadcast), phi_g in registers
//   k_vxc                 LDA (Slater + VWN5) and per-chunk partials of phi^T diag(w v) phi
//   k_coulomb             J (and K for Hartree-Fock) from the ERI pair matrix (HBM-bound GEMV)
//   k_scf_*               Fock build + CDIIS, generalized eigh (orthogonaliser X = L^-T of the
//                         overlap Cholesky, parallel cyclic Jacobi warm-started from the previous
//                         eigenvectors), density matrix, energy, convergence -- one workgroup each
// The SCF loop is enqueued without host synchronisation; a device flag freezes the state once
// |dE| < tol and |g| / sqrt(n_vir n_occ) < sqrt(tol) (pyscf's criterion).
#include <hip/hip_runtime.h>
#include <type_traits>
#include <vector>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "qex.h"
#include "prof.h"
#include "symeig.h"

namespace {

constexpr int NT = 256;
constexpr int NW = NT / 64;
thread_local char g_err3[512] = "";

// the SCF methods whose XC is a local density approximation evaluated in the density epilogue
__host__ __device__ inline bool is_lda(int method) { return method == QEX3D_RKS_LDA || method == QEX3D_RKS_LDA_X; }
// methods whose XC energy is a per-point sum, rho w e_xc, accumulated with the Vxc partials
__host__ __device__ inline bool has_point_exc(int method) { return is_lda(method) || method == QEX3D_RKS_NN_LOCAL; }

int fail3(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err3, sizeof(g_err3), fmt, ap);
  va_end(ap);
  return code;
}

int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail3(QEX_E_LAUNCH, "%s: %s", what, hipGetErrorString(e));
  return QEX_OK;
}

__device__ __forceinline__ double wsum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// ordered workgroup sum (every thread receives the total)
__device__ double bsum(double x, double* red) {
  x = wsum(x);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = x;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
  __syncthreads();
  return s;
}

__device__ __forceinline__ int pair_index(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// ---------------------------------------------------------------------------------------------
// Boys function F_m(T) = int_0^1 t^2m exp(-T t^2) dt for m = 0..4 (s/p quartets need L <= 4).
// Below BOYS_TMAX: 7-term Taylor expansion of F_4 about the nearest tabulated point (spacing
// 0.05, |dT| <= 0.025, truncation error < 2e-15 relative), then downward recursion
// F_m = (2T F_m+1 + exp(-T)) / (2m+1).  Above: F_0 = sqrt(pi/T)/2 and upward recursion.  The
// table holds F_0..F_11 at T_i = 0.05 i, made on the device once (series, then downward recursion).
// ---------------------------------------------------------------------------------------------
constexpr int BOYS_NM = 12;
constexpr double BOYS_H = 0.05;
constexpr int BOYS_NT = 801;  // T_i up to 40
constexpr double BOYS_TMAX = 39.9;
__device__ double g_boys_tab[BOYS_NT * BOYS_NM];

__global__ void k_boys_init() {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= BOYS_NT) return;
  const double T = BOYS_H * i;
  const int mm = BOYS_NM - 1;
  double term = 1.0 / (2 * mm + 1), sum = term;
  for (int k = 1; k < 400; ++k) {
    term *= 2.0 * T / (2 * mm + 2 * k + 1);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  const double et = exp(-T);
  double f = et * sum;
  g_boys_tab[i * BOYS_NM + mm] = f;
  for (int m = mm - 1; m >= 0; --m) {
    f = (2.0 * T * f + et) / (2 * m + 1);
    g_boys_tab[i * BOYS_NM + m] = f;
  }
}

struct Boys5 {
  double f0, f1, f2, f3, f4;
  __device__ __forceinline__ double operator[](int N) const {
    return N == 0 ? f0 : N == 1 ? f1 : N == 2 ? f2 : N == 3 ? f3 : f4;
  }
};

__device__ __forceinline__ Boys5 boys5(double T) {
  Boys5 F;
  const double et = exp(-T);
  if (T < BOYS_TMAX) {
    const int i = (int)(T * (1.0 / BOYS_H) + 0.5);
    const double dT = T - BOYS_H * i;
    const double* r = g_boys_tab + i * BOYS_NM;
    // F_4(T) = sum_k F_{4+k}(T_i) (-dT)^k / k!
    double f = r[10];
    f = fma(f, -dT / 6.0, r[9]);
    f = fma(f, -dT / 5.0, r[8]);
    f = fma(f, -dT / 4.0, r[7]);
    f = fma(f, -dT / 3.0, r[6]);
    f = fma(f, -dT / 2.0, r[5]);
    f = fma(f, -dT, r[4]);
    F.f4 = f;
    F.f3 = (2.0 * T * F.f4 + et) / 7.0;
    F.f2 = (2.0 * T * F.f3 + et) / 5.0;
    F.f1 = (2.0 * T * F.f2 + et) / 3.0;
    F.f0 = 2.0 * T * F.f1 + et;
  } else {
    const double i2t = 0.5 / T;
    F.f0 = 0.5 * sqrt(M_PI / T);
    F.f1 = (F.f0 - et) * i2t;
    F.f2 = (3.0 * F.f1 - et) * i2t;
    F.f3 = (5.0 * F.f2 - et) * i2t;
    F.f4 = (7.0 * F.f3 - et) * i2t;
  }
  return F;
}

// Hermite Coulomb integral R_tuv (t+u+v <= 4) in closed form:
//   R_tuv = sum_{i,j,k} c(t,i) c(u,j) c(v,k) (2aX)^(t-2i) (2aY)^(u-2j) (2aZ)^(v-2k) a^(i+j+k) (-1)^N F_N,
//   N = t+u+v-i-j-k, c(t,i) = t! / (i! (t-2i)!)  (the t-th X-derivative of F_0(a (X^2+Y^2+Z^2)))
__device__ __forceinline__ double hcoef(int t, int i) {
  // t <= 4: c(2,1) = 2, c(3,1) = 6, c(4,1) = 12, c(4,2) = 12, else 1
  return i == 0 ? 1.0 : (t == 2 ? 2.0 : t == 3 ? 6.0 : 12.0);
}
__device__ __forceinline__ double ipow(double x, int n) {
  double r = 1.0;
  for (int k = 0; k < n; ++k) r *= x;
  return r;
}
template <class FT>  // F[N] = F_N(T): a Boys5 or a pointer
__device__ __forceinline__ double rtuv(int t, int u, int v, double a, double X, double Y, double Z, const FT& F) {
  const double ax = 2.0 * a * X, ay = 2.0 * a * Y, az = 2.0 * a * Z;
  double r = 0.0;
  for (int i = 0; 2 * i <= t; ++i) {
    const double ci = hcoef(t, i) * ipow(ax, t - 2 * i);
    for (int j = 0; 2 * j <= u; ++j) {
      const double cj = ci * hcoef(u, j) * ipow(ay, u - 2 * j);
      for (int k = 0; 2 * k <= v; ++k) {
        const int N = t + u + v - i - j - k;
        const double c = cj * hcoef(v, k) * ipow(az, v - 2 * k) * ipow(a, i + j + k);
        r += (N & 1) ? -c * F[N] : c * F[N];
      }
    }
  }
  return r;
}

// Hermite expansion coefficients E_t for one Cartesian direction, l_a, l_b <= 1:
// K = exp(-mu X_AB^2), X_PA, X_PB, 1/(2p).
__device__ __forceinline__ void hermite_E1(int la, int lb, double K, double xpa, double xpb, double i2p, double* E) {
  if (la == 0 && lb == 0) {
    E[0] = K;
    E[1] = 0.0;
    E[2] = 0.0;
  } else if (la == 1 && lb == 0) {
    E[0] = xpa * K;
    E[1] = i2p * K;
    E[2] = 0.0;
  } else if (la == 0 && lb == 1) {
    E[0] = xpb * K;
    E[1] = i2p * K;
    E[2] = 0.0;
  } else {
    E[0] = (xpa * xpb + i2p) * K;
    E[1] = (xpa + xpb) * i2p * K;
    E[2] = i2p * i2p * K;
  }
}

// ---------------------------------------------------------------------------------------------
// basis description on the device
// ---------------------------------------------------------------------------------------------
struct Basis {
  int natm, nshell, nao;
  const double* xyz;     // [natm][3]
  const double* charge;  // [natm]
  const int* sh_atom;    // [nshell]
  const int* sh_l;       // [nshell]
  const int* sh_nprim;   // [nshell]
  const int* sh_off;     // [nshell] offset into exps / coefs
  const double* exps;
  const double* coefs;   // contraction coefficients including the primitive normalisation
  const int* ao_shell;   // [nao]
  const int* ao_lmn;     // [nao] packed lx | ly << 2 | lz << 4
};

__device__ __forceinline__ void ao_info(const Basis& b, int mu, int* sh, int* lx, int* ly, int* lz) {
  *sh = b.ao_shell[mu];
  const int c = b.ao_lmn[mu];
  *lx = c & 3;
  *ly = (c >> 2) & 3;
  *lz = (c >> 4) & 3;
}

// overlap and kinetic contributions of one primitive pair (exponents a, bb) to S_ij and T_ij
__device__ __forceinline__ void st_prim(const int* la, const int* lb, const double* A, const double* B, double a,
                                        double bb, double cc, double* s, double* t) {
  const double p = a + bb, mu = a * bb / p, i2p = 0.5 / p;
  double s1[3], dd[3];
  for (int d = 0; d < 3; ++d) {
    const double Pd = (a * A[d] + bb * B[d]) / p;
    const double xab = A[d] - B[d];
    const double K = exp(-mu * xab * xab);
    const double xpa = Pd - A[d], xpb = Pd - B[d];
    const double sq = sqrt(M_PI / p);
    // 1D overlaps s(la, j) for j = 0..3 from the Hermite E_0 of (la, j) (la <= 1)
    double Ei[4];
    {
      double E0[4][5] = {{0}}, E1[4][5] = {{0}};
      E0[0][0] = K;
      for (int jj = 0; jj < 3; ++jj)
        for (int tt = 0; tt <= jj + 1; ++tt)
          E0[jj + 1][tt] = (tt > 0 ? i2p * E0[jj][tt - 1] : 0.0) + xpb * E0[jj][tt] + (tt + 1) * E0[jj][tt + 1];
      for (int jj = 0; jj < 4; ++jj)
        for (int tt = 0; tt <= jj + 1; ++tt)
          E1[jj][tt] = (tt > 0 ? i2p * E0[jj][tt - 1] : 0.0) + xpa * E0[jj][tt] + (tt + 1) * (tt + 1 <= 4 ? E0[jj][tt + 1] : 0.0);
      for (int jj = 0; jj < 4; ++jj) Ei[jj] = (la[d] == 0 ? E0[jj][0] : E1[jj][0]) * sq;
    }
    const int jb = lb[d];
    s1[d] = Ei[jb];
    const double sm2 = jb >= 2 ? Ei[jb - 2] : 0.0;
    dd[d] = jb * (jb - 1) * sm2 - 2.0 * bb * (2 * jb + 1) * Ei[jb] + 4.0 * bb * bb * Ei[jb + 2];
  }
  *s += cc * s1[0] * s1[1] * s1[2];
  *t += cc * -0.5 * (dd[0] * s1[1] * s1[2] + s1[0] * dd[1] * s1[2] + s1[0] * s1[1] * dd[2]);
}

// nuclear-attraction contribution of one primitive pair and one nucleus C (charge Zc) to V_ij
__device__ __forceinline__ double v_prim_atom(const int* la, const int* lb, const double* A, const double* B, double a,
                                              double bb, double cc, const double* C, double Zc) {
  const double p = a + bb, mu = a * bb / p, i2p = 0.5 / p;
  double P[3], Ex[3], Ey[3], Ez[3];
  for (int d = 0; d < 3; ++d) P[d] = (a * A[d] + bb * B[d]) / p;
  hermite_E1(la[0], lb[0], exp(-mu * (A[0] - B[0]) * (A[0] - B[0])), P[0] - A[0], P[0] - B[0], i2p, Ex);
  hermite_E1(la[1], lb[1], exp(-mu * (A[1] - B[1]) * (A[1] - B[1])), P[1] - A[1], P[1] - B[1], i2p, Ey);
  hermite_E1(la[2], lb[2], exp(-mu * (A[2] - B[2]) * (A[2] - B[2])), P[2] - A[2], P[2] - B[2], i2p, Ez);
  const int tx = la[0] + lb[0], ty = la[1] + lb[1], tz = la[2] + lb[2];
  const double X = P[0] - C[0], Y = P[1] - C[1], Z = P[2] - C[2];
  const Boys5 F = boys5(p * (X * X + Y * Y + Z * Z));
  double acc = 0.0;
  for (int t0 = 0; t0 <= tx; ++t0)
    for (int u0 = 0; u0 <= ty; ++u0)
      for (int v0 = 0; v0 <= tz; ++v0) {
        const double e = Ex[t0] * Ey[u0] * Ez[v0];
        if (e != 0.0) acc += e * rtuv(t0, u0, v0, p, X, Y, Z, F);
      }
  return -cc * Zc * 2.0 * M_PI / p * acc;
}

// one-electron integrals, one wave per AO pair (i >= j): lanes take the primitive pairs of S and T
// and the (primitive pair, nucleus) terms of V; wave sums in a fixed order
__global__ __launch_bounds__(256) void k_int1e(Basis b, double* S, double* Tm, double* V) {
  const int idx = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int n = b.nao;
  if (idx >= n * (n + 1) / 2) return;  // whole waves
  int i = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > idx) --i;
  while ((i + 1) * (i + 2) / 2 <= idx) ++i;
  const int j = idx - i * (i + 1) / 2;
  int si, sj, la[3], lb[3];
  ao_info(b, i, &si, &la[0], &la[1], &la[2]);
  ao_info(b, j, &sj, &lb[0], &lb[1], &lb[2]);
  const double* A = b.xyz + 3 * b.sh_atom[si];
  const double* B = b.xyz + 3 * b.sh_atom[sj];
  const int npa = b.sh_nprim[si], npb = b.sh_nprim[sj], npp = npa * npb;
  const int oa = b.sh_off[si], ob = b.sh_off[sj];
  double s = 0.0, t = 0.0, v = 0.0;
  for (int r = lane; r < npp; r += 64) {
    const int pa = r / npb, pb = r % npb;
    st_prim(la, lb, A, B, b.exps[oa + pa], b.exps[ob + pb], b.coefs[oa + pa] * b.coefs[ob + pb], &s, &t);
  }
  for (int q = lane; q < npp * b.natm; q += 64) {
    const int r = q / b.natm, c = q % b.natm, pa = r / npb, pb = r % npb;
    v += v_prim_atom(la, lb, A, B, b.exps[oa + pa], b.exps[ob + pb], b.coefs[oa + pa] * b.coefs[ob + pb],
                     b.xyz + 3 * c, b.charge[c]);
  }
  s = wsum(s);
  t = wsum(t);
  v = wsum(v);
  if (lane == 0) {
    S[i * n + j] = S[j * n + i] = s;
    Tm[i * n + j] = Tm[j * n + i] = t;
    V[i * n + j] = V[j * n + i] = v;
  }
}

// ---------------------------------------------------------------------------------------------
// Two-electron integrals (ij|kl) as the pair matrix eri[P][Q], P = pair(i >= j), by
// McMurchie-Davidson over s/p shells.  The main pass works on shell quartets: a group of 8 lanes
// per (bra shell pair, ket shell pair), each lane taking every 8th primitive quartet.  A quartet
// computes F_0..F_L and the Hermite integrals R_tuv (t+u+v <= L) once; then, for every AO component
// pair (ij) of the bra shell pair, the bra contraction W = sum_L1 E^ij_L1 R_{L1+L2} over the
// nonzero Hermite coefficients of (ij) only, and for every component pair (kl) of the ket a dot
// with the ket's nonzero coefficients; the group sums its lanes by a fixed xor tree.  The (bra,
// ket) angular classes (0: ss, 1: sp or ps, 2: pp) and the component indices are compile-time
// constants.  Passes: primitive pair records; shell-pair lists; (ij|ij) for every AO pair; the
// Schwarz factor of every shell pair; shell pairs sorted by class (bra pairs further by their
// primitive-pair count); per ket shell pair the list of bra shell pairs holding a pair P >= Q's
// first pair (symmetry: only Q <= P is written, eri[P][Q] and eri[Q][P] both) whose Schwarz
// product reaches ERI_SCHWARZ_TOL; the output cleared; the main pass, one launch per bra class (pp
// bras in two: ket classes ss/sp, and pp|pp in three passes over the ket's first component).
// ---------------------------------------------------------------------------------------------
constexpr int SREC = 16;  // doubles per (shell pair, primitive pair) record
constexpr int EREC = 13;  // the leading doubles of a ket record, staged in LDS by the main pass
constexpr double TWO_PI_2_5 = 34.98683665524972;  // 2 pi^(5/2)

// Hermite index k <-> (t, u, v), ordered by N = t+u+v, then by u+v, then by v
__host__ __device__ constexpr int h_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int h_index(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
// (t, u, v) of Hermite index k as constant expressions
__host__ __device__ constexpr int h_n(int k) {
  int N = 0;
  while (h_count(N) <= k) ++N;
  return N;
}
__host__ __device__ constexpr int h_rem(int k) { return k - h_count(h_n(k) - 1); }
__host__ __device__ constexpr int h_s(int r) {
  int s = 0;
  while ((s + 1) * (s + 2) / 2 <= r) ++s;
  return s;
}
__host__ __device__ constexpr int h_tv(int k) { return h_rem(k) - h_s(h_rem(k)) * (h_s(h_rem(k)) + 1) / 2; }
__host__ __device__ constexpr int h_tu(int k) { return h_s(h_rem(k)) - h_tv(k); }
__host__ __device__ constexpr int h_tt(int k) { return h_n(k) - h_s(h_rem(k)); }
constexpr bool h_roundtrip() {
  for (int k = 0; k < 35; ++k)
    if (h_index(h_tt(k), h_tu(k), h_tv(k)) != k) return false;
  return true;
}
static_assert(h_roundtrip(), "Hermite index decomposition");
__device__ constexpr int HT[35] = {0, 1, 0, 0, 2, 1, 1, 0, 0, 0, 3, 2, 2, 1, 1, 1, 0, 0, 0, 0,
                                   4, 3, 3, 2, 2, 2, 1, 1, 1, 1, 0, 0, 0, 0, 0};
__device__ constexpr int HU[35] = {0, 0, 1, 0, 0, 1, 0, 2, 1, 0, 0, 1, 0, 2, 1, 0, 3, 2, 1, 0,
                                   0, 1, 0, 2, 1, 0, 3, 2, 1, 0, 4, 3, 2, 1, 0};
__device__ constexpr int HV[35] = {0, 0, 0, 1, 0, 0, 1, 0, 1, 2, 0, 0, 1, 0, 1, 2, 0, 1, 2, 3,
                                  0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4};

// F_0..F_L(T): Taylor expansion of F_L about the nearest tabulated point, downward recursion below
// (F_0 alone needs no exponential); above the table, F_0 = sqrt(pi/T)/2 and upward recursion.
template <int L>
__device__ __forceinline__ void boys_upto(double T, double* F) {
  if (T < BOYS_TMAX) {
    const int i = (int)(T * (1.0 / BOYS_H) + 0.5);
    const double dT = T - BOYS_H * i;
    const double* r = g_boys_tab + i * BOYS_NM + L;
    double f = r[6];
    f = fma(f, dT * (-1.0 / 6.0), r[5]);
    f = fma(f, dT * (-1.0 / 5.0), r[4]);
    f = fma(f, dT * (-1.0 / 4.0), r[3]);
    f = fma(f, dT * (-1.0 / 3.0), r[2]);
    f = fma(f, dT * -0.5, r[1]);
    f = fma(f, -dT, r[0]);
    F[L] = f;
    if (L > 0) {
      const double et = exp(-T), t2 = 2.0 * T;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = fma(t2, F[m + 1], et) * (1.0 / (2 * m + 1));
    }
  } else {
    const double rt = rsqrt(T);
    F[0] = 0.88622692545275801 * rt;  // sqrt(pi) / 2 / sqrt(T)
    if (L > 0) {
      const double et = exp(-T), i2t = 0.5 * rt * rt;
#pragma unroll
      for (int m = 1; m <= L; ++m) F[m] = ((2 * m - 1) * F[m - 1] - et) * i2t;
    }
  }
}

// R_tuv for all t+u+v <= L (closed form of the Hermite Coulomb integrals, constant indices)
template <int L>
__device__ __forceinline__ void rtable(double a, double X, double Y, double Z, const double* F, double* R) {
#pragma unroll
  for (int k = 0; k < h_count(L); ++k) R[k] = rtuv(HT[k], HU[k], HV[k], a, X, Y, Z, F);
}

// E^{ab}_{tuv} as a dense vector over t+u+v <= la+lb (per-direction E from hermite_E1)
template <int LAB>
__device__ __forceinline__ void e_dense(const int* la, const int* lb, const double* PA, const double* PB, double i2p,
                                        double* E) {
  double Ex[3], Ey[3], Ez[3];
  hermite_E1(la[0], lb[0], 1.0, PA[0], PB[0], i2p, Ex);
  hermite_E1(la[1], lb[1], 1.0, PA[1], PB[1], i2p, Ey);
  hermite_E1(la[2], lb[2], 1.0, PA[2], PB[2], i2p, Ez);
#pragma unroll
  for (int k = 0; k < h_count(LAB); ++k) E[k] = (HT[k] <= 2 ? Ex[HT[k]] : 0.0) * (HU[k] <= 2 ? Ey[HU[k]] : 0.0) *
                                              (HV[k] <= 2 ? Ez[HV[k]] : 0.0);
}

struct EriArgs {
  Basis bs;
  int nao, npair, nsp, maxp;
  int screen;        // Schwarz screening (bases of >= ERI_SCREEN_MIN_PAIRS AO pairs)
  int k_off, p_off;  // first ket shell pair / screened AO pair of this molecule in the batch's grids
  int m_off;         // first ket shell pair of this molecule in its main-pass launch
  long rec_off;      // first primitive-pair record of this molecule in the batch's record grid
  int* ao_first;     // [nshell] first AO of each shell
  double* rec;       // [nsp][maxp^2][SREC], the shell pair's na * nb primitive pairs first
  int* bra;          // [nsp] shell pairs sorted by class, then by primitive-pair count
  int* ket;          // [nsp] shell pairs sorted by class
  int* cnt;          // [0..2] bra class sizes, [3..5] ket class sizes
  double* sq;        // [nsp] Schwarz factor max sqrt|(ij|ij)| of each shell pair
  double* diag;      // [npair] (ij|ij) (diagonal pass)
  int* live;         // [nsp][nsp] per sorted ket: its unscreened bra shell pairs holding a P >= qmin, in bra order
  int* live_cnt;     // [nsp][3] their counts per bra class
  double* eri;
};
static_assert(sizeof(EriArgs) % 8 == 0, "EriArgs is copied in 8-byte words");

// (ij|kl) with sqrt|(ij|ij)| sqrt|(kl|kl)| (shell-pair maxima) below this are written as zero: every
// skipped integral is below it in magnitude (pyscf screens its direct SCF at 1e-13)
constexpr double ERI_SCHWARZ_TOL = 1e-15;
// primitive quartets whose c_ab c_cd 2 pi^(5/2) / (p q sqrt(p + q)) is below this (1e-20 when a p
// function is involved) are skipped
constexpr double ERI_PRIM_TOL = 1e-17;
// bases with fewer AO pairs are not screened: their diagonal pass would cost more than it saves
constexpr int ERI_SCREEN_MIN_PAIRS = 2048;

// x >= y with idx = x (x + 1) / 2 + y
__device__ __forceinline__ void tri_split(int idx, int* x, int* y) {
  int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > idx) --r;
  while ((r + 1) * (r + 2) / 2 <= idx) ++r;
  *x = r;
  *y = idx - r * (r + 1) / 2;
}

// lanes per AO pair in the diagonal pass
constexpr int ERI_GS = 8;

__device__ __forceinline__ void lmn3(int packed, int* l) {
  l[0] = packed & 3;
  l[1] = (packed >> 2) & 3;
  l[2] = (packed >> 4) & 3;
}

// (P|P) for the diagonal pass: GS lanes (a group) share one bra AO pair P = (ij); member m takes
// primitive quartets m, m + GS, ... of the shell pair with itself and the group's partial sums are
// combined by a fixed xor tree.  CB: 0 (ss), 1 (sp), 2 (pp) pair.
template <int CB, int GS>
__device__ void eri_diag_body(const EriArgs& g, int P, int i, int j, int m) {
  constexpr int NB = h_count(CB), NR = h_count(2 * CB);
  int la[3], lb[3];
  lmn3(g.bs.ao_lmn[i], la);
  lmn3(g.bs.ao_lmn[j], lb);
  const int sa = g.bs.ao_shell[i], sb = g.bs.ao_shell[j];  // sa >= sb: shells are in AO order
  const int pp = g.maxp * g.maxp;
  const double* r1 = g.rec + (size_t)(sa * (sa + 1) / 2 + sb) * pp * SREC;
  const int npa = g.bs.sh_nprim[sa] * g.bs.sh_nprim[sb], nq = npa * npa;
  double acc = 0.0;
  for (int qi = m; qi < nq; qi += GS) {
    const int a = qi / npa, c = qi - a * npa;
    const double* x1 = r1 + a * SREC;
    const double* x2 = r1 + c * SREC;
    const double c1 = x1[4], c2 = x2[4];
    if (c1 == 0.0 || c2 == 0.0) continue;
    const double p = x1[0], q = x2[0];
    const double rs = rsqrt(p + q);
    const double pref = x1[12] * x2[12] * (TWO_PI_2_5 * rs);
    if (fabs(pref) < (CB == 0 ? ERI_PRIM_TOL : 1e-3 * ERI_PRIM_TOL)) continue;
    double E1[NB], E2[NB];
    e_dense<CB>(la, lb, x1 + 5, x1 + 8, x1[11], E1);
    e_dense<CB>(la, lb, x2 + 5, x2 + 8, x2[11], E2);
    const double alpha = p * q * (rs * rs);
    const double X = x1[1] - x2[1], Y = x1[2] - x2[2], Z = x1[3] - x2[3];
    double F[2 * CB + 1];
    boys_upto<2 * CB>(alpha * (X * X + Y * Y + Z * Z), F);
    double R[NR];
    rtable<2 * CB>(alpha, X, Y, Z, F, R);
    double v = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < NB; ++k2) {
      double w = 0.0;
#pragma unroll
      for (int k1 = 0; k1 < NB; ++k1)
        w = fma(E1[k1], R[h_index(HT[k1] + HT[k2], HU[k1] + HU[k2], HV[k1] + HV[k2])], w);
      v = fma(E2[k2], ((HT[k2] + HU[k2] + HV[k2]) & 1) ? -w : w, v);
    }
    acc = fma(pref, v, acc);
  }
#pragma unroll
  for (int o = 1; o < GS; o <<= 1) acc += __shfl_xor(acc, o, 64);
  if (m == 0) g.diag[P] = acc;
}

// compile-time loop: f(std::integral_constant<int, 0>), ..., f(std::integral_constant<int, N - 1>)
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn&& f) {
  if constexpr (N > 0) {
    static_for<N - 1>(f);
    f(std::integral_constant<int, N - 1>{});
  }
}

// Nonzero Hermite coefficients of one AO component pair of a shell pair (K = 1; the records fold
// the Gaussian product factor into c).  Component BC of class C: C = 0 no p function; C = 1 one p
// function along BC, on the first shell when `first` is set (coefficients PA_d or PB_d, 1/2p);
// C = 2 the first shell's p along BC / 3 and the second's along BC % 3 (along one direction: the
// three coefficients of E^{11}; else the product of two E^{10}/E^{01} pairs).
// hsum<C, BC, K2>(...) = sum over those coefficients e_tuv of e_tuv * R[(t, u, v) + k2].
template <int C, int BC, int K2>
__device__ __forceinline__ double hsum(bool first, const double* PA, const double* PB, double i2p, const double* R) {
  constexpr int T2 = h_tt(K2), U2 = h_tu(K2), V2 = h_tv(K2);
  if constexpr (C == 0) {
    return R[K2];
  } else if constexpr (C == 1) {
    constexpr int D = BC;
    constexpr int up = h_index(T2 + (D == 0), U2 + (D == 1), V2 + (D == 2));
    const double e0 = first ? PA[D] : PB[D];
    return fma(e0, R[K2], i2p * R[up]);
  } else {
    constexpr int DX = BC / 3, DY = BC % 3;
    if constexpr (DX == DY) {
      constexpr int u1 = h_index(T2 + (DX == 0), U2 + (DX == 1), V2 + (DX == 2));
      constexpr int u2 = h_index(T2 + 2 * (DX == 0), U2 + 2 * (DX == 1), V2 + 2 * (DX == 2));
      const double e0 = fma(PA[DX], PB[DX], i2p), e1 = (PA[DX] + PB[DX]) * i2p, e2 = i2p * i2p;
      return fma(e0, R[K2], fma(e1, R[u1], e2 * R[u2]));
    } else {
      constexpr int ux = h_index(T2 + (DX == 0), U2 + (DX == 1), V2 + (DX == 2));
      constexpr int uy = h_index(T2 + (DY == 0), U2 + (DY == 1), V2 + (DY == 2));
      constexpr int uxy = h_index(T2 + (DX == 0) + (DY == 0), U2 + (DX == 1) + (DY == 1), V2 + (DX == 2) + (DY == 2));
      const double a0 = PA[DX], b0 = PB[DY];
      return fma(a0 * b0, R[K2], fma(i2p * b0, R[ux], fma(a0 * i2p, R[uy], i2p * i2p * R[uxy])));
    }
  }
}

// The shell quartet (SP = (sa, sb) | (sc, sd)) of classes (CB, CK) for one group of GS lanes, member
// m; r2: the ket shell pair's primitive-pair records (EREC doubles each); KCC >= 0 (pp|pp): only the
// ket components whose first index is KCC.
template <int CB, int CK, int GS, int KCC = -1>
__device__ void eri_shell_body(const EriArgs& g, int SP, int sa, int sb, const double* r2, int sc, int sd, int m) {
  constexpr int L = CB + CK, NR = h_count(L), NK = h_count(CK);
  constexpr int NBC = CB == 0 ? 1 : CB == 1 ? 3 : 9;
  constexpr int NKC = CK == 0 ? 1 : CK == 1 ? 3 : (KCC >= 0 ? 3 : 9);
  constexpr int KC0 = (CK == 2 && KCC >= 0) ? 3 * KCC : 0;
  const bool pa = g.bs.sh_l[sa] == 1, pc = g.bs.sh_l[sc] == 1;  // class 1: the p function is on the first shell
  const int pp = g.maxp * g.maxp;
  const double* r1 = g.rec + (size_t)SP * pp * SREC;
  const int npc = g.bs.sh_nprim[sc] * g.bs.sh_nprim[sd], nq = g.bs.sh_nprim[sa] * g.bs.sh_nprim[sb] * npc;
  double acc[NBC][NKC];
#pragma unroll
  for (int x = 0; x < NBC; ++x)
#pragma unroll
    for (int y = 0; y < NKC; ++y) acc[x][y] = 0.0;
  for (int qi = m; qi < nq; qi += GS) {
    const int a = qi / npc, c = qi - a * npc;
    const double* x1 = r1 + a * SREC;
    const double* x2 = r2 + c * EREC;
    const double c1 = x1[4], c2 = x2[4];
    if (c1 == 0.0 || c2 == 0.0) continue;
    const double p = x1[0], q = x2[0];
    // the quartet's (ss|ss)-type magnitude bound; p functions add Hermite factors (distances,
    // 1/2p), so their quartets get a threshold three orders lower
    const double rs = rsqrt(p + q);
    const double pref = x1[12] * x2[12] * (TWO_PI_2_5 * rs);
    if (fabs(pref) < (L == 0 ? ERI_PRIM_TOL : 1e-3 * ERI_PRIM_TOL)) continue;
    const double alpha = p * q * (rs * rs);
    const double X = x1[1] - x2[1], Y = x1[2] - x2[2], Z = x1[3] - x2[3];
    double F[L + 1];
    boys_upto<L>(alpha * (X * X + Y * Y + Z * Z), F);
    double R[NR];
    rtable<L>(alpha, X, Y, Z, F, R);
    static_for<NBC>([&](auto bc) {
      constexpr int BC = decltype(bc)::value;
      // W[k2] = (-1)^|k2| sum_L1 E^ij_L1 R_{L1 + k2}
      double W[NK];
      static_for<NK>([&](auto k2) {
        constexpr int K2 = decltype(k2)::value;
        const double w = hsum<CB, BC, K2>(pa, x1 + 5, x1 + 8, x1[11], R);
        W[K2] = ((h_tt(K2) + h_tu(K2) + h_tv(K2)) & 1) ? -w : w;
      });
      static_for<NKC>([&](auto kc) {
        constexpr int KI = decltype(kc)::value;
        acc[BC][KI] = fma(pref, hsum<CK, KC0 + KI, 0>(pc, x2 + 5, x2 + 8, x2[11], W), acc[BC][KI]);
      });
    });
  }
  const int a0 = g.ao_first[sa], b0 = g.ao_first[sb], c0 = g.ao_first[sc], d0 = g.ao_first[sd];
#pragma unroll
  for (int x = 0; x < NBC; ++x)
#pragma unroll
    for (int y = 0; y < NKC; ++y) {
      double val = acc[x][y];
#pragma unroll
      for (int o = 1; o < GS; o <<= 1) val += __shfl_xor(val, o, 64);
      const int i = CB == 0 ? a0 : CB == 1 ? (pa ? a0 + x : a0) : a0 + x / 3;
      const int j = CB == 0 ? b0 : CB == 1 ? (pa ? b0 : b0 + x) : b0 + x % 3;
      const int kc = KC0 + y;
      const int k = CK == 0 ? c0 : CK == 1 ? (pc ? c0 + kc : c0) : c0 + kc / 3;
      const int l = CK == 0 ? d0 : CK == 1 ? (pc ? d0 : d0 + kc) : d0 + kc % 3;
      if (m != 0 || i < j || k < l) continue;  // same p shell: only i >= j, k >= l
      const int P = i * (i + 1) / 2 + j, Q = k * (k + 1) / 2 + l;
      if (Q > P) continue;
      g.eri[(size_t)P * g.npair + Q] = val;
      g.eri[(size_t)Q * g.npair + P] = val;
    }
}

// the molecule b of a batch with off(b) <= idx < off(b + 1) (off non-decreasing from off(0) = 0)
template <class Off>
__device__ __forceinline__ int eri_find(int B, long idx, Off off) {
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off(mid) <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Main pass for bra class CB: blockIdx.y = sorted ket shell pair over the batch's molecules, whose
// class must be in KMASK; the gridDim.x waves of a ket take chunks of 64 / GS consecutive live bra
// shell pairs of class CB in turn, GS lanes per shell quartet (screened entries stay as the zero the
// output was cleared to).
template <int CB, int KMASK, int GS>
__global__ __launch_bounds__(64) void k_eri_cls(const EriArgs* __restrict__ gs, int B) {
  const int b = eri_find(B, blockIdx.y, [&](int x) { return (long)gs[x].m_off; });
  const EriArgs& g = gs[b];
  const int ky = blockIdx.y - g.m_off;
  const int kc0 = g.cnt[3], kc1 = g.cnt[4];
  const int ck = ky < kc0 ? 0 : ky < kc0 + kc1 ? 1 : 2;
  if (!((KMASK >> ck) & 1)) return;
  const int SQ = g.ket[ky];
  int sc, sd;
  tri_split(SQ, &sc, &sd);
  const int* lc3 = g.live_cnt + 3 * ky;
  const int nlive = lc3[CB];
  const int* list = g.live + (size_t)ky * g.nsp + (CB == 0 ? 0 : CB == 1 ? lc3[0] : lc3[0] + lc3[1]);
  const int m = threadIdx.x % GS;
  if (nlive == 0) return;
  // the ket shell pair's primitive-pair records, shared by the whole wave (LDS instead of a global
  // load per primitive quartet)
  __shared__ double kr[QEX3D_MAX_NPRIM * QEX3D_MAX_NPRIM * EREC];
  {
    const int npc = g.bs.sh_nprim[sc] * g.bs.sh_nprim[sd];
    const double* src = g.rec + (size_t)SQ * g.maxp * g.maxp * SREC;
    for (int t = threadIdx.x; t < npc * EREC; t += 64) kr[t] = src[(t / EREC) * SREC + t % EREC];
    __syncthreads();
  }
  for (int b0 = blockIdx.x * (64 / GS); b0 < nlive; b0 += gridDim.x * (64 / GS)) {
    const int bx = b0 + threadIdx.x / GS;
    if (bx >= nlive) break;
    const int SP = list[bx];
    int sa, sb;
    tri_split(SP, &sa, &sb);
    if (ck == 0) {
      eri_shell_body<CB, 0, GS>(g, SP, sa, sb, kr, sc, sd, m);
    } else if (ck == 1) {
      eri_shell_body<CB, 1, GS>(g, SP, sa, sb, kr, sc, sd, m);
    } else if constexpr (CB == 2) {
      eri_shell_body<2, 2, GS, 0>(g, SP, sa, sb, kr, sc, sd, m);
      eri_shell_body<2, 2, GS, 1>(g, SP, sa, sb, kr, sc, sd, m);
      eri_shell_body<2, 2, GS, 2>(g, SP, sa, sb, kr, sc, sd, m);
    } else {
      eri_shell_body<CB, 2, GS>(g, SP, sa, sb, kr, sc, sd, m);
    }
  }
}

// The main pass with GS lanes per shell quartet and xw waves per ket shell pair over nket kets.
template <int GS>
void eri_main(const EriArgs* gs, int B, int nket, int xw, hipStream_t st) {
  const dim3 eg((unsigned)xw, (unsigned)nket);
  hipLaunchKernelGGL((k_eri_cls<2, 4, GS>), eg, dim3(64), 0, st, gs, B);
  hipLaunchKernelGGL((k_eri_cls<2, 3, GS>), eg, dim3(64), 0, st, gs, B);
  hipLaunchKernelGGL((k_eri_cls<1, 7, GS>), eg, dim3(64), 0, st, gs, B);
  hipLaunchKernelGGL((k_eri_cls<0, 7, GS>), eg, dim3(64), 0, st, gs, B);
}

// Work lists of the main pass: one workgroup per sorted ket shell pair ky compacts, class segment
// by class segment of the sorted bra list, the bra shell pairs holding a pair P >= qmin whose
// Schwarz product passes.
__global__ __launch_bounds__(256) void k_eri_live(const EriArgs* __restrict__ gs, int B) {
  __shared__ int wtot[4];
  __shared__ int base;
  const int bm = eri_find(B, blockIdx.x, [&](int x) { return (long)gs[x].k_off; });
  const EriArgs& g = gs[bm];
  const int ky = blockIdx.x - g.k_off, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int SQ = g.ket[ky];
  int sc, sd;
  tri_split(SQ, &sc, &sd);
  const int qmin = g.ao_first[sc] * (g.ao_first[sc] + 1) / 2 + g.ao_first[sd];
  const double sk = g.sq[SQ];
  int* out = g.live + (size_t)ky * g.nsp;
  int seg0 = 0, off = 0;
  for (int c = 0; c < 3; ++c) {
    const int seg1 = seg0 + g.cnt[c];
    if (tid == 0) base = 0;
    __syncthreads();
    for (int b0 = seg0; b0 < seg1; b0 += 256) {
      const int bx = b0 + tid;
      bool flag = false;
      if (bx < seg1) {
        const int SP = g.bra[bx];
        int sa, sb;
        tri_split(SP, &sa, &sb);
        const int i = g.ao_first[sa] + (g.bs.sh_l[sa] == 0 ? 0 : 2);
        const int j = g.ao_first[sb] + (g.bs.sh_l[sb] == 0 ? 0 : 2);
        flag = i * (i + 1) / 2 + j >= qmin && g.sq[SP] * sk >= ERI_SCHWARZ_TOL;
      }
      const unsigned long long msk = __ballot(flag);
      if (lane == 0) wtot[wv] = __popcll(msk);
      __syncthreads();
      int wb = 0;
      for (int w = 0; w < wv; ++w) wb += wtot[w];
      if (flag) out[off + base + wb + __popcll(msk & ((1ull << lane) - 1))] = g.bra[bx];
      __syncthreads();
      if (tid == 0) base += wtot[0] + wtot[1] + wtot[2] + wtot[3];
      __syncthreads();
    }
    if (tid == 0) g.live_cnt[3 * ky + c] = base;
    off += base;
    seg0 = seg1;
    __syncthreads();
  }
}

// Diagonal pass over the screened molecules' AO pairs: (ij|ij) for every AO pair P = (ij), into
// diag[P].
__global__ __launch_bounds__(64) void k_eri_diag(const EriArgs* __restrict__ gs, int B, int total) {
  const int t = blockIdx.x * (64 / ERI_GS) + threadIdx.x / ERI_GS, m = threadIdx.x % ERI_GS;
  if (t >= total) return;
  const EriArgs& g = gs[eri_find(B, t, [&](int x) { return (long)gs[x].p_off; })];
  const int P = t - g.p_off;
  int i, j;
  tri_split(P, &i, &j);
  const int cb = (g.bs.ao_lmn[i] != 0) + (g.bs.ao_lmn[j] != 0);
  if (cb == 0) eri_diag_body<0, ERI_GS>(g, P, i, j, m);
  else if (cb == 1) eri_diag_body<1, ERI_GS>(g, P, i, j, m);
  else eri_diag_body<2, ERI_GS>(g, P, i, j, m);
}

// sq[SP] = max over the AO pairs (ij) of shell pair SP of sqrt|(ij|ij)| (1 for unscreened molecules)
__global__ void k_eri_schwarz(const EriArgs* __restrict__ gs, int B, int total) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const EriArgs& g = gs[eri_find(B, t, [&](int x) { return (long)gs[x].k_off; })];
  const int SP = t - g.k_off;
  if (!g.screen) {
    g.sq[SP] = 1.0;
    return;
  }
  int sa, sb;
  tri_split(SP, &sa, &sb);
  const int na = g.bs.sh_l[sa] == 0 ? 1 : 3, nb = g.bs.sh_l[sb] == 0 ? 1 : 3;
  double m = 0.0;
  for (int x = 0; x < na; ++x)
    for (int y = 0; y < nb; ++y) {
      const int i = g.ao_first[sa] + x, j = g.ao_first[sb] + y;
      if (j > i) continue;
      const int P = i * (i + 1) / 2 + j;
      m = fmax(m, sqrt(fabs(g.diag[P])));
    }
  g.sq[SP] = m;
}

// Shell-pair primitive records: p, P (3), c = c_a c_b exp(-mu |AB|^2), P - A (3), P - B (3), 1/(2p), c/p;
// record r = pa * nb + pb of the shell pair's maxp^2 slots for r < na * nb, the rest unused (c = 0).
// Primitive pairs whose Gaussian product prefactor is below 1e-20 get c = 0 (they contribute
// < 1e-18 to any integral) and are skipped.
__global__ void k_eri_records(const EriArgs* __restrict__ gs, int B, long total) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const EriArgs& g = gs[eri_find(B, t, [&](int x) { return gs[x].rec_off; })];
  const Basis& b = g.bs;
  const long idx = t - g.rec_off;
  const int pp = g.maxp * g.maxp;
  const int SP = (int)(idx / pp), r = (int)(idx % pp);
  int sa, sb;
  tri_split(SP, &sa, &sb);
  double* o = g.rec + idx * SREC;
  const int nb = b.sh_nprim[sb];
  if (r >= b.sh_nprim[sa] * nb) {
    o[4] = 0.0;
    return;
  }
  const int pa = r / nb, pb = r % nb;
  const double* A = b.xyz + 3 * b.sh_atom[sa];
  const double* Bc = b.xyz + 3 * b.sh_atom[sb];
  const double ea = b.exps[b.sh_off[sa] + pa], eb = b.exps[b.sh_off[sb] + pb];
  const double p = ea + eb, mu = ea * eb / p;
  const double cab = b.coefs[b.sh_off[sa] + pa] * b.coefs[b.sh_off[sb] + pb];
  const double r2 = (A[0] - Bc[0]) * (A[0] - Bc[0]) + (A[1] - Bc[1]) * (A[1] - Bc[1]) + (A[2] - Bc[2]) * (A[2] - Bc[2]);
  const double K = exp(-mu * r2);
  o[0] = p;
  for (int d = 0; d < 3; ++d) {
    const double Pd = (ea * A[d] + eb * Bc[d]) / p;
    o[1 + d] = Pd;
    o[5 + d] = Pd - A[d];
    o[8 + d] = Pd - Bc[d];
  }
  o[11] = 0.5 / p;
  o[4] = fabs(cab) * K * (M_PI / p) * sqrt(M_PI / p) < 1e-20 ? 0.0 : cab * K;
  o[12] = o[4] / p;
}

// One workgroup of 1024 per molecule: ao_first (first AO of each shell), then the work lists of
// shell pairs, each a stable partition by class l(a) + l(b), the bra list within a class by
// primitive-pair count, largest first; cnt[0..2] bra and [3..5] ket class sizes.
__global__ __launch_bounds__(1024) void k_eri_setup(const EriArgs* __restrict__ gs) {
  __shared__ int wtot[16];
  __shared__ int base;
  __shared__ unsigned long long keys;
  const EriArgs& g = gs[blockIdx.x];
  const Basis& b = g.bs;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int s = tid; s < b.nshell; s += 1024) {
    int f = 0;
    for (int t = 0; t < s; ++t) f += b.sh_l[t] == 0 ? 1 : 3;
    g.ao_first[s] = f;
  }
  const int n = g.nsp;
  for (int list = 0; list < 2; ++list) {
    int* out = list == 0 ? g.bra : g.ket;
    int off = 0;
    for (int c = 0; c < 3; ++c) {
      // the bra pairs of a class are further ordered by their primitive-pair count (largest first), so
      // the lanes of a wave loop over the same number of primitive quartets
      if (tid == 0) {
        base = 0;
        keys = 0ull;
      }
      __syncthreads();
      for (int idx = tid; list == 0 && idx < n; idx += 1024) {
        int x, y;
        tri_split(idx, &x, &y);
        if (b.sh_l[x] + b.sh_l[y] == c) atomicOr(&keys, 1ull << (b.sh_nprim[x] * b.sh_nprim[y] - 1));
      }
      __syncthreads();
      const unsigned long long present = list == 0 ? keys : 1ull;
      for (int v = 63; v >= 0; --v) {
        if (!((present >> v) & 1ull)) continue;
        for (int c0 = 0; c0 < n; c0 += 1024) {
          const int idx = c0 + tid;
          bool flag = false;
          if (idx < n) {
            int x, y;
            tri_split(idx, &x, &y);
            flag = b.sh_l[x] + b.sh_l[y] == c && (list == 1 || b.sh_nprim[x] * b.sh_nprim[y] == v + 1);
          }
          const unsigned long long m = __ballot(flag);
          if (lane == 0) wtot[wv] = __popcll(m);
          __syncthreads();
          int wb = 0;
          for (int w = 0; w < wv; ++w) wb += wtot[w];
          if (flag) out[off + base + wb + __popcll(m & ((1ull << lane) - 1))] = idx;
          __syncthreads();
          if (tid == 0)
            for (int w = 0; w < 16; ++w) base += wtot[w];
          __syncthreads();
        }
      }
      if (tid == 0) g.cnt[list * 3 + c] = base;
      off += base;
      __syncthreads();
    }
  }
}

// The pair matrices cleared before the main pass (screened entries are not written); blockIdx.y =
// molecule.
__global__ void k_eri_clear(const EriArgs* __restrict__ gs) {
  const EriArgs& g = gs[blockIdx.y];
  const size_t n = (size_t)g.npair * g.npair;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    g.eri[t] = 0.0;
}

// EriArgs records into device memory from kernel arguments (no host copy to wait for)
constexpr int ERI_PUT_N = 12;
struct EriPut {
  EriArgs r[ERI_PUT_N];
};
__global__ void k_put_eri(EriPut pk, int count, EriArgs* dst) {
  constexpr int W = sizeof(EriArgs) / 8;
  for (int t = threadIdx.x; t < count * W; t += blockDim.x) {
    const int i = t / W, w = t % W;
    reinterpret_cast<unsigned long long*>(dst + i)[w] = reinterpret_cast<const unsigned long long*>(&pk.r[i])[w];
  }
}

// ---------------------------------------------------------------------------------------------
// grid and AO values
// ---------------------------------------------------------------------------------------------
// cell-function step of the fuzzy partition: Becke's three iterations of p(mu) = 1.5 mu - 0.5 mu^3,
// or Stratmann-Scuseria-Frisch (CPL 257, 213 (1996), eq. 14): the odd polynomial
// (35 z - 35 z^3 + 21 z^5 - 5 z^7) / 16, z = mu / 0.64, and sign(mu) for |mu| >= 0.64
__device__ __forceinline__ double cell_step(int scheme, double mu) {
  if (scheme == QEX3D_PARTITION_STRATMANN) {
    constexpr double A = 0.64;
    if (mu <= -A) return -1.0;
    if (mu >= A) return 1.0;
    const double z = mu / A, z2 = z * z;
    return (1.0 / 16.0) * (z * (35.0 + z2 * (-35.0 + z2 * (21.0 - 5.0 * z2))));
  }
  for (int it = 0; it < 3; ++it) mu = 1.5 * mu - 0.5 * mu * mu * mu;
  return mu;
}

__global__ void k_grid(Basis b, int scheme, int nrad, const double* rr, const double* rw, int nang, const double* ax,
                       const double* aw, double* coords, double* weights) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nrad * nang;
  if (g >= per * b.natm) return;
  const int a = (int)(g / per);
  const int ir = (int)((g % per) / nang), ia = (int)(g % nang);
  const double* C = b.xyz + 3 * a;
  double x[3];
  for (int d = 0; d < 3; ++d) x[d] = C[d] + rr[ir] * ax[3 * ia + d];
  // fuzzy cell functions P_A = prod_B (1 - step(mu_AB)) / 2
  double Pown = 0.0, Psum = 0.0;
  for (int A = 0; A < b.natm; ++A) {
    const double* RA = b.xyz + 3 * A;
    const double dA = sqrt((x[0] - RA[0]) * (x[0] - RA[0]) + (x[1] - RA[1]) * (x[1] - RA[1]) + (x[2] - RA[2]) * (x[2] - RA[2]));
    double PA = 1.0;
    for (int Bq = 0; Bq < b.natm; ++Bq) {
      if (Bq == A) continue;
      const double* RB = b.xyz + 3 * Bq;
      const double dB = sqrt((x[0] - RB[0]) * (x[0] - RB[0]) + (x[1] - RB[1]) * (x[1] - RB[1]) + (x[2] - RB[2]) * (x[2] - RB[2]));
      const double rab = sqrt((RA[0] - RB[0]) * (RA[0] - RB[0]) + (RA[1] - RB[1]) * (RA[1] - RB[1]) + (RA[2] - RB[2]) * (RA[2] - RB[2]));
      PA *= 0.5 * (1.0 - cell_step(scheme, (dA - dB) / rab));
    }
    Psum += PA;
    if (A == a) Pown = PA;
  }
  for (int d = 0; d < 3; ++d) coords[3 * g + d] = x[d];
  weights[g] = rw[ir] * aw[ia] * (Psum > 0.0 ? Pown / Psum : 0.0);
}

__global__ void k_eval_ao(Basis b, long ng, const double* coords, double* ao) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ng * b.nshell) return;
  const long g = idx / b.nshell;
  const int s = (int)(idx % b.nshell);
  const double* C = b.xyz + 3 * b.sh_atom[s];
  const double dx = coords[3 * g] - C[0], dy = coords[3 * g + 1] - C[1], dz = coords[3 * g + 2] - C[2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  double rad = 0.0;
  for (int q = 0; q < b.sh_nprim[s]; ++q) rad += b.coefs[b.sh_off[s] + q] * exp(-b.exps[b.sh_off[s] + q] * r2);
  // first AO of the shell: shells are stored in AO order
  int mu0 = 0;
  for (int t = 0; t < s; ++t) mu0 += b.sh_l[t] == 0 ? 1 : 3;
  double* row = ao + g * b.nao;
  if (b.sh_l[s] == 0) {
    row[mu0] = rad;
  } else {
    row[mu0] = rad * dx;
    row[mu0 + 1] = rad * dy;
    row[mu0 + 2] = rad * dz;
  }
}

// The same values rounded to FP32 in the padded layout of the FP32 grid path: rows of ld floats
// (ld = nao rounded up to 16), rows past ng and columns past nao zero.  Thread per (row, shell);
// the last shell's thread of a row also clears the row's padding columns.
__global__ void k_eval_ao32(Basis b, long ng, long rows, int ld, const double* coords, float* ao) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * b.nshell) return;
  const long g = idx / b.nshell;
  const int s = (int)(idx % b.nshell);
  int mu0 = 0;
  for (int t = 0; t < s; ++t) mu0 += b.sh_l[t] == 0 ? 1 : 3;
  const int nc = b.sh_l[s] == 0 ? 1 : 3;
  float* row = ao + g * ld;
  if (s == b.nshell - 1)
    for (int c = b.nao; c < ld; ++c) row[c] = 0.0f;
  if (g >= ng) {
    for (int c = 0; c < nc; ++c) row[mu0 + c] = 0.0f;
    return;
  }
  const double* C = b.xyz + 3 * b.sh_atom[s];
  const double dx = coords[3 * g] - C[0], dy = coords[3 * g + 1] - C[1], dz = coords[3 * g + 2] - C[2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  double rad = 0.0;
  for (int q = 0; q < b.sh_nprim[s]; ++q) rad += b.coefs[b.sh_off[s] + q] * exp(-b.exps[b.sh_off[s] + q] * r2);
  if (nc == 1) {
    row[mu0] = (float)rad;
  } else {
    row[mu0] = (float)(rad * dx);
    row[mu0 + 1] = (float)(rad * dy);
    row[mu0 + 2] = (float)(rad * dz);
  }
}

// rho_g = phi_g^T D phi_g for a tile of 64 points: P = Phi_tile D on FP64 matrix cores
// (v_mfma_f64_16x16x4f64; lane l supplies A[l%16][l/16], B[l/16][l%16], receives D[4v+l/16][l%16]),
// then rho = rowsum(Phi_tile * P).  The AO tile is staged in LDS (row stride npad + 1).
typedef double d4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ d4_t mfma64(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// AO tile of 64 points x n (contiguous rows of the [ng][n] matrix) from HBM: every thread issues its
// loads at once (16-byte, up to TILE_V per thread), the caller does other work, then tile_store
// writes them to the LDS tile (row stride S, columns n..npad-1 and rows past ng zero).
template <int NTH>
struct TileRegs {
  static constexpr int V = (64 * QEX3D_MAX_NAO / 2 + NTH - 1) / NTH;
  double2 v[V];
};
template <int NTH>
__device__ __forceinline__ void tile_load(const double* ao, long g0, long ng, int n, TileRegs<NTH>& r) {
  const long rows = ng - g0 < 64 ? ng - g0 : 64;
  const long total = rows * n;  // valid doubles of the tile
  const double* src = ao + g0 * n;  // g0 is a multiple of 64: 16-byte aligned
#pragma unroll
  for (int k = 0; k < TileRegs<NTH>::V; ++k) {
    const long e = 2 * ((long)threadIdx.x + (long)k * NTH);
    if (e + 1 < total)
      r.v[k] = *reinterpret_cast<const double2*>(src + e);
    else
      r.v[k] = make_double2(e < total ? src[e] : 0.0, 0.0);
  }
}
template <int NTH>
__device__ __forceinline__ void tile_store(const TileRegs<NTH>& r, int n, int npad, int S, double* ph) {
  const int tot = 64 * n;
#pragma unroll
  for (int k = 0; k < TileRegs<NTH>::V; ++k) {
    const int e = 2 * (threadIdx.x + k * NTH);
    if (e < tot) ph[(e / n) * S + e % n] = r.v[k].x;
    if (e + 1 < tot) ph[((e + 1) / n) * S + (e + 1) % n] = r.v[k].y;
  }
  for (int t = threadIdx.x; t < 64 * (npad - n); t += NTH) ph[(t / (npad - n)) * S + n + t % (npad - n)] = 0.0;
}

// Workgroup of NT_RHO threads (one wave per 16-column block of D, n <= 96): wave w computes
// T[:, 16w:16w+16] = Phi_tile D[:, 16w:16w+16] for the 64 points (4 independent accumulators, its
// block of D held in registers as the MFMA B operand), then its partial rho = rowsum(Phi * T).
constexpr int NT_RHO = 64 * (QEX3D_MAX_NAO / 16);
__device__ void lda_point(int fn, double rho, double* e, double* v);
// With wv given: the LDA at the 64 points in the epilogue, wv = w v / 2 and ex = rho w e per point
// (the Vxc kernel then only reads them).
__device__ __forceinline__ void rho_body(long bid, long ng, int n, const double* ao, const double* dm, double* rho,
                                         const double* w = nullptr, double* wv_out = nullptr, double* ex_out = nullptr,
                                         int fn = QEX3D_LDA_SVWN5) {
  extern __shared__ double sm[];
  const int npad = (n + 15) & ~15, S = npad + 1, nb = npad / 16, ks = npad / 4;
  double* ph = sm;
  __shared__ double part[NT_RHO / 64][64];
  const long g0 = bid * 64;
  TileRegs<NT_RHO> r;
  tile_load(ao, g0, ng, n, r);
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = wv * 16 + (l & 15);
  // B operand of step k (k-slab 4k..4k+3): lane l supplies D[4k + l/16][col]
  double bv[QEX3D_MAX_NAO / 4];
#pragma unroll
  for (int k = 0; k < QEX3D_MAX_NAO / 4; ++k) {
    const int kk = 4 * k + (l >> 4);
    bv[k] = (wv < nb && k < ks && kk < n && col < n) ? dm[kk * n + col] : 0.0;
  }
  tile_store(r, n, npad, S, ph);
  __syncthreads();
  if (wv < nb) {
    d4_t acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[rb] = (d4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < QEX3D_MAX_NAO / 4; ++k) {
      if (k < ks) {
        const int kk = 4 * k + (l >> 4);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) acc[rb] = mfma64(ph[(rb * 16 + (l & 15)) * S + kk], bv[k], acc[rb]);
      }
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rb * 16 + 4 * v + (l >> 4);
        double x = ph[row * S + col] * acc[rb][v];
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        if ((l & 15) == 0) part[wv][row] = x;
      }
  }
  __syncthreads();
  if (threadIdx.x < 64 && g0 + threadIdx.x < ng) {
    const long g = g0 + threadIdx.x;
    double x = 0.0;
    for (int q = 0; q < nb; ++q) x += part[q][threadIdx.x];
    rho[g] = x;
    if (wv_out) {
      double e, v;
      lda_point(fn, x, &e, &v);
      const double ww = w[g];
      wv_out[g] = 0.5 * ww * v;
      ex_out[g] = x * ww * e;
    }
  }
}

__global__ __launch_bounds__(NT_RHO) void k_rho(long ng, int n, const double* ao, const double* dm, double* rho,
                                                const double* w, double* wv, double* ex, int fn) {
  rho_body(blockIdx.x, ng, n, ao, dm, rho, w, wv, ex, fn);
}

// rho from the occupied orbitals (inside the SCF D = 2 C_occ C_occ^T, so rho_g = 2 |phi_g^T C_occ|^2):
// 2 n_g n n_occ FLOPs instead of 2 n_g n^2.  Workgroup of NT threads (4 waves): wave w computes the
// 16 points of row block w against every 16-column block of C_occ (B operand in registers per
// block), squares and sums across the block's columns.  LDA epilogue as in rho_body.
template <int KS>
__device__ __forceinline__ void rho_occ_mfma(const double* ph, int S, const double* C, int n, int nocc, int l, int wv,
                                             double* rsum) {
  // rsum[v] accumulates, for row 16 wv + 4 v + l/16, the squares of this lane's column of each block
  for (int cb = 0; cb * 16 < nocc; ++cb) {
    const int col = cb * 16 + (l & 15);
    double bv[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int kk = 4 * k + (l >> 4);
      bv[k] = (kk < n && col < nocc) ? C[kk * n + col] : 0.0;
    }
    d4_t acc = (d4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < KS; ++k) acc = mfma64(ph[(wv * 16 + (l & 15)) * S + 4 * k + (l >> 4)], bv[k], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) rsum[v] = fma(acc[v], acc[v], rsum[v]);
  }
}

__device__ __forceinline__ void rho_occ_body(long bid, long ng, int n, int nocc, const double* ao, const double* C,
                                             double* rho, const double* w, double* wv_out, double* ex_out, int fn) {
  extern __shared__ double sm[];
  const int npad = (n + 15) & ~15, S = npad + 1;
  double* ph = sm;
  __shared__ double part[64];
  const long g0 = bid * 64;
  {
    TileRegs<NT> r;
    tile_load(ao, g0, ng, n, r);
    tile_store(r, n, npad, S, ph);
  }
  __syncthreads();
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double rsum[4] = {0.0, 0.0, 0.0, 0.0};
  switch (npad / 16) {
    case 1: rho_occ_mfma<4>(ph, S, C, n, nocc, l, wv, rsum); break;
    case 2: rho_occ_mfma<8>(ph, S, C, n, nocc, l, wv, rsum); break;
    case 3: rho_occ_mfma<12>(ph, S, C, n, nocc, l, wv, rsum); break;
    case 4: rho_occ_mfma<16>(ph, S, C, n, nocc, l, wv, rsum); break;
    case 5: rho_occ_mfma<20>(ph, S, C, n, nocc, l, wv, rsum); break;
    default: rho_occ_mfma<24>(ph, S, C, n, nocc, l, wv, rsum); break;
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double x = rsum[v];
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    if ((l & 15) == 0) part[wv * 16 + 4 * v + (l >> 4)] = 2.0 * x;
  }
  __syncthreads();
  if (threadIdx.x < 64 && g0 + threadIdx.x < ng) {
    const long g = g0 + threadIdx.x;
    const double x = part[threadIdx.x];
    rho[g] = x;
    if (wv_out) {
      double e, v;
      lda_point(fn, x, &e, &v);
      const double ww = w[g];
      wv_out[g] = 0.5 * ww * v;
      ex_out[g] = x * ww * e;
    }
  }
}

// LDA, spin unpolarised; e per electron and v = d(rho e)/d rho.  fn = QEX3D_LDA_SVWN5: Slater
// exchange + VWN5 correlation; QEX3D_LDA_SLATER: Slater exchange alone (pyscf's xc "lda").
__device__ void lda_point(int fn, double rho, double* e, double* v) {
  if (!(rho > 1e-30)) {
    *e = 0.0;
    *v = 0.0;
    return;
  }
  const double cx = -0.75 * cbrt(3.0 / M_PI);
  const double ex = cx * cbrt(rho), vx = 4.0 / 3.0 * ex;
  if (fn == QEX3D_LDA_SLATER) {
    *e = ex;
    *v = vx;
    return;
  }
  const double A = 0.0310907, x0 = -0.10498, b = 3.72744, c = 12.9352;
  const double rs = cbrt(3.0 / (4.0 * M_PI * rho));
  const double x = sqrt(rs);
  const double X = x * x + b * x + c, X0 = x0 * x0 + b * x0 + c;
  const double Q = sqrt(4.0 * c - b * b);
  const double at = atan(Q / (2.0 * x + b));
  const double ec = A * (log(x * x / X) + 2.0 * b / Q * at - b * x0 / X0 * (log((x - x0) * (x - x0) / X) + 2.0 * (b + 2.0 * x0) / Q * at));
  const double den = (2.0 * x + b) * (2.0 * x + b) + Q * Q;
  const double dec = A * (2.0 / x - (2.0 * x + b) / X - 4.0 * b / den -
                          b * x0 / X0 * (2.0 / (x - x0) - (2.0 * x + b) / X - 4.0 * (b + 2.0 * x0) / den));
  *e = ex + ec;
  *v = vx + ec - x / 6.0 * dec;
}

// Partials of V = phi^T diag(w v / 2) phi (symmetrised later), sum(rho w), sum(rho w e).
// Workgroup b processes grid chunks b, b + gridDim, ... (CH points each) and keeps its 16x16
// output tiles in registers (wave w owns tiles w, w+4, ...), so partials are per workgroup.
constexpr int CH = 64;
constexpr int NT_VXC = 512;
constexpr int NW_VXC = NT_VXC / 64;
constexpr int VX_TILES = 5;  // tiles per wave: ceil((npad/16)^2 / 8) for npad <= 96
// wv_given: per-point w v / 2 (from the LDA epilogue of the density kernel, or the neural XC);
// ex_given (optional): per-point rho w e, summed into the energy partial (zero without it).
// Workgroup bid of nblk (chunks bid, bid + nblk, ...).
__device__ __forceinline__ void vxc_body(int bid, int nblk, long ng, int n, const double* ao, const double* w,
                                         const double* rho, const double* wv_given, const double* ex_given,
                                         double* vpart, double* spart) {
  extern __shared__ double sm[];
  const int npad = (n + 15) & ~15, S = npad + 1, nb = npad / 16;
  double* ph = sm;                    // CH x S
  double* wvs = ph + (size_t)CH * S;  // CH
  __shared__ double red[NW_VXC * 2];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  d4_t acc[VX_TILES];
#pragma unroll
  for (int t = 0; t < VX_TILES; ++t) acc[t] = (d4_t){0.0, 0.0, 0.0, 0.0};
  double nel = 0.0, exc = 0.0;
  const long nchunk = (ng + CH - 1) / CH;
  for (long c = bid; c < nchunk; c += nblk) {
    const long g0 = c * CH;
    TileRegs<NT_VXC> tr;
    tile_load(ao, g0, ng, n, tr);  // in flight during the point loop below
    __syncthreads();
    for (int t = threadIdx.x; t < CH; t += NT_VXC) {
      const long g = g0 + t;
      double r = 0.0, ww = 0.0;
      if (g < ng) {
        r = rho[g];
        ww = w[g];
        if (ex_given) exc += ex_given[g];
      }
      wvs[t] = g < ng ? wv_given[g] : 0.0;
      nel += r * ww;
    }
    tile_store(tr, n, npad, S, ph);
    __syncthreads();
    // k outer, tiles inner: the wave's accumulators are independent, so its MFMAs pipeline
    for (int k0 = 0; k0 < CH; k0 += 4) {
      const int kk = k0 + (l >> 4);
      const double wk = wvs[kk];
      const double* rowk = ph + kk * S + (l & 15);
#pragma unroll
      for (int t = 0; t < VX_TILES; ++t) {
        const int tile = wv + NW_VXC * t;
        if (tile < nb * nb) {
          const int bi = tile / nb, bj = tile % nb;
          acc[t] = mfma64(rowk[bi * 16] * wk, rowk[bj * 16], acc[t]);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < VX_TILES; ++t) {
    const int tile = wv + NW_VXC * t;
    if (tile < nb * nb) {
      const int bi = tile / nb, bj = tile % nb;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = bi * 16 + 4 * v + (l >> 4), col = bj * 16 + (l & 15);
        if (row < n && col < n) vpart[(size_t)bid * n * n + row * n + col] = acc[t][v];
      }
    }
  }
  const double s0 = bsum(nel, red), s1 = bsum(exc, red + NW_VXC);
  if (threadIdx.x == 0) {
    spart[2 * bid] = s0;
    spart[2 * bid + 1] = s1;
  }
}

__global__ __launch_bounds__(NT_VXC) void k_vxc(long ng, int n, const double* ao, const double* w, const double* rho,
                                            const double* wv_given, const double* ex_given, double* vpart,
                                            double* spart) {
  vxc_body(blockIdx.x, gridDim.x, ng, n, ao, w, rho, wv_given, ex_given, vpart, spart);
}

size_t rho_lds(int n) { return (size_t)64 * (((n + 15) & ~15) + 1) * 8; }
size_t vxc_lds(int n) { return ((size_t)CH * (((n + 15) & ~15) + 1) + CH) * 8; }
constexpr int VXC_BLOCKS = 256;

void rho_dispatch(long ng, int n, const double* ao, const double* dm, double* rho, hipStream_t st,
                  const double* w = nullptr, double* wv = nullptr, double* ex = nullptr, int fn = QEX3D_LDA_SVWN5) {
  hipLaunchKernelGGL(k_rho, dim3((unsigned)((ng + 63) / 64)), dim3(NT_RHO), rho_lds(n), st, ng, n, ao, dm, rho, w, wv,
                     ex, fn);
}

__host__ __device__ inline int vxc_blocks(long ng) {
  const long nchunk = (ng + CH - 1) / CH;
  return (int)(nchunk < VXC_BLOCKS ? nchunk : VXC_BLOCKS);
}

// FP32 grid path sizes: AO row length, upper 16x16 tiles of the n x n Vxc, workgroups per molecule
// (4 waves, each taking 16-point units in turn; at most X32_BLOCKS so the partials stay small)
constexpr int X32_BLOCKS = 128;
__host__ __device__ inline int x32_ld(int n) { return (n + 15) & ~15; }
__host__ __device__ inline int x32_tiles(int n) {
  const int nb = x32_ld(n) / 16;
  return nb * (nb + 1) / 2;
}
__host__ __device__ inline int x32_blocks(long ng) {
  const long wg = (ng + 63) / 64;
  return (int)(wg < X32_BLOCKS ? wg : X32_BLOCKS);
}

// ---------------------------------------------------------------------------------------------
// global (amplitude-encoded) neural XC: x = rho[ng] -> Dense(h1) -> gelu -> Dense(h2) -> gelu ->
// Dense(1) -> -scale swish; excsum = output, vrho = d excsum / d rho.  W1 is ng x h1 (the large
// operand): one streaming pass for z1 = W1^T rho and one for vrho = W1 dz1, both HBM-bound.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double gelu_t(double x) {
  const double k = 0.7978845608028654;  // sqrt(2 / pi)
  return 0.5 * x * (1.0 + tanh(k * (x + 0.044715 * x * x * x)));
}
__device__ __forceinline__ double gelu_t_grad(double x) {
  const double k = 0.7978845608028654;
  const double t = tanh(k * (x + 0.044715 * x * x * x));
  return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * k * (1.0 + 3.0 * 0.044715 * x * x);
}

struct NNOff {
  size_t W1, b1, W2, b2, W3, b3, total;  // offsets (doubles) into the packed parameters
};
__host__ __device__ inline NNOff nn_offsets(long ng, int h1, int h2) {
  NNOff o;
  o.W1 = 0;
  o.b1 = o.W1 + (size_t)ng * h1;
  o.W2 = o.b1 + h1;
  o.b2 = o.W2 + (size_t)h1 * h2;
  o.W3 = o.b2 + h2;
  o.b3 = o.W3 + h2;
  o.total = o.b3 + 1;
  return o;
}

constexpr int GNN_BLOCKS = 1024;
__host__ __device__ inline int gnn_blocks(long ng) {
  const long b = (ng + 15) / 16;
  return (int)(b < GNN_BLOCKS ? b : GNN_BLOCKS);
}

// partial z1 over a contiguous chunk of grid rows per workgroup (8 independent row streams per thread)
__device__ __forceinline__ void gnn_l1_body(int bid, int nblk, long ng, int h1, const double* rho, const double* W1,
                                            double* zpart) {
  const long per = (ng + nblk - 1) / nblk;
  const long g0 = (long)bid * per;
  const long g1 = g0 + per < ng ? g0 + per : ng;
  for (int c = threadIdx.x; c < h1; c += NT) {
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long g = g0;
    for (; g + 7 < g1; g += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = fma(rho[g + u], W1[(size_t)(g + u) * h1 + c], a[u]);
    }
    for (; g < g1; ++g) a[0] = fma(rho[g], W1[(size_t)g * h1 + c], a[0]);
    zpart[(size_t)bid * h1 + c] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  }
}

__global__ __launch_bounds__(NT) void k_gnn_l1(long ng, int h1, const double* rho, const double* W1, double* zpart) {
  gnn_l1_body(blockIdx.x, gridDim.x, ng, h1, rho, W1, zpart);
}

// z1 = sum of the partials (fixed order) + b1; 64 columns per workgroup, 16 partial streams each
__device__ __forceinline__ void gnn_red_body(int bid, int nb, int h1, const double* zpart, const double* b1,
                                             double* z1) {
  __shared__ double acc[16][64];
  const int cl = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int c = bid * 64 + cl;
  double v = 0.0;
  if (c < h1)
    for (int b = q; b < nb; b += 16) v += zpart[(size_t)b * h1 + c];
  acc[q][cl] = v;
  __syncthreads();
  if (q == 0 && c < h1) {
    double t = 0.0;
    for (int k = 0; k < 16; ++k) t += acc[k][cl];
    z1[c] = t + b1[c];
  }
}

__global__ __launch_bounds__(1024) void k_gnn_red(int nb, int h1, const double* zpart, const double* b1, double* z1) {
  gnn_red_body(blockIdx.x, nb, h1, zpart, b1, z1);
}

// the small layers and the backward pass down to dz1 (one workgroup of 1024 threads)
__device__ __forceinline__ void gnn_mid_body(int h1, int h2, const double* P, NNOff o, double scale, const double* z1,
                                             double* dz1, double* excsum) {
  extern __shared__ double sm[];
  double* a1 = sm;        // gelu(z1)   [h1]
  double* z2 = a1 + h1;   // [h2]
  double* d2 = z2 + h2;   // dz2        [h2]
  __shared__ double red[16];
  __shared__ double gout;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
  for (int c = t; c < h1; c += blockDim.x) a1[c] = gelu_t(z1[c]);
  __syncthreads();
  for (int j = t; j < h2; j += blockDim.x) {
    double z = 0.0;
    for (int c = 0; c < h1; ++c) z = fma(a1[c], P[o.W2 + (size_t)c * h2 + j], z);
    z2[j] = z + P[o.b2 + j];
  }
  __syncthreads();
  double part = 0.0;
  for (int j = t; j < h2; j += blockDim.x) part = fma(gelu_t(z2[j]), P[o.W3 + j], part);
  part = wsum(part);
  if (lane == 0) red[wv] = part;
  __syncthreads();
  if (t == 0) {
    double z3 = 0.0;
    for (int k = 0; k < nw; ++k) z3 += red[k];
    z3 += P[o.b3];
    const double sg = 1.0 / (1.0 + exp(-z3));
    *excsum = -scale * z3 * sg;
    gout = -scale * (sg + z3 * sg * (1.0 - sg));
  }
  __syncthreads();
  for (int j = t; j < h2; j += blockDim.x) d2[j] = P[o.W3 + j] * gout * gelu_t_grad(z2[j]);
  __syncthreads();
  // dz1[c] = gelu'(z1[c]) sum_j W2[c][j] dz2[j]: one wave per row of W2 (coalesced)
  for (int c = wv; c < h1; c += nw) {
    double v = 0.0;
    for (int j = lane; j < h2; j += 64) v = fma(P[o.W2 + (size_t)c * h2 + j], d2[j], v);
    v = wsum(v);
    if (lane == 0) dz1[c] = v * gelu_t_grad(z1[c]);
  }
}

__global__ __launch_bounds__(1024) void k_gnn_mid(int h1, int h2, const double* P, NNOff o, double scale,
                                                  const double* z1, double* dz1, double* excsum) {
  gnn_mid_body(h1, h2, P, o, scale, z1, dz1, excsum);
}

// vrho[g] = sum_c W1[g][c] dz1[c]; writes wv[g] = w[g] vrho[g] / 2 (one wave per grid row)
__host__ __device__ inline int gnn_vjp_blocks(long ng) {
  const long vb = (ng + NW - 1) / NW;
  return (int)(vb < 2048 ? vb : 2048);
}

__device__ __forceinline__ void gnn_vjp_body(int bid, int nblk, long ng, int h1, const double* W1, const double* dz1,
                                             const double* w, double* wv) {
  extern __shared__ double sdz[];
  for (int c = threadIdx.x; c < h1; c += NT) sdz[c] = dz1[c];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (long g = (long)bid * NW + (threadIdx.x >> 6); g < ng; g += (long)nblk * NW) {
    const double* row = W1 + (size_t)g * h1;
    double v = 0.0;
    for (int c = lane; c < h1; c += 64) v = fma(row[c], sdz[c], v);
    v = wsum(v);
    if (lane == 0) wv[g] = w ? 0.5 * w[g] * v : v;  // without weights: vrho itself
  }
}

__global__ __launch_bounds__(NT) void k_gnn_vjp(long ng, int h1, const double* W1, const double* dz1, const double* w,
                                                double* wv) {
  gnn_vjp_body(blockIdx.x, gridDim.x, ng, h1, W1, dz1, w, wv);
}

void gnn_dispatch(long ng, int h1, int h2, double scale, const double* P, const double* rho, const double* w,
                  double* zpart, double* z1, double* dz1, double* excsum, double* wv, hipStream_t st) {
  const NNOff o = nn_offsets(ng, h1, h2);
  const int nb = gnn_blocks(ng);
  hipLaunchKernelGGL(k_gnn_l1, dim3(nb), dim3(NT), 0, st, ng, h1, rho, P + o.W1, zpart);
  hipLaunchKernelGGL(k_gnn_red, dim3((h1 + 63) / 64), dim3(1024), 0, st, nb, h1, zpart, P + o.b1, z1);
  hipLaunchKernelGGL(k_gnn_mid, dim3(1), dim3(1024), (size_t)(h1 + 2 * h2) * 8, st, h1, h2, P, o, scale, z1, dz1,
                     excsum);
  hipLaunchKernelGGL(k_gnn_vjp, dim3((unsigned)gnn_vjp_blocks(ng)), dim3(NT), (size_t)h1 * 8, st, ng, h1, P + o.W1,
                     dz1, w, wv);
}

// ---------------------------------------------------------------------------------------------
// per-point ("NN", local) neural XC: exc_g = -scale swish(y), y = MLP(rho_g) (Dense(h1) gelu Dense(h2)
// gelu Dense(1) on the scalar rho_g), vrho_g = d exc_g / d rho_g.  Workgroup: NNL_P points, 16 threads
// per point.  LDS: the layer-1 activations of the tile, then (reusing them) W1_u gelu'(z1_u); the
// layer-2 cotangents W3_j gelu'(z2_j).  Per point 2 h1 h2 FMAs forward and 2 h1 h2 backward.
// ---------------------------------------------------------------------------------------------
constexpr int NNL_P = 16;
constexpr int NNL_NT = NNL_P * 16;
__host__ __device__ inline size_t nn_local_lds(int h1, int h2) { return (size_t)NNL_P * ((h1 + 1) + (h2 + 1)) * 8; }
__host__ __device__ inline long nn_local_blocks(long ng) { return (ng + NNL_P - 1) / NNL_P; }

__device__ void nn_local_body(long bid, long ng, int h1, int h2, double scale, const double* __restrict__ P,
                              const double* __restrict__ rho, const double* __restrict__ w, double* wv, double* ex) {
  extern __shared__ double sm[];
  const int S1 = h1 + 1, S2 = h2 + 1;
  double* a1 = sm;                     // NNL_P x S1: gelu(z1), then W1 gelu'(z1)
  double* g2 = a1 + (size_t)NNL_P * S1;  // NNL_P x S2: W3 gelu'(z2)
  const double* W1 = P;
  const double* b1 = W1 + h1;
  const double* W2 = b1 + h1;
  const double* b2 = W2 + (size_t)h1 * h2;
  const double* W3 = b2 + h2;
  const double b3 = W3[h2];
  const int tid = threadIdx.x, p = tid >> 4, q = tid & 15;
  const long g = bid * NNL_P + p;
  const bool valid = g < ng;
  const double x = valid ? rho[g] : 0.0;
  for (int u = q; u < h1; u += 16) a1[p * S1 + u] = gelu_t(fma(x, W1[u], b1[u]));
  __syncthreads();
  double outp = 0.0;
  for (int j = q; j < h2; j += 16) {
    double z = b2[j];
    for (int u = 0; u < h1; ++u) z = fma(a1[p * S1 + u], W2[(size_t)u * h2 + j], z);
    outp = fma(gelu_t(z), W3[j], outp);
    g2[p * S2 + j] = W3[j] * gelu_t_grad(z);
  }
  outp += __shfl_xor(outp, 1, 64);
  outp += __shfl_xor(outp, 2, 64);
  outp += __shfl_xor(outp, 4, 64);
  outp += __shfl_xor(outp, 8, 64);
  __syncthreads();  // every thread of the tile is done with a1
  for (int u = q; u < h1; u += 16) a1[p * S1 + u] = W1[u] * gelu_t_grad(fma(x, W1[u], b1[u]));
  __syncthreads();
  // d y / d rho = sum_u W1_u gelu'(z1_u) sum_j W2[u][j] W3_j gelu'(z2_j); this thread sums its j's
  double acc = 0.0;
  for (int u = 0; u < h1; ++u) {
    double t = 0.0;
    for (int j = q; j < h2; j += 16) t = fma(W2[(size_t)u * h2 + j], g2[p * S2 + j], t);
    acc = fma(t, a1[p * S1 + u], acc);
  }
  acc += __shfl_xor(acc, 1, 64);
  acc += __shfl_xor(acc, 2, 64);
  acc += __shfl_xor(acc, 4, 64);
  acc += __shfl_xor(acc, 8, 64);
  if (q == 0 && valid) {
    const double y = outp + b3;
    const double sg = 1.0 / (1.0 + exp(-y));
    const double exc = -scale * y * sg;
    const double dexc = -scale * sg * (1.0 + y * (1.0 - sg));
    if (w) {  // the nr_rks inputs: w vrho / 2 and rho w exc
      const double ww = w[g];
      wv[g] = 0.5 * ww * (dexc * acc);
      ex[g] = x * ww * exc;
    } else {  // eval_xc: vrho and exc themselves
      wv[g] = dexc * acc;
      ex[g] = exc;
    }
  }
}

__global__ __launch_bounds__(NNL_NT) void k_nn_local(long ng, int h1, int h2, double scale, const double* P,
                                                     const double* rho, const double* w, double* wv, double* ex) {
  nn_local_body(blockIdx.x, ng, h1, h2, scale, P, rho, w, wv, ex);
}

__global__ void k_copy1(const double* src, double* dst) {
  if (threadIdx.x == 0) *dst = *src;
}

constexpr int JT = 64;  // Coulomb tile edge (pair-matrix rows and columns)
__host__ __device__ inline int jtiles(int npair) {
  const int nb = (npair + JT - 1) / JT;
  return nb * (nb + 1) / 2;
}

// J_ij = sum_kl (ij|kl) D_kl over the pair matrix (row P = pair(i,j)); K (optional) by gathers.
__host__ __device__ inline int coulomb_blocks(int npair) {
  const int b = (npair + 3) / 4;  // four rows per workgroup pass
  return b < 1024 ? b : 1024;
}

// workgroup bid of nblk: D' staged in LDS, then one wave per row P of the pair matrix (rows
// bid * NW + w, stepping by nblk * NW), 16-byte loads, wave-level reduction only.
__device__ __forceinline__ void coulomb_body(int bid, int nblk, int n, int npair, const double* eri, const double* dm,
                                             double* J, double* K) {
  extern __shared__ double sDp[];  // D'_Q = D_kl (2 - delta_kl)
  __shared__ double red[NW];
  for (int q = threadIdx.x; q < npair; q += NT) {
    int k = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
    while (k * (k + 1) / 2 > q) --k;
    while ((k + 1) * (k + 2) / 2 <= q) ++k;
    const int l = q - k * (k + 1) / 2;
    sDp[q] = k == l ? dm[k * n + l] : dm[k * n + l] + dm[l * n + k];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool even = (npair & 1) == 0;  // rows start 16-byte aligned
  for (int P = bid * NW + wv; P < npair; P += nblk * NW) {
    const double* row = eri + (size_t)P * npair;
    double s0 = 0.0, s1 = 0.0;
    if (even) {
      const double2* r2 = reinterpret_cast<const double2*>(row);
      const int h = npair / 2;
      int q = lane;
      for (; q + 64 < h; q += 128) {
        const double2 a = r2[q], b = r2[q + 64];
        s0 = fma(a.x, sDp[2 * q], s0);
        s1 = fma(a.y, sDp[2 * q + 1], s1);
        s0 = fma(b.x, sDp[2 * q + 128], s0);
        s1 = fma(b.y, sDp[2 * q + 129], s1);
      }
      for (; q < h; q += 64) {
        const double2 a = r2[q];
        s0 = fma(a.x, sDp[2 * q], s0);
        s1 = fma(a.y, sDp[2 * q + 1], s1);
      }
    } else {
      for (int q = lane; q < npair; q += 64) s0 = fma(row[q], sDp[q], s0);
    }
    const double sum = wsum(s0 + s1);
    if (lane == 0) {
      int i = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
      while (i * (i + 1) / 2 > P) --i;
      while ((i + 1) * (i + 2) / 2 <= P) ++i;
      const int j = P - i * (i + 1) / 2;
      J[i * n + j] = sum;
      J[j * n + i] = sum;
    }
  }
  if (K) {
    // K_ij = sum_kl (ik|jl) D_kl (Hartree-Fock only; O(n^4) gathers)
    for (int ij = bid; ij < n * n; ij += nblk) {
      const int i = ij / n, j = ij % n;
      double s = 0.0;
      for (int kl = threadIdx.x; kl < n * n; kl += NT) {
        const int k = kl / n, l = kl % n;
        s = fma(eri[(size_t)pair_index(i, k) * npair + pair_index(j, l)], dm[k * n + l], s);
      }
      s = bsum(s, red);
      if (threadIdx.x == 0) K[i * n + j] = s;
    }
  }
}

__global__ __launch_bounds__(NT) void k_coulomb(int n, int npair, const double* eri, const double* dm, double* J,
                                                double* K) {
  coulomb_body(blockIdx.x, gridDim.x, n, npair, eri, dm, J, K);
}

// ---------------------------------------------------------------------------------------------
// small dense linear algebra in one workgroup (row-major n x n in global memory, n <= 128)
// ---------------------------------------------------------------------------------------------
// C = op(A) op(B) (+ beta C); ta/tb transpose flags
__device__ void mm(int n, const double* A, bool ta, const double* B, bool tb, double* C, double alpha, double beta) {
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int i = idx / n, j = idx % n;
    double s = 0.0;
    for (int k = 0; k < n; ++k) {
      const double a = ta ? A[k * n + i] : A[i * n + k];
      const double b = tb ? B[j * n + k] : B[k * n + j];
      s = fma(a, b, s);
    }
    C[idx] = alpha * s + (beta != 0.0 ? beta * C[idx] : 0.0);
  }
  __syncthreads();
}

__device__ double dot_all(int len, const double* a, const double* b, double* red) {
  double s = 0.0;
  for (int t = threadIdx.x; t < len; t += blockDim.x) s = fma(a[t], b[t], s);
  return bsum(s, red);
}

// C[M x N] = alpha op(A) op(B), K the inner dimension, on FP64 matrix cores: one 16 x 16 output
// tile per wave at a time (lane l supplies A[l%16][l/16] and B[l/16][l%16] of each 16 x 4 x 16
// step and receives C[4v + l/16][l%16] in element v).  Operands in LDS or global memory.
__device__ void mfma_gemm(int M, int N, int K, const double* A, int lda, bool ta, const double* B, int ldb, bool tb,
                          double* C, int ldc, double alpha) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int mb = (M + 15) / 16, nb = (N + 15) / 16;
  for (int tile = wv; tile < mb * nb; tile += nw) {
    const int bi = tile / nb, bj = tile % nb;
    const int r = bi * 16 + (lane & 15), c = bj * 16 + (lane & 15);
    d4_t acc = (d4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 4) {
      const int k = k0 + (lane >> 4);
      double a = 0.0, b = 0.0;
      if (r < M && k < K) a = ta ? A[k * lda + r] : A[r * lda + k];
      if (c < N && k < K) b = tb ? B[c * ldb + k] : B[k * ldb + c];
      acc = mfma64(a, b, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = bi * 16 + 4 * v + (lane >> 4), col = bj * 16 + (lane & 15);
      if (row < M && col < N) C[row * ldc + col] = alpha * acc[v];
    }
  }
  __syncthreads();
}

// Cholesky S = L L^T (in place, lower), then X = L^-T (upper): C = X V solves F C = S C e.
__device__ void orthogonaliser(int n, const double* S, double* L, double* X) {
  const int NT = blockDim.x;
  for (int t = threadIdx.x; t < n * n; t += NT) L[t] = S[t];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x == 0) L[k * n + k] = sqrt(L[k * n + k]);
    __syncthreads();
    const double dkk = L[k * n + k];
    for (int i = k + 1 + threadIdx.x; i < n; i += NT) L[i * n + k] /= dkk;
    __syncthreads();
    for (int idx = threadIdx.x; idx < (n - k - 1) * (n - k - 1); idx += NT) {
      const int i = k + 1 + idx / (n - k - 1), j = k + 1 + idx % (n - k - 1);
      if (j <= i) L[i * n + j] -= L[i * n + k] * L[j * n + k];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < n * n; t += NT) {
    const int i = t / n, j = t % n;
    if (j > i) L[t] = 0.0;
  }
  __syncthreads();
  // Linv (lower) column by column: Linv[:, c] solves L x = e_c; X = Linv^T
  for (int c = threadIdx.x; c < n; c += NT) {
    for (int i = 0; i < n; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= L[i * n + k] * X[c * n + k];  // X row c holds Linv[:, c]
      X[c * n + i] = i < c ? 0.0 : s / L[i * n + i];
    }
  }
  __syncthreads();
}

// One molecule of an SCF launch.  The records of a batch sit in device memory; every per-cycle
// kernel finds its molecule from the block index (single-workgroup kernels: blockIdx.x,
// multi-workgroup kernels: blockIdx.y) and exits past that molecule's own block count.
struct Scf3 {
  int n, npair, nocc, max_cycle, diis_space, method;  // method 0 RKS-LDA, 1 RHF, 2 RKS with a global network
  long ng;
  int nchunk;
  double conv_tol, enuc;
  const double* S;
  const double* H;
  double* X;   // orthogonaliser
  double* Lc;  // Cholesky factor of S (scratch)
  int* status_out;
  const double* eri;
  const double* ao;
  const double* w;
  const float* ao32;  // QEX3D_PRECISION_FP32_GRID: AO values [rows][x32_ld(n)]
  int prec, x32blk;   // grid precision; workgroups of the FP32 grid kernels for this molecule
  // state (workspace)
  double *F, *Fd, *D, *Dlast, *C, *Cp, *mo, *J, *K, *Vxc, *vpart, *spart, *rho, *tmp1, *tmp2;
  double *diisF, *diisE, *diisB;  // [space][n*n], [space][n*n], [(space+1)^2]
  double* scal;                    // [0] e_tot, [1] e_last, [2] nelec, [3] exc (LDA), [4] gnorm, [5] exc (network)
  int* ist;                        // [0] frozen, [1] cycles done, [2] diis count, [3] diis head, [4] converged
  double* e_hist;                  // [max_cycle + 1]
  // global neural XC (method 2)
  const double* nn;
  int h1, h2;
  double nn_scale;
  double *zpart, *z1, *dz1, *wv, *ex;
  double* eigf;   // eigensolver scratch (3 n^2)
  double* jpart;  // Coulomb tile partials (RKS): jtiles(npair) x 2 x JT
};

// Fock + CDIIS (one workgroup of NT_SCF threads).  cycle < 0: plain F = H + Veff (no DIIS).
constexpr int NT_SCF = 1024;
constexpr int NW_SCF = NT_SCF / 64;
__global__ __launch_bounds__(NT_SCF) void k_scf_fock(const Scf3* __restrict__ ss, int cycle) {
  const Scf3 s = ss[blockIdx.x];
  if (s.ist[0]) return;
  __shared__ double redm[QEX3D_MAX_DIIS][NW_SCF];
  const int n = s.n, nn = n * n;
  for (int t = threadIdx.x; t < nn; t += NT_SCF) {
    double f = s.H[t] + s.J[t];
    if (s.method != 1) f += s.Vxc[t];
    if (s.method == 1) f -= 0.5 * s.K[t];
    s.F[t] = f;
  }
  __syncthreads();
  if (cycle < 1 || s.diis_space < 1) {
    for (int t = threadIdx.x; t < nn; t += NT_SCF) s.Fd[t] = s.F[t];
    __syncthreads();
    return;
  }
  // error e = F D S - S D F
  mfma_gemm(n, n, n, s.F, n, false, s.D, n, false, s.tmp1, n, 1.0);     // FD
  mfma_gemm(n, n, n, s.tmp1, n, false, s.S, n, false, s.tmp2, n, 1.0);  // FDS
  // S D F = (F D S)^T for symmetric F, D, S
  const int cnt = s.ist[2], head = s.ist[3];
  const int slot = head;
  double* Fi = s.diisF + (size_t)slot * nn;
  double* Ei = s.diisE + (size_t)slot * nn;
  for (int t = threadIdx.x; t < nn; t += NT_SCF) {
    const int i = t / n, j = t % n;
    Ei[t] = s.tmp2[i * n + j] - s.tmp2[j * n + i];
    Fi[t] = s.F[t];
  }
  __syncthreads();
  const int m = cnt + 1 < s.diis_space ? cnt + 1 : s.diis_space;
  // B (stored by slot) row/column of the new vector: the m dot products in one pass
  const int M = s.diis_space;
  {
    double d[QEX3D_MAX_DIIS];
#pragma unroll
    for (int q = 0; q < QEX3D_MAX_DIIS; ++q) d[q] = 0.0;
    for (int t = threadIdx.x; t < nn; t += NT_SCF) {
      const double e = Ei[t];
#pragma unroll
      for (int q = 0; q < QEX3D_MAX_DIIS; ++q)
        if (q < m) d[q] = fma(e, s.diisE[(size_t)q * nn + t], d[q]);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < QEX3D_MAX_DIIS; ++q) {
      const double x = wsum(d[q]);
      if (l == 0) redm[q][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < m) {
      const int q = threadIdx.x;
      double v = 0.0;
      for (int k = 0; k < NW_SCF; ++k) v += redm[q][k];
      s.diisB[slot * M + q] = v;
      s.diisB[q * M + slot] = v;
    }
    __syncthreads();
  }
  // solve [0 -1; -1 B] [l; c] = [-1; 0] in the chronological order of the stored vectors
  __shared__ double Bm[(QEX3D_MAX_DIIS + 1) * (QEX3D_MAX_DIIS + 2)];
  __shared__ double coef[QEX3D_MAX_DIIS + 1];
  __shared__ int order[QEX3D_MAX_DIIS];
  if (threadIdx.x == 0) {
    const int oldest = m < M ? 0 : (slot + 1) % M;
    for (int q = 0; q < m; ++q) order[q] = (oldest + q) % M;
    const int dim = m + 1, ld = m + 2;
    for (int r = 0; r < dim; ++r)
      for (int c = 0; c < dim; ++c) {
        double v;
        if (r == 0 && c == 0)
          v = 0.0;
        else if (r == 0 || c == 0)
          v = -1.0;
        else
          v = s.diisB[order[r - 1] * M + order[c - 1]];
        Bm[r * ld + c] = v;
      }
    for (int r = 0; r < dim; ++r) Bm[r * ld + dim] = r == 0 ? -1.0 : 0.0;
    // Gaussian elimination with partial pivoting
    for (int c = 0; c < dim; ++c) {
      int pr = c;
      for (int r = c + 1; r < dim; ++r)
        if (fabs(Bm[r * ld + c]) > fabs(Bm[pr * ld + c])) pr = r;
      if (pr != c)
        for (int k = 0; k <= dim; ++k) {
          const double t = Bm[c * ld + k];
          Bm[c * ld + k] = Bm[pr * ld + k];
          Bm[pr * ld + k] = t;
        }
      const double pv = Bm[c * ld + c];
      for (int r = c + 1; r < dim; ++r) {
        const double f = Bm[r * ld + c] / pv;
        for (int k = c; k <= dim; ++k) Bm[r * ld + k] -= f * Bm[c * ld + k];
      }
    }
    for (int r = dim - 1; r >= 0; --r) {
      double v = Bm[r * ld + dim];
      for (int k = r + 1; k < dim; ++k) v -= Bm[r * ld + k] * coef[k];
      coef[r] = v / Bm[r * ld + r];
    }
    s.ist[2] = m;
    s.ist[3] = (slot + 1) % M;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nn; t += NT_SCF) {
    double f = 0.0;
    for (int q = 0; q < m; ++q) f = fma(coef[q + 1], s.diisF[(size_t)order[q] * nn + t], f);
    s.Fd[t] = f;
  }
  __syncthreads();
}

// generalized eigh of Fd: Fp = X^T Fd X (symmetrised), Fp = Z diag(e) Z^T in LDS, C = X Z,
// D = 2 C_occ C_occ^T.  One workgroup of 1024 threads.
constexpr int NT_EIG = 1024;
size_t eigh_lds(int n) { return (size_t)2 * n * (n + 1) * 8; }

// full = 0 (the SCF cycles): only the orbitals the cycle uses (the n_occ + 2 lowest; C's other
// columns zero, their energies unrefined); full = 1 (after the last cycle, every molecule, frozen or
// not): the whole solve, whose first n_occ + 2 columns equal the cycles' bit for bit.
__global__ __launch_bounds__(NT_EIG) void k_scf_eigh(const Scf3* __restrict__ ss, int full) {
  const Scf3 s = ss[blockIdx.x];
  if (s.ist[0] && !full) return;
  extern __shared__ double sm[];
  __shared__ qexeig::Smem sh;
  const int n = s.n, nn = n * n, S = n + 1;
  double* A = sm;
  double* Z = A + (size_t)n * S;
  mfma_gemm(n, n, n, s.Fd, n, false, s.X, n, false, s.tmp1, n, 1.0);  // Fd X
  mfma_gemm(n, n, n, s.X, n, true, s.tmp1, n, false, A, S, 1.0);      // X^T Fd X (into LDS)
  for (int t = threadIdx.x; t < nn; t += blockDim.x) {
    const int i = t / n, j = t % n;
    if (i < j) {
      const double v = 0.5 * (A[i * S + j] + A[j * S + i]);
      A[i * S + j] = v;
      A[j * S + i] = v;
    }
  }
  __syncthreads();
  qexeig::symeig(n, A, Z, S, s.mo, s.eigf, sh, full ? n : s.nocc + 2);
  mfma_gemm(n, n, n, s.X, n, false, Z, S, false, s.C, n, 1.0);  // C = X Z
  mfma_gemm(n, n, s.nocc, s.C, n, false, s.C, n, true, s.D, n, 2.0);  // D = 2 C_occ C_occ^T
}

// batched generalized eigh F C = S C e, one workgroup per matrix (scratch: 8 n^2 doubles each)
__global__ __launch_bounds__(NT_EIG) void k_gen_eigh(int n, const double* F, const double* Sm, double* mo, double* C,
                                                     double* scratch) {
  extern __shared__ double sm[];
  __shared__ qexeig::Smem sh;
  const size_t nn = (size_t)n * n, b = blockIdx.x;
  const int S = n + 1;
  double* w = scratch + b * 8 * nn;
  double *L = w, *X = w + nn, *t1 = w + 2 * nn, *t2 = w + 3 * nn, *Zg = w + 4 * nn, *f = w + 5 * nn;
  orthogonaliser(n, Sm + b * nn, L, X);
  mfma_gemm(n, n, n, F + b * nn, n, false, X, n, false, t1, n, 1.0);
  mfma_gemm(n, n, n, X, n, true, t1, n, false, t2, n, 1.0);
  double* A = sm;
  double* Z = A + (size_t)n * S;
  for (int t = threadIdx.x; t < (int)nn; t += blockDim.x) {
    const int i = t / n, j = t % n;
    A[i * S + j] = 0.5 * (t2[t] + t2[j * n + i]);
  }
  __syncthreads();
  qexeig::symeig(n, A, Z, S, mo + b * n, f, sh);
  mfma_gemm(n, n, n, X, n, false, Z, S, false, C + b * nn, n, 1.0);
  (void)Zg;
}

// Reverse pass of the symmetric eigendecomposition A = V diag(w) V^T with degenerate pairs masked
// (qedft/train/td/generalized_eigensolver.py:158-219, degen_eigh_bwd): with U = V^T (exact) or U = V
// (the reference's code: its transpose with axes (-2, -1) leaves a 2-D array unchanged),
//   M = F o (U gV) + diag(gw),  F_ij = 1 / (w_j - w_i), 0 where |w_j - w_i| < eps^0.6,
//   gA = sym(V M U);
// either cotangent may be absent.  One workgroup per matrix; products on FP64 matrix cores.
__global__ __launch_bounds__(NT_EIG) void k_eigh_vjp(int n, const double* w, const double* v, const double* gw,
                                                     const double* gv, double* gA, double* ws, int exact) {
  const size_t nn = (size_t)n * n, b = blockIdx.x;
  const double* W = w + b * n;
  const double* V = v + b * nn;
  double* M = ws + b * 2 * nn;
  double* T = M + nn;
  double* G = gA + b * nn;
  const bool ut = exact != 0;  // U = V^T: op(U) reads V transposed
  if (gv) mfma_gemm(n, n, n, V, n, ut, gv + b * nn, n, false, M, n, 1.0);  // U gV
  const double thr = 4.05381695970951e-10;  // float64 eps^0.6, eps = 2^-52
  for (size_t t = threadIdx.x; t < nn; t += blockDim.x) {
    const int i = (int)(t / n), j = (int)(t % n);
    const double f = W[j] - W[i];
    double m = gv && fabs(f) >= thr ? M[t] / f : 0.0;
    if (i == j && gw) m += gw[b * n + i];
    M[t] = m;
  }
  __syncthreads();
  mfma_gemm(n, n, n, M, n, false, V, n, ut, T, n, 1.0);  // M U
  mfma_gemm(n, n, n, V, n, false, T, n, false, G, n, 1.0);  // V M U
  for (size_t t = threadIdx.x; t < nn; t += blockDim.x) {
    const int i = (int)(t / n), j = (int)(t % n);
    if (i < j) {
      const double s = 0.5 * (G[i * n + j] + G[j * n + i]);
      G[i * n + j] = s;
      G[j * n + i] = s;
    }
  }
}

// Energies and convergence (one workgroup of NT_SCF threads).
__global__ __launch_bounds__(NT_SCF) void k_scf_energy(const Scf3* __restrict__ ss, int cycle) {
  const Scf3 s = ss[blockIdx.x];
  if (s.ist[0]) return;
  __shared__ double red[NW_SCF];
  const int n = s.n, nn = n * n, nocc = s.nocc;
  double e1 = 0.0, e2 = 0.0;
  for (int t = threadIdx.x; t < nn; t += NT_SCF) {
    e1 = fma(s.D[t], s.H[t], e1);
    double g = 0.5 * s.J[t];
    if (s.method == 1) g -= 0.25 * s.K[t];
    e2 = fma(s.D[t], g, e2);
  }
  e1 = bsum(e1, red);
  e2 = bsum(e2, red);
  const double etot = e1 + e2 + (has_point_exc(s.method) ? s.scal[3] : s.method == QEX3D_RKS_NN_GLOBAL ? s.scal[5] : 0.0) +
                      s.enuc;
  if (cycle < 0) {
    if (threadIdx.x == 0) {
      s.scal[0] = etot;
      s.e_hist[0] = etot;
    }
    return;
  }
  // orbital gradient norm with the new Fock matrix: |C_vir^T F C_occ|^2 via the S^-1 projector
  // = tr(Co^T F X X^T F Co) - |Co^T F Co|^2, times 4 (the factor 2 of pyscf's get_grad)
  for (int t = threadIdx.x; t < nn; t += NT_SCF) {
    double f = s.H[t] + s.J[t];
    if (s.method != 1) f += s.Vxc[t];
    if (s.method == 1) f -= 0.5 * s.K[t];
    s.F[t] = f;
  }
  __syncthreads();
  mfma_gemm(n, nocc, n, s.F, n, false, s.C, n, false, s.tmp1, n, 1.0);    // tmp1 = F C_occ   (n x nocc)
  mfma_gemm(n, nocc, n, s.X, n, true, s.tmp1, n, false, s.tmp2, n, 1.0);  // tmp2 = X^T F C_occ
  mfma_gemm(nocc, nocc, n, s.C, n, true, s.tmp1, n, false, s.Cp, n, 1.0); // Cp = C_occ^T F C_occ
  double a = 0.0, b = 0.0;
  for (int t = threadIdx.x; t < n * nocc; t += NT_SCF) {
    const double v = s.tmp2[(t / nocc) * n + t % nocc];
    a = fma(v, v, a);
  }
  for (int t = threadIdx.x; t < nocc * nocc; t += NT_SCF) {
    const double v = s.Cp[(t / nocc) * n + t % nocc];
    b = fma(v, v, b);
  }
  a = bsum(a, red);
  b = bsum(b, red);
  if (threadIdx.x == 0) {
    const double g2 = fmax(0.0, 4.0 * (a - b));
    const int nvir = n - s.nocc;
    const double gnorm = sqrt(g2) / sqrt((double)(nvir > 0 ? nvir * s.nocc : 1));
    const double elast = s.scal[0];
    s.scal[1] = elast;
    s.scal[0] = etot;
    s.scal[4] = gnorm;
    s.e_hist[cycle + 1] = etot;
    s.ist[1] = cycle + 1;
    if (fabs(etot - elast) < s.conv_tol && gnorm < sqrt(s.conv_tol)) {
      s.ist[0] = 1;
      s.ist[4] = 1;
    } else if (cycle + 1 >= s.max_cycle) {
      s.ist[0] = 1;  // out of cycles: frozen, not converged (a batch runs to its longest member)
    }
  }
}

// state flags and scalars cleared; J, Vxc, K and D zeroed for the core-Hamiltonian guess F = H
__global__ __launch_bounds__(NT) void k_scf_init(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.x];
  if (threadIdx.x < 8) s.ist[threadIdx.x] = 0;
  if (threadIdx.x < 8) s.scal[threadIdx.x] = 0.0;
  const int nn = s.n * s.n;
  for (int t = threadIdx.x; t < nn; t += NT) {
    s.J[t] = 0.0;
    s.Vxc[t] = 0.0;
    s.K[t] = 0.0;
    s.D[t] = 0.0;
  }
}

// Cholesky and triangular inverse on LDS copies (2 n^2 doubles of dynamic LDS), X written out
__global__ __launch_bounds__(NT_SCF) void k_scf_orthogonaliser(const Scf3* __restrict__ ss) {
  extern __shared__ double sm[];
  const Scf3 s = ss[blockIdx.x];
  const int nn = s.n * s.n;
  double* Ls = sm;
  double* Xs = sm + nn;
  orthogonaliser(s.n, s.S, Ls, Xs);
  for (int t = threadIdx.x; t < nn; t += NT_SCF) {
    s.X[t] = Xs[t];
    s.Lc[t] = Ls[t];
  }
}

// fill e_hist beyond the last cycle with the final energy (frozen state repeats); status out
__global__ __launch_bounds__(NT) void k_scf_finish(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.x];
  for (int c = threadIdx.x; c <= s.max_cycle; c += blockDim.x)
    if (c > s.ist[1]) s.e_hist[c] = s.scal[0];
  if (threadIdx.x < 8) s.status_out[threadIdx.x] = s.ist[threadIdx.x];
}

// molecules of the batch still iterating (one thread; the host reads it when it polls)
__global__ void k_scf_active(const Scf3* __restrict__ ss, int B, int* out) {
  if (threadIdx.x != 0) return;
  int c = 0;
  for (int b = 0; b < B; ++b) c += ss[b].ist[0] ? 0 : 1;
  out[0] = c;
}

// The records reach the device as kernel arguments, PUT_N at a time (no host copy, stream ordered).
constexpr int PUT_N = 4;
static_assert(sizeof(Scf3) % 8 == 0, "Scf3 is copied in 8-byte words");
struct ScfPut {
  Scf3 r[PUT_N];
};
__global__ void k_put_scf(ScfPut pk, int count, Scf3* dst) {
  constexpr int W = sizeof(Scf3) / 8;
  for (int t = threadIdx.x; t < count * W; t += blockDim.x) {
    const int i = t / W, w = t % W;
    reinterpret_cast<unsigned long long*>(dst + i)[w] = reinterpret_cast<const unsigned long long*>(&pk.r[i])[w];
  }
}

// ---- per-molecule wrappers of the grid and Coulomb kernels (blockIdx.y = molecule) ----
// the SCF's density on the grid from its occupied orbitals (D = 2 C_occ C_occ^T by construction)
__global__ __launch_bounds__(NT) void k_rho_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.method == QEX3D_RHF || s.prec || s.ist[0] || (long)blockIdx.x >= (s.ng + 63) / 64) return;
  if (is_lda(s.method))
    rho_occ_body(blockIdx.x, s.ng, s.n, s.nocc, s.ao, s.C, s.rho, s.w, s.wv, s.ex,
                 s.method == QEX3D_RKS_LDA_X ? QEX3D_LDA_SLATER : QEX3D_LDA_SVWN5);
  else
    rho_occ_body(blockIdx.x, s.ng, s.n, s.nocc, s.ao, s.C, s.rho, nullptr, nullptr, nullptr, QEX3D_LDA_SVWN5);
}

__global__ __launch_bounds__(NNL_NT) void k_nn_local_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.method != QEX3D_RKS_NN_LOCAL || s.ist[0] || (long)blockIdx.x >= nn_local_blocks(s.ng)) return;
  nn_local_body(blockIdx.x, s.ng, s.h1, s.h2, s.nn_scale, s.nn, s.rho, s.w, s.wv, s.ex);
}

__global__ __launch_bounds__(NT_VXC) void k_vxc_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.method == QEX3D_RHF || s.prec || s.ist[0] || (int)blockIdx.x >= s.nchunk) return;
  vxc_body(blockIdx.x, s.nchunk, s.ng, s.n, s.ao, s.w, s.rho, s.wv, has_point_exc(s.method) ? s.ex : nullptr,
           s.vpart, s.spart);
}

__global__ __launch_bounds__(NT) void k_gnn_l1_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  const int nb = gnn_blocks(s.ng);
  if (s.method != QEX3D_RKS_NN_GLOBAL || s.ist[0] || (int)blockIdx.x >= nb) return;
  gnn_l1_body(blockIdx.x, nb, s.ng, s.h1, s.rho, s.nn + nn_offsets(s.ng, s.h1, s.h2).W1, s.zpart);
}

__global__ __launch_bounds__(1024) void k_gnn_red_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.method != QEX3D_RKS_NN_GLOBAL || s.ist[0] || (int)blockIdx.x >= (s.h1 + 63) / 64) return;
  gnn_red_body(blockIdx.x, gnn_blocks(s.ng), s.h1, s.zpart, s.nn + nn_offsets(s.ng, s.h1, s.h2).b1, s.z1);
}

__global__ __launch_bounds__(1024) void k_gnn_mid_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.method != QEX3D_RKS_NN_GLOBAL || s.ist[0]) return;
  gnn_mid_body(s.h1, s.h2, s.nn, nn_offsets(s.ng, s.h1, s.h2), s.nn_scale, s.z1, s.dz1, s.scal + 5);
}

__global__ __launch_bounds__(NT) void k_gnn_vjp_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  const int nb = gnn_vjp_blocks(s.ng);
  if (s.method != QEX3D_RKS_NN_GLOBAL || s.ist[0] || (int)blockIdx.x >= nb) return;
  gnn_vjp_body(blockIdx.x, nb, s.ng, s.h1, s.nn + nn_offsets(s.ng, s.h1, s.h2).W1, s.dz1, s.w, s.wv);
}

// Hartree-Fock (J and K): the full pair matrix, row by row
__global__ __launch_bounds__(NT) void k_coulomb_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  const int nb = coulomb_blocks(s.npair);
  if (s.method != QEX3D_RHF || s.ist[0] || (int)blockIdx.x >= nb) return;
  coulomb_body(blockIdx.x, nb, s.n, s.npair, s.eri, s.D, s.J, s.K);
}

// Kohn-Sham (J only): the lower block triangle of the symmetric pair matrix, half its bytes.
// Tile t = (pb, qb), qb <= pb, covers rows P in block pb and columns Q in block qb (JT each): its
// row sums sum_Q (P|Q) D'_Q and, off the diagonal, its column sums sum_P (P|Q) D'_P, which stand
// for the upper triangle's (Q|P).  k_coulomb_jsum_b adds a row's partials in a fixed order.
__device__ __forceinline__ double dprime(int q, int n, int npair, const double* dm) {
  if (q >= npair) return 0.0;
  int k = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
  while (k * (k + 1) / 2 > q) --k;
  while ((k + 1) * (k + 2) / 2 <= q) ++k;
  const int l = q - k * (k + 1) / 2;
  return k == l ? dm[k * n + l] : dm[k * n + l] + dm[l * n + k];
}
__global__ __launch_bounds__(256) void k_coulomb_tiles_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  const int npair = s.npair, t = blockIdx.x;
  if (s.method == QEX3D_RHF || s.ist[0] || t >= jtiles(npair)) return;
  int pb = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (pb * (pb + 1) / 2 > t) --pb;
  while ((pb + 1) * (pb + 2) / 2 <= t) ++pb;
  const int qb = t - pb * (pb + 1) / 2;
  __shared__ double dq[JT], dp[JT], colp[4][JT];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid < JT) dq[tid] = dprime(qb * JT + tid, s.n, npair, s.D);
  else if (tid < 2 * JT) dp[tid - JT] = dprime(pb * JT + tid - JT, s.n, npair, s.D);
  __syncthreads();
  const int Q = qb * JT + lane;
  const double dQ = dq[lane];
  double* rowp = s.jpart + (size_t)t * 2 * JT;
  double cs = 0.0;
  for (int r = wv; r < JT; r += 4) {
    const int P = pb * JT + r;
    const double v = (P < npair && Q < npair) ? s.eri[(size_t)P * npair + Q] : 0.0;
    const double rs = wsum(v * dQ);
    if (lane == 0) rowp[r] = rs;
    cs = fma(v, dp[r], cs);
  }
  if (pb != qb) {
    colp[wv][lane] = cs;
    __syncthreads();
    if (tid < JT) rowp[JT + tid] = ((colp[0][tid] + colp[1][tid]) + colp[2][tid]) + colp[3][tid];
  }
}
__global__ __launch_bounds__(256) void k_coulomb_jsum_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  const int npair = s.npair, P = blockIdx.x * 256 + threadIdx.x;
  if (s.method == QEX3D_RHF || s.ist[0] || P >= npair) return;
  const int nbk = (npair + JT - 1) / JT, pb = P / JT, r = P % JT;
  double sum = 0.0;
  for (int qb = 0; qb <= pb; ++qb) sum += s.jpart[(size_t)(pb * (pb + 1) / 2 + qb) * 2 * JT + r];
  for (int p2 = pb + 1; p2 < nbk; ++p2) sum += s.jpart[(size_t)(p2 * (p2 + 1) / 2 + pb) * 2 * JT + JT + r];
  int i = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > P) --i;
  while ((i + 1) * (i + 2) / 2 <= P) ++i;
  const int j = P - i * (i + 1) / 2;
  s.J[i * s.n + j] = sum;
  s.J[j * s.n + i] = sum;
}

struct Lay3 {
  size_t F, Fd, D, Dlast, C, Cp, mo, J, K, Vxc, vpart, spart, rho, tmp1, tmp2, diisF, diisE, diisB, scal, ist;
  size_t zpart, z1, dz1, wv, ex, eigf, jpart, total;
};

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

Lay3 layout3(int n, long ng, int space, int h1 = 0) {
  Lay3 L{};
  size_t off = 0;
  const size_t nn = (size_t)n * n * 8;
  const long nchunk = vxc_blocks(ng);
  auto take = [&](size_t b) {
    const size_t o = off;
    off = al(off + b);
    return o;
  };
  L.F = take(nn);
  L.Fd = take(nn);
  L.D = take(nn);
  L.Dlast = take(nn);
  L.C = take(nn);
  L.Cp = take(nn);
  L.mo = take((size_t)n * 8);
  L.J = take(nn);
  L.K = take(nn);
  L.Vxc = take(nn);
  const size_t v64 = (size_t)nchunk * nn, v32 = (size_t)x32_blocks(ng) * x32_tiles(n) * 256 * 8;
  L.vpart = take(v64 > v32 ? v64 : v32);
  L.spart = take((size_t)nchunk * 16);
  L.rho = take((size_t)ng * 8);
  L.tmp1 = take(nn);
  L.tmp2 = take(nn);
  L.diisF = take((size_t)space * nn);
  L.diisE = take((size_t)space * nn);
  L.diisB = take((size_t)space * space * 8);
  L.scal = take(64);
  L.ist = take(64);
  L.zpart = take(h1 > 0 ? (size_t)gnn_blocks(ng) * h1 * 8 : 0);
  L.z1 = take((size_t)h1 * 8);
  L.dz1 = take((size_t)h1 * 8);
  L.wv = take((size_t)ng * 8);
  L.ex = take((size_t)ng * 8);
  L.eigf = take((size_t)3 * n * n * 8);
  L.jpart = take((size_t)jtiles(n * (n + 1) / 2) * 2 * JT * 8);
  L.total = off;
  return L;
}

// the Boys table is made once per device (the first integral call waits for it)
bool g_boys_ready[64] = {false};
void ensure_boys(hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (g_boys_ready[dev]) return;
  hipLaunchKernelGGL(k_boys_init, dim3((BOYS_NT + 255) / 256), dim3(256), 0, st);
  // wait once, so that launches on any other stream also see the finished table
  if (hipStreamSynchronize(st) == hipSuccess) g_boys_ready[dev] = true;
}

Basis make_basis(const qex3d_basis* b) {
  Basis d{};
  d.natm = b->natm;
  d.nshell = b->nshell;
  d.nao = b->nao;
  d.xyz = b->atom_coords;
  d.charge = b->atom_charges;
  d.sh_atom = b->shell_atom;
  d.sh_l = b->shell_l;
  d.sh_nprim = b->shell_nprim;
  d.sh_off = b->shell_offset;
  d.exps = b->exponents;
  d.coefs = b->coefficients;
  d.ao_shell = b->ao_shell;
  d.ao_lmn = b->ao_lmn;
  return d;
}

int check_basis(const qex3d_basis* b) {
  if (!b) return fail3(QEX_E_BADARG, "null basis");
  if (b->natm < 1 || b->nshell < 1 || b->nao < 1) return fail3(QEX_E_BADARG, "empty basis");
  if (b->nao > QEX3D_MAX_NAO) return fail3(QEX_E_UNSUPPORTED, "nao %d exceeds %d", b->nao, QEX3D_MAX_NAO);
  if (b->max_nprim < 1 || b->max_nprim > QEX3D_MAX_NPRIM)
    return fail3(QEX_E_BADARG, "max_nprim must be in [1, %d]", QEX3D_MAX_NPRIM);
  if (!b->atom_coords || !b->atom_charges || !b->shell_atom || !b->shell_l || !b->shell_nprim || !b->shell_offset ||
      !b->exponents || !b->coefficients || !b->ao_shell || !b->ao_lmn)
    return fail3(QEX_E_BADARG, "missing basis array");
  return QEX_OK;
}

// P[e] = sum_c vpart[c][e] for 32 entries per workgroup (8 interleaved partial sums combined in a
// fixed order); the extra last workgroup sums the (nelec, exc) chunk partials by a fixed tree.
__host__ __device__ inline int vxc_sum_blocks(int nn) { return (nn + 31) / 32 + 1; }

__device__ __forceinline__ void vxc_sum_body(int bid, int nblk, int nn, int nchunk, const double* vpart,
                                             const double* spart, double* P, double* out2) {
  __shared__ double red[8][33];
  __shared__ double s2[2][256];
  const int t = threadIdx.x;
  if (bid == nblk - 1) {
    s2[0][t] = t < nchunk ? spart[2 * t] : 0.0;
    s2[1][t] = t < nchunk ? spart[2 * t + 1] : 0.0;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) {
        s2[0][t] += s2[0][t + h];
        s2[1][t] += s2[1][t + h];
      }
      __syncthreads();
    }
    if (t == 0) {
      out2[0] = s2[0][0];
      out2[1] = s2[1][0];
    }
    return;
  }
  const int e = bid * 32 + (t & 31), part = t >> 5;
  double a = 0.0;
  if (e < nn)
    for (int c = part; c < nchunk; c += 8) a += vpart[(size_t)c * nn + e];
  red[part][t & 31] = a;
  __syncthreads();
  if (t < 32 && e < nn) {
    double v = 0.0;
    for (int q = 0; q < 8; ++q) v += red[q][t];
    P[e] = v;
  }
}

__global__ __launch_bounds__(256) void k_vxc_sum(int nn, int nchunk, const double* vpart, const double* spart,
                                                 double* P, double* out2) {
  vxc_sum_body(blockIdx.x, gridDim.x, nn, nchunk, vpart, spart, P, out2);
}

__global__ __launch_bounds__(256) void k_vxc_sum_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  const int nn = s.n * s.n, nb = vxc_sum_blocks(nn);
  if (s.method == QEX3D_RHF || s.prec || s.ist[0] || (int)blockIdx.x >= nb) return;
  vxc_sum_body(blockIdx.x, nb, nn, s.nchunk, s.vpart, s.spart, s.tmp2, s.scal + 2);
}

// vmat = P + P^T
__device__ __forceinline__ void sym_add_body(int t, int n, const double* P, double* vmat) {
  if (t < n * n) vmat[t] = P[t] + P[(t % n) * n + t / n];
}

__global__ void k_sym_add(int n, const double* P, double* vmat) {
  sym_add_body(blockIdx.x * blockDim.x + threadIdx.x, n, P, vmat);
}

__global__ __launch_bounds__(256) void k_sym_add_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.method == QEX3D_RHF || s.prec || s.ist[0]) return;
  sym_add_body(blockIdx.x * blockDim.x + threadIdx.x, s.n, s.tmp2, s.Vxc);
}

// ---------------------------------------------------------------------------------------------
// FP32 grid path (QEX3D_PRECISION_FP32_GRID).  The AO values are FP32 rows of x32_ld(n) floats
// (rows padded to a multiple of 64).  Each wave of a workgroup works alone on 16-point units
// (unit u = points 16u..16u+15; the units of wave w of workgroup b are b*4+w, then every
// 4*x32blk-th after it):
//   1. the unit's 16 x ld AO block goes HBM -> registers (issued one unit ahead) -> the wave's LDS
//      rows (row stride ld + 4 floats);
//   2. rho_p = 2 sum_o (phi_p . C_o)^2 on FP64 matrix cores (FP32 AO values, FP64 C_occ from LDS and
//      FP64 sums), the K dimension permuted so that a lane reads 4 consecutive AO values of its point
//      in one 16-byte LDS read;
//   3. the LDA (or the given w v) at the 16 points in FP64;
//   4. the unit's rank-16 update of the upper 16x16 tiles of V = phi^T diag(w v) phi on f32 MFMA.
// A wave sums its units in FP32; the 4 waves are summed in FP64 in a fixed order into the
// workgroup's partial, and k_x32_sum_b sums the partials in FP64 in a fixed order.
// f32 MFMA lane layout: A[l % 16][l / 16], B[l / 16][l % 16], D[4 (l / 16) + v][l % 16].
// ---------------------------------------------------------------------------------------------
typedef float f4_t __attribute__((ext_vector_type(4)));
template <class T>
using gptr = __attribute__((address_space(1))) T*;  // pointer into global memory
__device__ __forceinline__ f4_t mfma32(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
constexpr int NT_X32 = 256;
enum { X32_FUSED = 0, X32_RHO = 1, X32_V = 2 };  // LDA in one pass; rho only; V from the given w v
// dynamic LDS: the 4 waves' 16-row AO blocks (row stride ld + 4 floats), then C_occ in FP64
// (ld rows of 16 * ceil(nocc / 16) columns, zero padded)
size_t x32_tile_bytes(int n) { return (size_t)64 * (x32_ld(n) + 4) * 4; }
size_t x32_lds(int n, int nocc) { return x32_tile_bytes(n) + (size_t)x32_ld(n) * 16 * ((nocc + 15) / 16) * 8; }

// Order of the wave's LDS writes and reads among its own lanes.  The LDS executes one wave's
// instructions in issue order, so only the compiler must be kept from moving LDS accesses across
// this point; no wait is needed (a workgroup-scope fence here would also wait for the next unit's
// global loads in flight, serialising every unit on HBM latency).
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

template <int MODE, int NB>
__device__ void x32_body(const Scf3& s, int bid, int nblk) {
  extern __shared__ double x32_sm[];
  constexpr int LD = 16 * NB, S = LD + 4, T = NB * (NB + 1) / 2, NV = NB;
  __shared__ double sred[2][NT_X32 / 64];
  float* smf = reinterpret_cast<float*>(x32_sm);
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, q = l >> 4, i = l & 15;
  float* tile = smf + wv * 16 * S;
  const long ng = s.ng, nunit = (ng + 15) / 16;
  const int fn = s.method == QEX3D_RKS_LDA_X ? QEX3D_LDA_SLATER : QEX3D_LDA_SVWN5;
  const int nob = (s.nocc + 15) / 16;
  const bool point_exc = has_point_exc(s.method);
  f4_t acc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) acc[t] = (f4_t){0.0f, 0.0f, 0.0f, 0.0f};
  double nel = 0.0, exc = 0.0;
  // C_occ (FP64) into LDS once: Cs[mu][o] for mu < LD, o < 16 nob, zero outside n x nocc
  double* Cs = reinterpret_cast<double*>(smf + 64 * S);
  const int CW = 16 * nob;
  if (MODE != X32_V) {
    for (int t = threadIdx.x; t < LD * CW; t += NT_X32) {
      const int mu = t / CW, o = t % CW;
      Cs[t] = (mu < s.n && o < s.nocc) ? s.C[mu * s.n + o] : 0.0;
    }
    __syncthreads();
  }
  // global (not generic) pointers: their loads count only against vmcnt, so waits for LDS results
  // do not also wait for the next unit's loads in flight
  const gptr<const f4_t> ao4 = (gptr<const f4_t>)s.ao32;
  const gptr<const double> gw = (gptr<const double>)s.w;
  const long ustride = (long)nblk * 4;
  long u = (long)bid * 4 + wv;
  f4_t pre[NV];
  if (u < nunit) {
#pragma unroll
    for (int k = 0; k < NV; ++k) pre[k] = ao4[u * 4 * LD + l + 64 * k];
  }
  for (; u < nunit; u += ustride) {
    // 1. stage the unit (float4 e of the 16 x LD block: row e / (LD / 4), column 4 (e % (LD / 4)))
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int e = l + 64 * k;
      *reinterpret_cast<f4_t*>(tile + (e / (LD / 4)) * S + 4 * (e % (LD / 4))) = pre[k];
    }
    const long g = u * 16 + i;  // this lane's point in steps 2-3 (lanes i = 0..15 of each q group alike)
    // the unit's per-point inputs are loaded before the next unit's AO block, so waiting for them
    // leaves that prefetch in flight
    const bool pt = l < 16 && g < ng;
    const double ww = pt ? gw[g] : 0.0;
    double rin = 0.0, wvin = 0.0, exin = 0.0;
    if (MODE == X32_V && pt) {
      rin = ((gptr<const double>)s.rho)[g];
      wvin = ((gptr<const double>)s.wv)[g];
      if (point_exc) exin = ((gptr<const double>)s.ex)[g];
    }
    {
      // unconditional (the last unit reloads itself), so the compiler's wait for the per-point loads
      // above counts these as still in flight instead of waiting for everything
      const long un = u + ustride < nunit ? u + ustride : u;
#pragma unroll
      for (int k = 0; k < NV; ++k) pre[k] = ao4[un * 4 * LD + l + 64 * k];
    }
    wave_lds_sync();
    double rp = 0.0;
    if (MODE != X32_V) {
      // 2. D[p][o] = sum_mu phi[p][mu] C[mu][o] on FP64 matrix cores (the FP32 AO values are exact in
      //    FP64, so rho is a smooth function of C and the SCF energy carries no rounding noise from
      //    it): lane (point i, K slot q) supplies phi[i][16 c + 4 q + j] and C[16 c + 4 q + j][16 ob + i]
      double sq[4] = {0.0, 0.0, 0.0, 0.0};
      for (int ob = 0; ob < nob; ++ob) {
        d4_t ra[2] = {(d4_t){0.0, 0.0, 0.0, 0.0}, (d4_t){0.0, 0.0, 0.0, 0.0}};
#pragma unroll 1
        for (int c = 0; c < NB; ++c) {
          const f4_t x = *reinterpret_cast<const f4_t*>(tile + i * S + 16 * c + 4 * q);
          const double* cb = Cs + (16 * c + 4 * q) * CW + 16 * ob + i;
#pragma unroll
          for (int j = 0; j < 4; ++j) ra[j & 1] = mfma64((double)x[j], cb[j * CW], ra[j & 1]);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double d = ra[0][v] + ra[1][v];
          sq[v] = fma(d, d, sq[v]);
        }
      }
      // FP64 layout D[4 v + l / 16][l % 16]: sum over the 16 orbitals (lanes of the q group) gives
      // rho / 2 at point 4 v + q
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = sq[v];
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        sq[v] = x;
      }
      // lane i takes point i = 4 v + q from lane 16 (i % 4), slot i / 4
      const int src = 16 * (i & 3);
      const double t0 = __shfl(sq[0], src, 64), t1 = __shfl(sq[1], src, 64), t2 = __shfl(sq[2], src, 64),
                   t3 = __shfl(sq[3], src, 64);
      const int sl = i >> 2;
      rp = 2.0 * (sl == 0 ? t0 : sl == 1 ? t1 : sl == 2 ? t2 : t3);
    }
    // 3. the XC at the unit's points (lanes 0..15)
    double wvp = 0.0;
    if (pt) {
      if (MODE == X32_RHO) {
        s.rho[g] = rp;
      } else if (MODE == X32_FUSED) {
        double e, v;
        lda_point(fn, rp, &e, &v);
        wvp = ww * v;
        nel = fma(rp, ww, nel);
        exc = fma(rp * ww, e, exc);
      } else {
        wvp = 2.0 * wvin;
        nel = fma(rin, ww, nel);
        exc += exin;
      }
    }
    if (MODE != X32_RHO) {
      // 4. V[16 bi + r][16 bj + c] += sum_p (w v)_p phi[p][16 bi + r] phi[p][16 bj + c], bi <= bj,
      //    4 points per MFMA (K slot q = point 4 ks + q)
      const float wf = (float)wvp;
#pragma unroll 1
      for (int ks = 0; ks < 4; ++ks) {
        const float wk = __shfl(wf, 4 * ks + q, 64);
        const float* rowp = tile + (4 * ks + q) * S + i;
        float a[NB], aw[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          a[c] = rowp[16 * c];
          aw[c] = a[c] * wk;
        }
        int t = 0;
#pragma unroll
        for (int bi = 0; bi < NB; ++bi)
#pragma unroll
          for (int bj = bi; bj < NB; ++bj) {
            acc[t] = mfma32(aw[bi], a[bj], acc[t]);
            ++t;
          }
      }
    }
    wave_lds_sync();  // the next unit's staging overwrites these rows
  }
  if (MODE == X32_RHO) return;
  // the 4 waves' FP32 tiles through the LDS (NB tiles per pass), summed per element in FP64 in wave
  // order into the workgroup partial
  __syncthreads();
  float* red = smf;  // [4 waves][NB tiles][256]
  double* vp = s.vpart + (size_t)bid * T * 256;
  constexpr int NP = (T + NB - 1) / NB;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
#pragma unroll
    for (int t = 0; t < T; ++t)
      if (t / NB == p)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[(wv * NB + t - p * NB) * 256 + v * 64 + l] = acc[t][v];
    __syncthreads();
    const int cnt = T - p * NB < NB ? T - p * NB : NB;
    for (int e = threadIdx.x; e < cnt * 256; e += NT_X32) {
      double x = 0.0;
#pragma unroll
      for (int w = 0; w < NT_X32 / 64; ++w) x += (double)red[w * NB * 256 + e];
      vp[p * NB * 256 + e] = x;
    }
    __syncthreads();
  }
  const double a = wsum(nel), b = wsum(exc);
  if (l == 0) {
    sred[0][wv] = a;
    sred[1][wv] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int w = 0; w < NT_X32 / 64; ++w) {
      x += sred[0][w];
      y += sred[1][w];
    }
    s.spart[2 * bid] = x;
    s.spart[2 * bid + 1] = y;
  }
}

// one kernel per AO block count NB (registers sized for it); molecules of other widths exit
template <int MODE, int NB>
__global__ __launch_bounds__(NT_X32) void k_x32_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.prec != QEX3D_PRECISION_FP32_GRID || s.method == QEX3D_RHF || s.ist[0]) return;
  if ((MODE == X32_FUSED) != is_lda(s.method) || x32_ld(s.n) != 16 * NB) return;
  if ((int)blockIdx.x >= s.x32blk) return;
  x32_body<MODE, NB>(s, blockIdx.x, s.x32blk);
}

template <int MODE, int NB>
void x32_launch1(const long* blk, const size_t* lds, unsigned nmol, hipStream_t st, const Scf3* dss) {
  if (lds[NB] > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_x32_b<MODE, NB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds[NB]);
  hipLaunchKernelGGL((k_x32_b<MODE, NB>), dim3((unsigned)blk[NB], nmol), dim3(NT_X32), lds[NB], st, dss);
}
template <int MODE>
void x32_launch(int nbmask, const long* blk, const size_t* lds, unsigned nmol, hipStream_t st, const Scf3* dss) {
  if (nbmask & 2) x32_launch1<MODE, 1>(blk, lds, nmol, st, dss);
  if (nbmask & 4) x32_launch1<MODE, 2>(blk, lds, nmol, st, dss);
  if (nbmask & 8) x32_launch1<MODE, 3>(blk, lds, nmol, st, dss);
  if (nbmask & 16) x32_launch1<MODE, 4>(blk, lds, nmol, st, dss);
  if (nbmask & 32) x32_launch1<MODE, 5>(blk, lds, nmol, st, dss);
  if (nbmask & 64) x32_launch1<MODE, 6>(blk, lds, nmol, st, dss);
}

// Vxc = the fixed-order FP64 sum of the workgroup partials (packed upper tiles, element
// t * 256 + v * 64 + l = tile t's D[4 (l / 16) + v][l % 16]), written to both triangles; the extra
// last workgroup sums (nelec, exc) as in vxc_sum_body.
__global__ __launch_bounds__(256) void k_x32_sum_b(const Scf3* __restrict__ ss) {
  const Scf3 s = ss[blockIdx.y];
  if (s.prec != QEX3D_PRECISION_FP32_GRID || s.method == QEX3D_RHF || s.ist[0]) return;
  const int n = s.n, nb = x32_ld(n) / 16, T = x32_tiles(n), ne = T * 256;
  const int nblk = vxc_sum_blocks(ne);
  const int bid = blockIdx.x;
  if (bid >= nblk) return;
  __shared__ double red[8][33];
  __shared__ double s2[2][256];
  const int t = threadIdx.x, nparts = s.x32blk;
  if (bid == nblk - 1) {
    s2[0][t] = t < nparts ? s.spart[2 * t] : 0.0;
    s2[1][t] = t < nparts ? s.spart[2 * t + 1] : 0.0;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) {
        s2[0][t] += s2[0][t + h];
        s2[1][t] += s2[1][t + h];
      }
      __syncthreads();
    }
    if (t == 0) {
      s.scal[2] = s2[0][0];
      s.scal[3] = s2[1][0];
    }
    return;
  }
  const int e = bid * 32 + (t & 31), part = t >> 5;
  double a = 0.0;
  if (e < ne)
    for (int c = part; c < nparts; c += 8) a += s.vpart[(size_t)c * ne + e];
  red[part][t & 31] = a;
  __syncthreads();
  if (t < 32 && e < ne) {
    double v = 0.0;
    for (int k = 0; k < 8; ++k) v += red[k][t];
    int tt = e >> 8, bi = 0;
    while (tt >= nb - bi) {
      tt -= nb - bi;
      ++bi;
    }
    const int bj = bi + tt, ll = e & 63, vv = (e >> 6) & 3;
    const int row = 16 * bi + 4 * (ll >> 4) + vv, col = 16 * bj + (ll & 15);
    if (row < n && col < n) {
      s.Vxc[row * n + col] = v;
      if (bi != bj) s.Vxc[col * n + row] = v;
    }
  }
}

// Vxc from the k_vxc chunk partials (nchunk <= VXC_BLOCKS = 256); P is n^2 scratch
void vxc_reduce(int n, int nchunk, const double* vpart, const double* spart, double* P, double* vmat, double* out2,
                hipStream_t st) {
  const int nn = n * n;
  hipLaunchKernelGGL(k_vxc_sum, dim3(vxc_sum_blocks(nn)), dim3(256), 0, st, nn, nchunk, vpart, spart, P, out2);
  hipLaunchKernelGGL(k_sym_add, dim3((nn + 255) / 256), dim3(256), 0, st, n, P, vmat);
}

}  // namespace

extern "C" {

const char* qex3d_last_error(void) { return g_err3; }

int qex3d_int1e(const qex3d_basis* b, double* S, double* T, double* V, void* stream) {
  int rc = check_basis(b);
  if (rc) return rc;
  if (!S || !T || !V) return fail3(QEX_E_BADARG, "null output");
  const int np = b->nao * (b->nao + 1) / 2;
  ensure_boys((hipStream_t)stream);
  qexprof::Scope ps(qexprof::K_INT1E, (hipStream_t)stream);
  hipLaunchKernelGGL(k_int1e, dim3((np + 3) / 4), dim3(256), 0, (hipStream_t)stream, make_basis(b), S, T, V);
  return launch_status("qex3d_int1e");
}

// Per molecule: shell-pair records, ao_first [nshell], bra [nsp], ket [nsp], class sizes [8],
// Schwarz factors [nsp], (ij|ij) [npair], live lists [nsp][nsp] and their counts [nsp][3]; a batch
// adds its EriArgs records in front.
static size_t eri_mol_bytes(const qex3d_basis* b) {
  const long np = (long)b->nao * (b->nao + 1) / 2;
  const long nsp = (long)b->nshell * (b->nshell + 1) / 2;
  return al((size_t)nsp * b->max_nprim * b->max_nprim * SREC * 8) + al((size_t)b->nshell * 4) + al((size_t)nsp * 4) +
         al((size_t)nsp * 4) + al(64) + al((size_t)nsp * 8) + al((size_t)np * 8) + al((size_t)nsp * nsp * 4) +
         al((size_t)nsp * 3 * 4);
}

size_t qex3d_eri_batch_workspace_bytes(int32_t B, const qex3d_basis* bases) {
  if (B < 1 || B > QEX3D_MAX_ERI_BATCH || !bases) return 0;
  size_t s = al((size_t)B * sizeof(EriArgs));
  for (int i = 0; i < B; ++i) {
    if (check_basis(bases + i) != QEX_OK) return 0;
    s += eri_mol_bytes(bases + i);
  }
  return s;
}

size_t qex3d_eri_workspace_bytes(const qex3d_basis* b) { return qex3d_eri_batch_workspace_bytes(1, b); }

int qex3d_eri_batch(int32_t B, const qex3d_basis* bases, double* const* eri, void* workspace, size_t workspace_bytes,
                    void* stream) {
  if (B < 1 || B > QEX3D_MAX_ERI_BATCH) return fail3(QEX_E_BADARG, "batch size %d outside 1..%d", B, QEX3D_MAX_ERI_BATCH);
  if (!bases || !eri) return fail3(QEX_E_BADARG, "null bases or outputs");
  for (int i = 0; i < B; ++i) {
    const int rc = check_basis(bases + i);
    if (rc) return rc;
    if (!eri[i]) return fail3(QEX_E_BADARG, "null output of molecule %d", i);
  }
  const size_t need = qex3d_eri_batch_workspace_bytes(B, bases);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  hipStream_t st = (hipStream_t)stream;
  ensure_boys(st);
  char* w = (char*)workspace;
  EriArgs* dgs = (EriArgs*)w;
  w += al((size_t)B * sizeof(EriArgs));
  // lanes per shell quartet, per molecule (so a molecule's result does not depend on its batch): a
  // molecule has about nsp^2 / 2 shell quartets, each up to max_nprim^4 primitive quartets; small
  // ones split a quartet over a whole wave (latency), larger ones over 16 lanes (throughput).  The
  // records are ordered wave-split molecules first, so each main-pass launch covers a contiguous run.
  auto wide = [&](int i) {
    const double nsp = 0.5 * bases[i].nshell * (bases[i].nshell + 1);
    return 0.5 * nsp * nsp < 16384.0;
  };
  std::vector<int> order;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < B; ++i)
      if (wide(i) == (pass == 0)) order.push_back(i);
  std::vector<EriArgs> gs((size_t)B);
  long nrec = 0;
  int nket = 0, ndiag = 0, npmax = 0, nwide = 0, kwide = 0;
  for (int o = 0; o < B; ++o) {
    const int i = order[o];
    const qex3d_basis* b = bases + i;
    EriArgs& g = gs[o];
    g = EriArgs{};
    g.bs = make_basis(b);
    g.nao = b->nao;
    g.npair = b->nao * (b->nao + 1) / 2;
    g.nsp = b->nshell * (b->nshell + 1) / 2;
    g.maxp = b->max_nprim;
    g.screen = g.npair >= ERI_SCREEN_MIN_PAIRS;
    g.k_off = nket;
    g.p_off = ndiag;
    g.rec_off = nrec;
    if (wide(i)) {
      g.m_off = nket;
      ++nwide;
      kwide += g.nsp;
    } else {
      g.m_off = nket - kwide;
    }
    nrec += (long)g.nsp * g.maxp * g.maxp;
    nket += g.nsp;
    if (g.screen) ndiag += g.npair;
    npmax = g.npair > npmax ? g.npair : npmax;
    g.rec = (double*)w;
    w += al((size_t)g.nsp * g.maxp * g.maxp * SREC * 8);
    g.ao_first = (int*)w;
    w += al((size_t)b->nshell * 4);
    g.bra = (int*)w;
    w += al((size_t)g.nsp * 4);
    g.ket = (int*)w;
    w += al((size_t)g.nsp * 4);
    g.cnt = (int*)w;
    w += al(64);
    g.sq = (double*)w;
    w += al((size_t)g.nsp * 8);
    g.diag = (double*)w;
    w += al((size_t)g.npair * 8);
    g.live = (int*)w;
    w += al((size_t)g.nsp * g.nsp * 4);
    g.live_cnt = (int*)w;
    w += al((size_t)g.nsp * 3 * 4);
    g.eri = eri[i];
  }
  qexprof::Scope ps(qexprof::K_ERI, st);
  for (int b0 = 0; b0 < B; b0 += ERI_PUT_N) {
    EriPut pk{};
    const int cnt = B - b0 < ERI_PUT_N ? B - b0 : ERI_PUT_N;
    for (int i = 0; i < cnt; ++i) pk.r[i] = gs[b0 + i];
    hipLaunchKernelGGL(k_put_eri, dim3(1), dim3(256), 0, st, pk, cnt, dgs + b0);
  }
  hipLaunchKernelGGL(k_eri_records, dim3((unsigned)((nrec + 255) / 256)), dim3(256), 0, st, dgs, B, nrec);
  hipLaunchKernelGGL(k_eri_setup, dim3((unsigned)B), dim3(1024), 0, st, dgs);
  if (ndiag)
    hipLaunchKernelGGL(k_eri_diag, dim3((unsigned)((ndiag + 64 / ERI_GS - 1) / (64 / ERI_GS))), dim3(64), 0, st, dgs, B,
                       ndiag);
  hipLaunchKernelGGL(k_eri_schwarz, dim3((unsigned)((nket + 255) / 256)), dim3(256), 0, st, dgs, B, nket);
  hipLaunchKernelGGL(k_eri_live, dim3((unsigned)nket), dim3(256), 0, st, dgs, B);
  const unsigned clear_x = (unsigned)((size_t)npmax * npmax / 256 / 8 + 1 < 1024 ? (size_t)npmax * npmax / 256 / 8 + 1
                                                                                : 1024);
  hipLaunchKernelGGL(k_eri_clear, dim3(clear_x, (unsigned)B), dim3(256), 0, st, dgs);
  if (nwide) eri_main<64>(dgs, nwide, kwide, 64, st);
  if (nwide < B) eri_main<16>(dgs + nwide, B - nwide, nket - kwide, 16, st);
  return launch_status("qex3d_eri_batch");
}

int qex3d_eri(const qex3d_basis* b, double* eri, void* workspace, size_t workspace_bytes, void* stream) {
  return qex3d_eri_batch(1, b, &eri, workspace, workspace_bytes, stream);
}

int qex3d_grid(const qex3d_basis* b, int32_t nrad, const double* rad_r, const double* rad_w, int32_t nang,
               const double* ang_xyz, const double* ang_w, double* coords, double* weights, void* stream) {
  return qex3d_grid_partition(b, QEX3D_PARTITION_BECKE, nrad, rad_r, rad_w, nang, ang_xyz, ang_w, coords, weights,
                              stream);
}

int qex3d_grid_partition(const qex3d_basis* b, int32_t scheme, int32_t nrad, const double* rad_r, const double* rad_w,
                         int32_t nang, const double* ang_xyz, const double* ang_w, double* coords, double* weights,
                         void* stream) {
  int rc = check_basis(b);
  if (rc) return rc;
  if (scheme != QEX3D_PARTITION_BECKE && scheme != QEX3D_PARTITION_STRATMANN)
    return fail3(QEX_E_BADARG, "bad partition scheme %d", scheme);
  if (nrad < 1 || nang < 1 || !rad_r || !rad_w || !ang_xyz || !ang_w || !coords || !weights)
    return fail3(QEX_E_BADARG, "bad grid arguments");
  const long ng = (long)b->natm * nrad * nang;
  qexprof::Scope ps(qexprof::K_GRID, (hipStream_t)stream);
  hipLaunchKernelGGL(k_grid, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, (hipStream_t)stream, make_basis(b),
                     (int)scheme, nrad, rad_r, rad_w, nang, ang_xyz, ang_w, coords, weights);
  return launch_status("qex3d_grid");
}

int qex3d_eval_ao(const qex3d_basis* b, int64_t ngrid, const double* coords, double* ao, void* stream) {
  int rc = check_basis(b);
  if (rc) return rc;
  if (ngrid < 1 || !coords || !ao) return fail3(QEX_E_BADARG, "bad eval_ao arguments");
  const long total = (long)ngrid * b->nshell;
  qexprof::Scope ps(qexprof::K_AO, (hipStream_t)stream);
  hipLaunchKernelGGL(k_eval_ao, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     make_basis(b), (long)ngrid, coords, ao);
  return launch_status("qex3d_eval_ao");
}

int qex3d_eval_ao_f32(const qex3d_basis* b, int64_t ngrid, const double* coords, float* ao32, void* stream) {
  int rc = check_basis(b);
  if (rc) return rc;
  if (ngrid < 1 || !coords || !ao32) return fail3(QEX_E_BADARG, "bad eval_ao_f32 arguments");
  if (b->nao > QEX3D_MAX_NAO) return fail3(QEX_E_UNSUPPORTED, "nao must be <= %d", QEX3D_MAX_NAO);
  const long rows = QEX3D_AO32_ROWS((long)ngrid);
  const long total = rows * b->nshell;
  qexprof::Scope ps(qexprof::K_AO, (hipStream_t)stream);
  hipLaunchKernelGGL(k_eval_ao32, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     make_basis(b), (long)ngrid, rows, (int)QEX3D_AO32_LD(b->nao), coords, ao32);
  return launch_status("qex3d_eval_ao_f32");
}

int qex3d_eval_rho(int64_t ngrid, int32_t nao, const double* ao, const double* dm, double* rho, void* stream) {
  if (ngrid < 1 || nao < 1 || nao > QEX3D_MAX_NAO || !ao || !dm || !rho) return fail3(QEX_E_BADARG, "bad eval_rho arguments");
  {
    qexprof::Scope ps(qexprof::K_RHO, (hipStream_t)stream);
    rho_dispatch(ngrid, nao, ao, dm, rho, (hipStream_t)stream);
  }
  return launch_status("qex3d_eval_rho");
}

size_t qex3d_nr_rks_workspace_bytes(int64_t ngrid, int32_t nao) {
  if (ngrid < 1 || nao < 1) return 0;
  const long nb = vxc_blocks(ngrid);
  return al((size_t)ngrid * 8) + al((size_t)nb * nao * nao * 8) + al((size_t)nb * 16) + al((size_t)nao * nao * 8) +
         2 * al((size_t)ngrid * 8);
}

int qex3d_nr_rks(int64_t ngrid, int32_t nao, const double* ao, const double* weights, const double* dm,
                 double* nelec_excsum, double* vmat, void* workspace, size_t workspace_bytes, void* stream) {
  return qex3d_nr_rks_lda(QEX3D_LDA_SVWN5, ngrid, nao, ao, weights, dm, nelec_excsum, vmat, workspace, workspace_bytes,
                          stream);
}

int qex3d_nr_rks_lda(int32_t functional, int64_t ngrid, int32_t nao, const double* ao, const double* weights,
                     const double* dm, double* nelec_excsum, double* vmat, void* workspace, size_t workspace_bytes,
                     void* stream) {
  if (functional != QEX3D_LDA_SVWN5 && functional != QEX3D_LDA_SLATER) return fail3(QEX_E_BADARG, "bad LDA functional");
  if (ngrid < 1 || nao < 1 || nao > QEX3D_MAX_NAO || !ao || !weights || !dm || !nelec_excsum || !vmat)
    return fail3(QEX_E_BADARG, "bad nr_rks arguments");
  const size_t need = qex3d_nr_rks_workspace_bytes(ngrid, nao);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  hipStream_t st = (hipStream_t)stream;
  const int nb = vxc_blocks(ngrid);
  double* rho = (double*)workspace;
  double* vpart = (double*)((char*)workspace + al((size_t)ngrid * 8));
  double* spart = (double*)((char*)vpart + al((size_t)nb * nao * nao * 8));
  double* wv = (double*)((char*)spart + al((size_t)nb * 16) + al((size_t)nao * nao * 8));
  double* ex = (double*)((char*)wv + al((size_t)ngrid * 8));
  {
    qexprof::Scope ps(qexprof::K_RHO, st);
    rho_dispatch(ngrid, nao, ao, dm, rho, st, weights, wv, ex, functional);
  }
  {
    qexprof::Scope ps(qexprof::K_VXC, st);
    hipLaunchKernelGGL(k_vxc, dim3((unsigned)nb), dim3(NT_VXC), vxc_lds(nao), st, (long)ngrid, nao, ao, weights, rho,
                       wv, ex, vpart, spart);
  }
  {
    qexprof::Scope ps(qexprof::K_OTHER, st);
    double* P = (double*)((char*)spart + al((size_t)nb * 16));
    vxc_reduce(nao, nb, vpart, spart, P, vmat, nelec_excsum, st);
  }
  return launch_status("qex3d_nr_rks");
}

// grid size the SCF workspace is laid out for (Hartree-Fock keeps a 1-point dummy)
static long scf_ngrid(const qex3d_scf_problem* p) { return p->method == QEX3D_RHF ? 1 : (long)p->ngrid; }
static int scf_space(const qex3d_scf_problem* p) { return p->diis_space > 0 ? p->diis_space : 1; }
static int scf_h1(const qex3d_scf_problem* p) { return p->method == QEX3D_RKS_NN_GLOBAL ? p->nn_h1 : 0; }
static bool nn_widths_ok(int h1, int h2) {
  return h1 >= 1 && h2 >= 1 && h1 <= QEX3D_MAX_NN_HIDDEN && h2 <= QEX3D_MAX_NN_HIDDEN;
}

// one molecule's region: the state layout, then the Cholesky factor and the orthogonaliser
static size_t scf_mol_bytes(const qex3d_scf_problem* p) {
  if (!p || p->nao < 1 || p->nao > QEX3D_MAX_NAO || scf_ngrid(p) < 1) return 0;
  if (p->method == QEX3D_RKS_NN_GLOBAL && !nn_widths_ok(p->nn_h1, p->nn_h2)) return 0;
  return al(layout3(p->nao, scf_ngrid(p), scf_space(p), scf_h1(p)).total + al((size_t)p->nao * p->nao * 8) * 2);
}
// the records of a batch and the host-poll counter come first
static size_t scf_head_bytes(int B) { return al((size_t)B * sizeof(Scf3)) + al(64); }

size_t qex3d_scf_workspace_bytes(const qex3d_scf_problem* p) {
  const size_t m = scf_mol_bytes(p);
  return m ? scf_head_bytes(1) + m : 0;
}

size_t qex3d_scf_batch_workspace_bytes(int32_t batch, const qex3d_scf_problem* problems) {
  if (batch < 1 || batch > QEX3D_MAX_SCF_BATCH || !problems) return 0;
  size_t total = scf_head_bytes(batch);
  for (int b = 0; b < batch; ++b) {
    const size_t m = scf_mol_bytes(problems + b);
    if (!m) return 0;
    total += m;
  }
  return total;
}

size_t qex3d_nr_rks_nn_workspace_bytes(int64_t ngrid, int32_t nao, int32_t h1, int32_t h2) {
  if (ngrid < 1 || nao < 1 || !nn_widths_ok(h1, h2)) return 0;
  return qex3d_nr_rks_workspace_bytes(ngrid, nao) + al((size_t)gnn_blocks(ngrid) * h1 * 8) + 2 * al((size_t)h1 * 8) +
         al((size_t)ngrid * 8) + al(64);
}

int qex3d_nr_rks_nn(int64_t ngrid, int32_t nao, const double* ao, const double* weights, const double* dm,
                    const double* nn_params, int32_t h1, int32_t h2, double scale, double* nelec_excsum,
                    double* vmat, void* workspace, size_t workspace_bytes, void* stream) {
  if (ngrid < 1 || nao < 1 || nao > QEX3D_MAX_NAO || !ao || !weights || !dm || !nn_params || !nelec_excsum || !vmat)
    return fail3(QEX_E_BADARG, "bad nr_rks_nn arguments");
  if (!nn_widths_ok(h1, h2)) return fail3(QEX_E_BADARG, "hidden widths must be in [1, %d]", QEX3D_MAX_NN_HIDDEN);
  const size_t need = qex3d_nr_rks_nn_workspace_bytes(ngrid, nao, h1, h2);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  hipStream_t st = (hipStream_t)stream;
  const int nb = vxc_blocks(ngrid);
  char* w8 = (char*)workspace;
  double* rho = (double*)w8;
  double* vpart = (double*)(w8 + al((size_t)ngrid * 8));
  double* spart = (double*)((char*)vpart + al((size_t)nb * nao * nao * 8));
  char* q = w8 + qex3d_nr_rks_workspace_bytes(ngrid, nao);
  double* zpart = (double*)q;
  q += al((size_t)gnn_blocks(ngrid) * h1 * 8);
  double* z1 = (double*)q;
  q += al((size_t)h1 * 8);
  double* dz1 = (double*)q;
  q += al((size_t)h1 * 8);
  double* wv = (double*)q;
  q += al((size_t)ngrid * 8);
  double* exc = (double*)q;
  {
    qexprof::Scope ps(qexprof::K_RHO, st);
    rho_dispatch(ngrid, nao, ao, dm, rho, st);
  }
  {
    qexprof::Scope ps(qexprof::K_NN, st);
    gnn_dispatch(ngrid, h1, h2, scale, nn_params, rho, weights, zpart, z1, dz1, exc, wv, st);
  }
  {
    qexprof::Scope ps(qexprof::K_VXC, st);
    hipLaunchKernelGGL(k_vxc, dim3((unsigned)nb), dim3(NT_VXC), vxc_lds(nao), st, (long)ngrid, nao, ao, weights, rho, wv,
                       nullptr, vpart, spart);
  }
  {
    qexprof::Scope ps(qexprof::K_OTHER, st);
    double* P = (double*)((char*)spart + al((size_t)nb * 16));
    vxc_reduce(nao, nb, vpart, spart, P, vmat, nelec_excsum, st);
    hipLaunchKernelGGL(k_copy1, dim3(1), dim3(64), 0, st, exc, nelec_excsum + 1);
  }
  return launch_status("qex3d_nr_rks_nn");
}

int qex3d_nr_rks_nn_local(int64_t ngrid, int32_t nao, const double* ao, const double* weights, const double* dm,
                          const double* nn_params, int32_t h1, int32_t h2, double scale, double* nelec_excsum,
                          double* vmat, void* workspace, size_t workspace_bytes, void* stream) {
  if (ngrid < 1 || nao < 1 || nao > QEX3D_MAX_NAO || !ao || !weights || !dm || !nn_params || !nelec_excsum || !vmat)
    return fail3(QEX_E_BADARG, "bad nr_rks_nn_local arguments");
  if (h1 < 1 || h2 < 1 || h1 > QEX3D_MAX_NN_LOCAL_HIDDEN || h2 > QEX3D_MAX_NN_LOCAL_HIDDEN)
    return fail3(QEX_E_BADARG, "hidden widths must be in [1, %d]", QEX3D_MAX_NN_LOCAL_HIDDEN);
  const size_t need = qex3d_nr_rks_workspace_bytes(ngrid, nao);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  hipStream_t st = (hipStream_t)stream;
  const int nb = vxc_blocks(ngrid);
  double* rho = (double*)workspace;
  double* vpart = (double*)((char*)workspace + al((size_t)ngrid * 8));
  double* spart = (double*)((char*)vpart + al((size_t)nb * nao * nao * 8));
  double* wv = (double*)((char*)spart + al((size_t)nb * 16) + al((size_t)nao * nao * 8));
  double* ex = (double*)((char*)wv + al((size_t)ngrid * 8));
  {
    qexprof::Scope ps(qexprof::K_RHO, st);
    rho_dispatch(ngrid, nao, ao, dm, rho, st);
  }
  {
    qexprof::Scope ps(qexprof::K_NN, st);
    const size_t lds = nn_local_lds(h1, h2);
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_nn_local, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_nn_local, dim3((unsigned)nn_local_blocks(ngrid)), dim3(NNL_NT), lds, st, (long)ngrid, h1, h2,
                       scale, nn_params, rho, weights, wv, ex);
  }
  {
    qexprof::Scope ps(qexprof::K_VXC, st);
    hipLaunchKernelGGL(k_vxc, dim3((unsigned)nb), dim3(NT_VXC), vxc_lds(nao), st, (long)ngrid, nao, ao, weights, rho,
                       wv, ex, vpart, spart);
  }
  {
    qexprof::Scope ps(qexprof::K_OTHER, st);
    double* P = (double*)((char*)spart + al((size_t)nb * 16));
    vxc_reduce(nao, nb, vpart, spart, P, vmat, nelec_excsum, st);
  }
  return launch_status("qex3d_nr_rks_nn_local");
}

size_t qex3d_eval_xc_nn_workspace_bytes(int64_t ngrid, int32_t h1, int32_t h2, int32_t is_global) {
  if (ngrid < 1 || h1 < 1 || h2 < 1) return 0;
  if (!is_global) return 256;
  if (!nn_widths_ok(h1, h2)) return 0;
  return al((size_t)gnn_blocks(ngrid) * h1 * 8) + 2 * al((size_t)h1 * 8);
}

int qex3d_eval_xc_nn(int32_t is_global, int64_t ngrid, const double* rho, const double* nn_params, int32_t h1,
                     int32_t h2, double scale, double* exc, double* vrho, void* workspace, size_t workspace_bytes,
                     void* stream) {
  if (ngrid < 1 || !rho || !nn_params || !exc || !vrho) return fail3(QEX_E_BADARG, "bad eval_xc_nn arguments");
  hipStream_t st = (hipStream_t)stream;
  if (!is_global) {
    if (h1 < 1 || h2 < 1 || h1 > QEX3D_MAX_NN_LOCAL_HIDDEN || h2 > QEX3D_MAX_NN_LOCAL_HIDDEN)
      return fail3(QEX_E_BADARG, "hidden widths must be in [1, %d]", QEX3D_MAX_NN_LOCAL_HIDDEN);
    qexprof::Scope ps(qexprof::K_NN, st);
    const size_t lds = nn_local_lds(h1, h2);
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_nn_local, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_nn_local, dim3((unsigned)nn_local_blocks(ngrid)), dim3(NNL_NT), lds, st, (long)ngrid, h1, h2,
                       scale, nn_params, rho, (const double*)nullptr, vrho, exc);
    return launch_status("qex3d_eval_xc_nn");
  }
  if (!nn_widths_ok(h1, h2)) return fail3(QEX_E_BADARG, "hidden widths must be in [1, %d]", QEX3D_MAX_NN_HIDDEN);
  const size_t need = qex3d_eval_xc_nn_workspace_bytes(ngrid, h1, h2, is_global);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  char* q = (char*)workspace;
  double* zpart = (double*)q;
  q += al((size_t)gnn_blocks(ngrid) * h1 * 8);
  double* z1 = (double*)q;
  q += al((size_t)h1 * 8);
  double* dz1 = (double*)q;
  qexprof::Scope ps(qexprof::K_NN, st);
  gnn_dispatch(ngrid, h1, h2, scale, nn_params, rho, nullptr, zpart, z1, dz1, exc, vrho, st);
  return launch_status("qex3d_eval_xc_nn");
}

size_t qex3d_gen_eigh_workspace_bytes(int32_t batch, int32_t n) {
  if (batch < 1 || n < 1 || n > QEX3D_MAX_NAO) return 0;
  return al((size_t)batch * 8 * n * n * 8);
}

int qex3d_gen_eigh(int32_t batch, int32_t n, const double* F, const double* S, double* mo_energy, double* mo_coeff,
                   void* workspace, size_t workspace_bytes, void* stream) {
  if (batch < 1 || n < 1 || n > QEX3D_MAX_NAO || !F || !S || !mo_energy || !mo_coeff)
    return fail3(QEX_E_BADARG, "bad gen_eigh arguments (n must be in [1, %d])", QEX3D_MAX_NAO);
  const size_t need = qex3d_gen_eigh_workspace_bytes(batch, n);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  const size_t lds = eigh_lds(n);
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_gen_eigh, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  qexprof::Scope ps(qexprof::K_EIGH, (hipStream_t)stream);
  hipLaunchKernelGGL(k_gen_eigh, dim3(batch), dim3(NT_EIG), lds, (hipStream_t)stream, n, F, S, mo_energy, mo_coeff,
                     (double*)workspace);
  return launch_status("qex3d_gen_eigh");
}

size_t qex3d_eigh_vjp_workspace_bytes(int32_t batch, int32_t n) {
  if (batch < 1 || n < 1 || n > QEX3D_MAX_NAO) return 0;
  return al((size_t)batch * 2 * n * n * 8);
}

int qex3d_eigh_vjp(int32_t batch, int32_t n, const double* w, const double* v, const double* gw, const double* gv,
                   int32_t exact, double* gA, void* workspace, size_t workspace_bytes, void* stream) {
  if (batch < 1 || n < 1 || n > QEX3D_MAX_NAO || !w || !v || !gA)
    return fail3(QEX_E_BADARG, "bad eigh_vjp arguments (n must be in [1, %d])", QEX3D_MAX_NAO);
  const size_t need = qex3d_eigh_vjp_workspace_bytes(batch, n);
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  hipLaunchKernelGGL(k_eigh_vjp, dim3(batch), dim3(NT_EIG), 0, (hipStream_t)stream, n, w, v, gw, gv, gA,
                     (double*)workspace, exact);
  return launch_status("qex3d_eigh_vjp");
}

// argument checks of one molecule; `pre` prefixes the message (the molecule's index in a batch)
static int scf_check(const qex3d_scf_problem* p, const qex3d_scf_inputs* in, const qex3d_scf_outputs* out,
                     const char* pre) {
  const int n = p->nao;
  if (n < 1 || n > QEX3D_MAX_NAO) return fail3(QEX_E_UNSUPPORTED, "%snao must be in [1, %d]", pre, QEX3D_MAX_NAO);
  if (p->nelectron < 2 || (p->nelectron & 1) || p->nelectron / 2 > n)
    return fail3(QEX_E_BADARG, "%sclosed-shell nelectron required (got %d)", pre, p->nelectron);
  if (!is_lda(p->method) && p->method != QEX3D_RHF && p->method != QEX3D_RKS_NN_GLOBAL &&
      p->method != QEX3D_RKS_NN_LOCAL)
    return fail3(QEX_E_BADARG, "%sbad method", pre);
  if (p->method == QEX3D_RKS_NN_LOCAL &&
      (!in->nn_params || p->nn_h1 < 1 || p->nn_h2 < 1 || p->nn_h1 > QEX3D_MAX_NN_LOCAL_HIDDEN ||
       p->nn_h2 > QEX3D_MAX_NN_LOCAL_HIDDEN))
    return fail3(QEX_E_BADARG, "%snetwork parameters and hidden widths in [1, %d] required", pre,
                 QEX3D_MAX_NN_LOCAL_HIDDEN);
  if (p->precision != QEX3D_PRECISION_FP64 && p->precision != QEX3D_PRECISION_FP32_GRID)
    return fail3(QEX_E_BADARG, "%sbad precision %d", pre, p->precision);
  const bool f32 = p->method != QEX3D_RHF && p->precision == QEX3D_PRECISION_FP32_GRID;
  if (p->method != QEX3D_RHF && (p->ngrid < 1 || !in->weights || (f32 ? !in->ao_f32 : !in->ao)))
    return fail3(QEX_E_BADARG, "%sRKS needs a grid and its AO values (%s)", pre, f32 ? "ao_f32" : "ao");
  if (p->method == QEX3D_RKS_NN_GLOBAL && (!in->nn_params || !nn_widths_ok(p->nn_h1, p->nn_h2)))
    return fail3(QEX_E_BADARG, "%snetwork parameters and hidden widths in [1, %d] required", pre,
                 QEX3D_MAX_NN_HIDDEN);
  if (p->max_cycle < 1 || p->diis_space < 0 || p->diis_space > QEX3D_MAX_DIIS)
    return fail3(QEX_E_BADARG, "%sbad cycle/diis", pre);
  if (!in->S || !in->H || !in->eri) return fail3(QEX_E_BADARG, "%smissing S, H or eri", pre);
  if (!out->e_tot || !out->mo_energy || !out->mo_coeff || !out->dm || !out->status)
    return fail3(QEX_E_BADARG, "%smissing output array", pre);
  return QEX_OK;
}

// The SCF of B molecules of any sizes in one sequence of launches: every stage is one launch over
// the batch, each molecule computes exactly what it computes alone (same per-molecule block
// counts and reduction orders), and a molecule that has converged or used its cycles is frozen.
static int scf_run(int B, const qex3d_scf_problem* ps, const qex3d_scf_inputs* ins, const qex3d_scf_outputs* outs,
                   void* workspace, size_t workspace_bytes, hipStream_t st, const char* name) {
  if (!ps || !ins || !outs) return fail3(QEX_E_BADARG, "null argument");
  char pre[32] = "";
  for (int b = 0; b < B; ++b) {
    if (B > 1) snprintf(pre, sizeof(pre), "molecule %d: ", b);
    const int rc = scf_check(ps + b, ins + b, outs + b, pre);
    if (rc) return rc;
  }
  const size_t need = B == 1 ? qex3d_scf_workspace_bytes(ps) : qex3d_scf_batch_workspace_bytes(B, ps);
  if (!need) return fail3(QEX_E_BADARG, "bad problem sizes");
  if (!workspace || workspace_bytes < need) return fail3(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  char* ws = (char*)workspace;
  Scf3* dss = (Scf3*)ws;
  int* active = (int*)(ws + al((size_t)B * sizeof(Scf3)));
  size_t off = scf_head_bytes(B);
  std::vector<Scf3> rec(B);
  int nmax = 1, max_cycle = 0, sync_every = 0, npair_max = 1, h1_max = 1, h2_max = 1;
  long rho_blk = 0, vxc_blk = 0, sum_blk = 0, sym_blk = 0, j_blk = 1, l1_blk = 0, red_blk = 0, vjp_blk = 0;
  bool any_grid = false, any_nn = false, any_nnl = false, any_x32 = false, any_x32_nn = false, any_rhf = false;
  long x32_blk[8] = {0, 0, 0, 0, 0, 0, 0, 0}, x32_sum_blk = 0;
  int x32_mask_lda = 0, x32_mask_nn = 0;  // bit NB: a molecule of x32_ld(n) = 16 NB
  size_t x32_lds_nb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long nnl_blk = 0;
  size_t lds_nnl = 0;
  for (int b = 0; b < B; ++b) {
    const qex3d_scf_problem* p = ps + b;
    const int n = p->nao;
    const long ng = scf_ngrid(p);
    const Lay3 L = layout3(n, ng, scf_space(p), scf_h1(p));
    char* w = ws + off;
    off += scf_mol_bytes(p);
    Scf3& s = rec[b];
    s = Scf3{};
    s.n = n;
    s.npair = n * (n + 1) / 2;
    s.nocc = p->nelectron / 2;
    s.max_cycle = p->max_cycle;
    s.diis_space = p->diis_space;
    s.method = p->method;
    s.ng = ng;
    s.nchunk = vxc_blocks(ng);
    s.conv_tol = p->conv_tol;
    s.enuc = p->energy_nuc;
    s.S = ins[b].S;
    s.H = ins[b].H;
    s.eri = ins[b].eri;
    s.ao = ins[b].ao;
    s.w = ins[b].weights;
    s.ao32 = ins[b].ao_f32;
    s.prec = p->method == QEX3D_RHF ? QEX3D_PRECISION_FP64 : p->precision;
    s.x32blk = x32_blocks(ng);
    s.Lc = (double*)(w + L.total);
    s.X = (double*)(w + L.total + al((size_t)n * n * 8));
    s.status_out = outs[b].status;
    s.F = (double*)(w + L.F);
    s.Fd = (double*)(w + L.Fd);
    s.D = outs[b].dm;
    s.Dlast = (double*)(w + L.Dlast);
    s.C = outs[b].mo_coeff;
    s.Cp = (double*)(w + L.Cp);
    s.mo = outs[b].mo_energy;
    s.J = (double*)(w + L.J);
    s.K = (double*)(w + L.K);
    s.Vxc = (double*)(w + L.Vxc);
    s.vpart = (double*)(w + L.vpart);
    s.spart = (double*)(w + L.spart);
    s.rho = (double*)(w + L.rho);
    s.tmp1 = (double*)(w + L.tmp1);
    s.tmp2 = (double*)(w + L.tmp2);
    s.diisF = (double*)(w + L.diisF);
    s.diisE = (double*)(w + L.diisE);
    s.diisB = (double*)(w + L.diisB);
    s.scal = (double*)(w + L.scal);
    s.ist = (int*)(w + L.ist);
    s.e_hist = outs[b].e_tot;
    s.nn = ins[b].nn_params;
    s.h1 = p->method == QEX3D_RKS_NN_LOCAL ? p->nn_h1 : scf_h1(p);
    s.h2 = p->nn_h2;
    s.nn_scale = p->nn_scale;
    s.zpart = (double*)(w + L.zpart);
    s.z1 = (double*)(w + L.z1);
    s.dz1 = (double*)(w + L.dz1);
    s.wv = (double*)(w + L.wv);
    s.ex = (double*)(w + L.ex);
    s.eigf = (double*)(w + L.eigf);
    s.jpart = (double*)(w + L.jpart);
    nmax = n > nmax ? n : nmax;
    max_cycle = p->max_cycle > max_cycle ? p->max_cycle : max_cycle;
    if (p->sync_every > 0 && (sync_every == 0 || p->sync_every < sync_every)) sync_every = p->sync_every;
    npair_max = s.npair > npair_max ? s.npair : npair_max;
    j_blk = coulomb_blocks(s.npair) > j_blk ? coulomb_blocks(s.npair) : j_blk;
    if (p->method == QEX3D_RHF) any_rhf = true;
    if (s.prec == QEX3D_PRECISION_FP32_GRID) {
      any_x32 = true;
      if (!is_lda(p->method)) any_x32_nn = true;
      const int nbk = x32_ld(n) / 16;
      (is_lda(p->method) ? x32_mask_lda : x32_mask_nn) |= 1 << nbk;
      x32_blk[nbk] = s.x32blk > x32_blk[nbk] ? s.x32blk : x32_blk[nbk];
      x32_lds_nb[nbk] = x32_lds(n, s.nocc) > x32_lds_nb[nbk] ? x32_lds(n, s.nocc) : x32_lds_nb[nbk];
      x32_sum_blk = vxc_sum_blocks(x32_tiles(n) * 256) > x32_sum_blk ? vxc_sum_blocks(x32_tiles(n) * 256) : x32_sum_blk;
    } else if (p->method != QEX3D_RHF) {
      any_grid = true;
      rho_blk = (ng + 63) / 64 > rho_blk ? (ng + 63) / 64 : rho_blk;
      vxc_blk = s.nchunk > vxc_blk ? s.nchunk : vxc_blk;
      sum_blk = vxc_sum_blocks(n * n) > sum_blk ? vxc_sum_blocks(n * n) : sum_blk;
      sym_blk = (n * n + 255) / 256 > sym_blk ? (n * n + 255) / 256 : sym_blk;
    }
    if (p->method == QEX3D_RKS_NN_LOCAL) {
      any_nnl = true;
      nnl_blk = nn_local_blocks(ng) > nnl_blk ? nn_local_blocks(ng) : nnl_blk;
      lds_nnl = nn_local_lds(p->nn_h1, p->nn_h2) > lds_nnl ? nn_local_lds(p->nn_h1, p->nn_h2) : lds_nnl;
    }
    if (p->method == QEX3D_RKS_NN_GLOBAL) {
      any_nn = true;
      l1_blk = gnn_blocks(ng) > l1_blk ? gnn_blocks(ng) : l1_blk;
      red_blk = (s.h1 + 63) / 64 > red_blk ? (s.h1 + 63) / 64 : red_blk;
      vjp_blk = gnn_vjp_blocks(ng) > vjp_blk ? gnn_vjp_blocks(ng) : vjp_blk;
      h1_max = s.h1 > h1_max ? s.h1 : h1_max;
      h2_max = s.h2 > h2_max ? s.h2 : h2_max;
    }
  }
  if (off > need) return fail3(QEX_E_WORKSPACE, "internal layout exceeds workspace");
  const size_t lds_eigh = eigh_lds(nmax);
  if (lds_eigh + sizeof(qexeig::Smem) > 160 * 1024) return fail3(QEX_E_UNSUPPORTED, "nao too large for the LDS eigensolver");
  if (lds_eigh > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_scf_eigh, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_eigh);
  if (lds_nnl > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_nn_local_b, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_nnl);
  for (int b0 = 0; b0 < B; b0 += PUT_N) {
    ScfPut pk{};
    const int cnt = B - b0 < PUT_N ? B - b0 : PUT_N;
    for (int i = 0; i < cnt; ++i) pk.r[i] = rec[b0 + i];
    hipLaunchKernelGGL(k_put_scf, dim3(1), dim3(256), 0, st, pk, cnt, dss + b0);
  }
  const unsigned nb = (unsigned)B;
  auto veff = [&]() {
    if (any_x32) {
      // FP32 grid molecules: LDA in one pass, network methods' density first
      qexprof::Scope ps_(any_x32_nn ? qexprof::K_RHO : qexprof::K_VXC, st);
      x32_launch<X32_FUSED>(x32_mask_lda, x32_blk, x32_lds_nb, nb, st, dss);
      x32_launch<X32_RHO>(x32_mask_nn, x32_blk, x32_lds_nb, nb, st, dss);
    }
    if (any_grid) {
      qexprof::Scope ps_(qexprof::K_RHO, st);
      hipLaunchKernelGGL(k_rho_b, dim3((unsigned)rho_blk, nb), dim3(NT), rho_lds(nmax), st, dss);
    }
    if (any_grid || any_x32) {
      if (any_nnl) {
        qexprof::Scope ps_(qexprof::K_NN, st);
        hipLaunchKernelGGL(k_nn_local_b, dim3((unsigned)nnl_blk, nb), dim3(NNL_NT), lds_nnl, st, dss);
      }
      if (any_nn) {
        qexprof::Scope ps_(qexprof::K_NN, st);
        hipLaunchKernelGGL(k_gnn_l1_b, dim3((unsigned)l1_blk, nb), dim3(NT), 0, st, dss);
        hipLaunchKernelGGL(k_gnn_red_b, dim3((unsigned)red_blk, nb), dim3(1024), 0, st, dss);
        hipLaunchKernelGGL(k_gnn_mid_b, dim3(1, nb), dim3(1024), (size_t)(h1_max + 2 * h2_max) * 8, st, dss);
        hipLaunchKernelGGL(k_gnn_vjp_b, dim3((unsigned)vjp_blk, nb), dim3(NT), (size_t)h1_max * 8, st, dss);
      }
      if (any_grid) {
        {
          qexprof::Scope ps_(qexprof::K_VXC, st);
          hipLaunchKernelGGL(k_vxc_b, dim3((unsigned)vxc_blk, nb), dim3(NT_VXC), vxc_lds(nmax), st, dss);
        }
        // Vxc = P + P^T with P the fixed-order sum of the chunk partials; scal[2..3] = nelec, exc
        qexprof::Scope ps_(qexprof::K_OTHER, st);
        hipLaunchKernelGGL(k_vxc_sum_b, dim3((unsigned)sum_blk, nb), dim3(256), 0, st, dss);
        hipLaunchKernelGGL(k_sym_add_b, dim3((unsigned)sym_blk, nb), dim3(256), 0, st, dss);
      }
      if (any_x32) {
        if (any_x32_nn) {
          qexprof::Scope ps_(qexprof::K_VXC, st);
          x32_launch<X32_V>(x32_mask_nn, x32_blk, x32_lds_nb, nb, st, dss);
        }
        qexprof::Scope ps_(qexprof::K_OTHER, st);
        hipLaunchKernelGGL(k_x32_sum_b, dim3((unsigned)x32_sum_blk, nb), dim3(256), 0, st, dss);
      }
    }
    qexprof::Scope ps_(qexprof::K_COULOMB, st);
    if (any_rhf) hipLaunchKernelGGL(k_coulomb_b, dim3((unsigned)j_blk, nb), dim3(NT), (size_t)npair_max * 8, st, dss);
    if (any_grid || any_x32) {
      hipLaunchKernelGGL(k_coulomb_tiles_b, dim3((unsigned)jtiles(npair_max), nb), dim3(256), 0, st, dss);
      hipLaunchKernelGGL(k_coulomb_jsum_b, dim3((unsigned)((npair_max + 255) / 256), nb), dim3(256), 0, st, dss);
    }
  };
  auto per_mol = [&](int kind, auto kernel, auto... args) {
    qexprof::Scope ps_(kind, st);
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(NT), 0, st, dss, args...);
  };
  auto per_mol_wide = [&](int kind, auto kernel, auto... args) {
    qexprof::Scope ps_(kind, st);
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(NT_SCF), 0, st, dss, args...);
  };
  auto eigh = [&](int full) {
    qexprof::Scope ps_(qexprof::K_EIGH, st);
    hipLaunchKernelGGL(k_scf_eigh, dim3(nb), dim3(NT_EIG), lds_eigh, st, dss, full);
  };
  // core-Hamiltonian guess: F = H (k_scf_init zeroes J, K, Vxc and D with a kernel, not a memset)
  per_mol(qexprof::K_OTHER, k_scf_init);
  {
    qexprof::Scope ps_(qexprof::K_OTHER, st);
    const size_t lds_o = (size_t)2 * nmax * nmax * 8;
    if (lds_o > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_scf_orthogonaliser, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_o);
    hipLaunchKernelGGL(k_scf_orthogonaliser, dim3(nb), dim3(NT_SCF), lds_o, st, dss);
  }
  per_mol_wide(qexprof::K_FOCK, k_scf_fock, -1);
  eigh(0);
  veff();
  per_mol_wide(qexprof::K_ENERGY, k_scf_energy, -1);
  for (int c = 0; c < max_cycle; ++c) {
    per_mol_wide(qexprof::K_FOCK, k_scf_fock, c);
    eigh(0);
    veff();
    per_mol_wide(qexprof::K_ENERGY, k_scf_energy, c);
    if (sync_every > 0 && (c + 1) % sync_every == 0 && c + 1 < max_cycle) {
      // optional host check: stop enqueueing cycles once every molecule is frozen
      int32_t left = 0;
      hipLaunchKernelGGL(k_scf_active, dim3(1), dim3(64), 0, st, dss, B, active);
      if (hipMemcpyAsync(&left, active, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return fail3(QEX_E_LAUNCH, "convergence check failed");
      if (left == 0) break;
    }
  }
  eigh(1);  // all orbitals and orbital energies of the final Fock matrix
  per_mol(qexprof::K_OTHER, k_scf_finish);
  return launch_status(name);
}

int qex3d_scf(const qex3d_scf_problem* p, const qex3d_scf_inputs* in, const qex3d_scf_outputs* out, void* workspace,
              size_t workspace_bytes, void* stream) {
  return scf_run(1, p, in, out, workspace, workspace_bytes, (hipStream_t)stream, "qex3d_scf");
}

int qex3d_scf_batch(int32_t batch, const qex3d_scf_problem* problems, const qex3d_scf_inputs* inputs,
                    const qex3d_scf_outputs* outputs, void* workspace, size_t workspace_bytes, void* stream) {
  if (batch < 1 || batch > QEX3D_MAX_SCF_BATCH)
    return fail3(QEX_E_BADARG, "batch must be in [1, %d]", QEX3D_MAX_SCF_BATCH);
  return scf_run(batch, problems, inputs, outputs, workspace, workspace_bytes, (hipStream_t)stream,
                 "qex3d_scf_batch");
}

}  // extern "C"
