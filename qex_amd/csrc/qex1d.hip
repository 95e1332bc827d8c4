// qex1d.hip -- batched 1D Kohn-Sham SCF for gfx950 (MI355X).
//
// One SCF iteration over a batch of B systems on a shared uniform grid of N points is four
// launches (plus one MLP launch for the neural XC kind):
//   k_mlp        (LOCAL_MLP only) per-point MLP forward + input-VJP at the input density, and
//                forward at the previous iteration's output density (for its XC energy).
//   k_potential  grid (tiles, B): Hartree potential as a Toeplitz matvec from an LDS table,
//                XC potential, v_KS; and the deferred Hartree / XC energies of the previous
//                iteration's output density (second matvec in the same pass).
//   k_eigen      grid (K, B): one workgroup per wanted eigenpair.  The KS Hamiltonian
//                -1/2 L/dx^2 + diag(v) is pentadiagonal (5-point stencil), so eigenvalues come
//                from Sturm counts of its LDL^T factorisation (256 shifts per round, one per
//                lane) and eigenvectors from inverse iteration with a pivoted banded LU.
//   k_density    grid (B): Gram-Schmidt of the occupied orbitals, density, eigen energy, gap,
//                reflection, density mixing, convergence flags.
// k_finalize assembles the total energies after the loop.  Reductions run in a fixed order
// (wave butterflies + ordered cross-wave sums), so results do not depend on the batch.
#include <hip/hip_runtime.h>
#include <mutex>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qex.h"
#include "prof.h"
#include "band.h"

namespace qexprof {
bool g_on = false;
unsigned g_mask = 0xffffffffu;
Rec* g_rec = nullptr;
int g_n = 0, g_cap = 0;
hipEvent_t* g_pool = nullptr;
int g_pool_n = 0, g_pool_used = 0;
}  // namespace qexprof

namespace {

using qexband::ldlt_backward;
using qexband::ldlt_factor_forward;
using qexband::ldlt_forward;
using qexband::rcp64;
using qexband::sturm_count;
using qexband::Side;
using qexband::SideZ;
using qexband::sturm_side;
using qexband::sturm_join;
using qexband::middle_solve;
using qexband::ldlt_side_factor_forward;
using qexband::ldlt_side_forward;
using qexband::ldlt_side_backward;

constexpr int NT = 256;     // threads per workgroup (4 wave64)
constexpr int NWAVE = NT / 64;
constexpr int TILE = 64;    // output points per k_potential workgroup
constexpr double EPS = 2.220446049250313e-16;

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(QEX_E_LAUNCH, "%s: %s", what, hipGetErrorString(e));
  return QEX_OK;
}

using qexprof::Scope;

// --------------------------------------------------------------------------------------------
// device helpers
// --------------------------------------------------------------------------------------------
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

// Sum over the 64 lanes (all active), the same value in every lane: DPP moves within each row of
// 16 lanes (pairs, quads, half-row mirror, row mirror), then the four row sums in a fixed order.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

// Sums M values over the workgroup; every thread receives the totals.  Fixed order.
template <int M>
__device__ __forceinline__ void block_sum(double (&v)[M], double* red) {
#pragma unroll
  for (int m = 0; m < M; ++m) v[m] = wave_sum(v[m]);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
#pragma unroll
    for (int m = 0; m < M; ++m) red[w * M + m] = v[m];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NWAVE; ++q) s += red[q * M + m];
    v[m] = s;
  }
}

__device__ __forceinline__ double block_max(double x, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = x;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int q = 1; q < NWAVE; ++q) s = fmax(s, red[q]);
  return s;
}

// ---- 1D LDA for the exponential interaction (jax_dft xc; Baker et al. PRB 91, 235141) -------
struct XcOut {
  double e, v;  // energy per electron and v = e + n de/dn
};

__device__ XcOut lda_exp(double n) {
  const double A = 1.071295, kappa = 1.0 / 2.385345;
  const double y = M_PI * n / kappa;
  const double dydn = M_PI / kappa;
  // exchange
  const double cx = A / (2.0 * M_PI);
  double ex, dex;
  if (y > 1e-15) {
    const double l = log(1.0 + y * y);  // as in the reference (not log1p)
    ex = cx * (l / y - 2.0 * atan(y));
    dex = cx * (-l / (y * y));
  } else {
    ex = cx * (-y + y * y * y / 6.0);
    dex = cx * (-1.0 + 0.5 * y * y);
  }
  // correlation
  const double al = 2.0, be = -1.00077, ga = 6.26099, de = -11.9041, et = 9.62614, si = -1.48334, nu = 1.0;
  double ec, dec;
  if (y == 0.0) {
    ec = 0.0;
    dec = -A / M_PI / al;
  } else {
    const double s = sqrt(y);
    const double c3 = nu * M_PI * kappa * kappa / A;
    const double P = al + be * s + ga * y + de * y * s + et * y * y + si * y * y * s + c3 * y * y * y;
    const double dP = 0.5 * be / s + ga + 1.5 * de * s + 2.0 * et * y + 2.5 * si * y * s + 3.0 * c3 * y * y;
    ec = -A * y / M_PI / P;
    dec = -A / M_PI * (P - y * dP) / (P * P);
  }
  XcOut o;
  o.e = ex + ec;
  o.v = o.e + n * (dex + dec) * dydn;
  return o;
}

// ---- activations --------------------------------------------------------------------------
__device__ __forceinline__ double act_fwd(int kind, double z, double* dz) {
  switch (kind) {
    case QEX_ACT_SILU: {
      const double s = 1.0 / (1.0 + exp(-z));
      *dz = s * (1.0 + z * (1.0 - s));
      return z * s;
    }
    case QEX_ACT_TANH: {
      const double t = tanh(z);
      *dz = 1.0 - t * t;
      return t;
    }
    case QEX_ACT_SOFTPLUS: {
      *dz = 1.0 / (1.0 + exp(-z));
      return z > 0 ? z + log1p(exp(-z)) : log1p(exp(z));
    }
    case QEX_ACT_SIGMOID: {
      const double s = 1.0 / (1.0 + exp(-z));
      *dz = s * (1.0 - s);
      return s;
    }
    case QEX_ACT_RELU:
      *dz = z > 0 ? 1.0 : 0.0;
      return z > 0 ? z : 0.0;
    default: {  // GELU, tanh approximation
      const double k = 0.7978845608028654;
      const double u = k * (z + 0.044715 * z * z * z);
      const double t = tanh(u);
      *dz = 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * k * (1.0 + 3.0 * 0.044715 * z * z);
      return 0.5 * z * (1.0 + t);
    }
  }
}

// --------------------------------------------------------------------------------------------
// Kernel arguments
// --------------------------------------------------------------------------------------------
struct Dev1 {
  int B, N, T, K, M, W, nTiles, loop, reflect, xc, nlay, act, mlpout;
  int widths[QEX1D_MAX_MLP_LAYERS + 1];
  int woff[QEX1D_MAX_MLP_LAYERS];  // offset of W_l in the packed parameter array
  int boff[QEX1D_MAX_MLP_LAYERS];  // offset of b_l
  int wmax;                        // widest layer input (window or hidden)
  int hmax;                        // widest hidden layer
  double dx, tol, xc_c, xc_p, xc_eps, norm;
  double Td, Tb, Tc;  // kinetic stencil: diagonal, first and second off-diagonal
  // inputs
  const double* table;
  const double* vext;
  const double* n0;
  const int* ne;
  const int* win;
  const double* mlp;
  // outputs
  double* dens;
  double* etot;
  double* vh;
  double* vxc;
  double* exc;
  double* gap;
  uint8_t* conv;
  double* eigs_out;
  double* orb_out;  // optional [B][T][K][N]
  // workspace
  double* o_raw;     // [B][N] output density of the last computed iteration
  double* o_xc;      // [B][N] density the deferred XC energy is evaluated at
  double* vks;       // [B][N]
  double* hist;      // [B][M][N] density differences (ring)
  double* vec;       // [B][K][N] eigenvectors (2-norm 1)
  double* eig;       // [B][K]
  double* eig_prev;  // [B][K]
  double* eig_delta; // [B][K]
  double* terms;     // [B][T][4]  E_eig, E_ks, E_ext
  double* partH;     // [B][T][nTiles]
  double* partX;     // [B][T][nTiles]
  double* partF;     // [B][T][nTiles]  sum f(n_in) for GLOBAL xc_energy_density
  int* src;          // [B][T]
  int* st;           // [B][4]  frozen, pending, warm_valid, -
  double* mlp_e;     // [B][N]
  double* mlp_G;     // [B][N][W]
  double* mlp_eout;  // [B][N]
  // slot mode (continuous batching, qex1d_slots_step): every system (slot) runs its own iteration
  // kslot[s] (0..T-1 SCF iterations, T: the pass that only evaluates the last energies; negative or
  // > T: idle); null in batch mode, where every system runs the launch's iteration
  int* kslot;
  const int* admit;  // [B] 1: a new request was placed in the slot before this step
  double alpha0, decay;
};

// iteration of system s in this launch: the launch's, or the slot's own in slot mode
__device__ __forceinline__ int iter_of(const Dev1& a, int s, int k_launch) { return a.kslot ? a.kslot[s] : k_launch; }

__device__ __host__ __forceinline__ bool xc_is_local(int xc) {
  return xc != QEX1D_XC_POWER_GLOBAL && xc != QEX1D_XC_GLOBAL_MLP && xc != QEX1D_XC_EXTERNAL_GLOBAL;
}
__device__ __host__ __forceinline__ bool xc_is_external(int xc) {
  return xc == QEX1D_XC_EXTERNAL_LOCAL || xc == QEX1D_XC_EXTERNAL_GLOBAL;
}

// transform of the network output (QEX1D_MLP_OUTPUT_*) and its derivative
__device__ __forceinline__ double out_transform(int kind, double y, double* d) {
  if (kind == QEX1D_MLP_OUTPUT_NEGATIVITY) {
    const double s = 1.0 / (1.0 + exp(-y));
    *d = -(s * (1.0 + y * (1.0 - s)));
    return -y * s;
  }
  *d = 1.0;
  return y;
}

// pointwise XC at density x: energy density (local: e; global: f), potential contribution
__device__ __forceinline__ void xc_point(const Dev1& a, double x, double* e, double* v) {
  switch (a.xc) {
    case QEX1D_XC_LDA_EXP: {
      XcOut o = lda_exp(x);
      *e = o.e;
      *v = o.v;
      return;
    }
    case QEX1D_XC_POWER_LOCAL: {
      const double t = x + a.xc_eps;
      const double pw = pow(t, a.xc_p);
      *e = a.xc_c * pw;
      *v = a.xc_c * pw + x * a.xc_c * a.xc_p * pow(t, a.xc_p - 1.0);
      return;
    }
    case QEX1D_XC_POWER_GLOBAL: {
      const double t = x + a.xc_eps;
      *e = a.xc_c * pow(t, a.xc_p);
      *v = a.dx * a.xc_c * a.xc_p * pow(t, a.xc_p - 1.0);
      return;
    }
    default:
      *e = 0.0;
      *v = 0.0;
      return;
  }
}

// flip-and-average map inside the system's reflection window
__device__ __forceinline__ int mirror(int i, int lo, int hi) { return (i >= lo && i <= hi) ? lo + hi - i : i; }

__device__ __forceinline__ const double* input_density(const Dev1& a, int s, int k) {
  return k == 0 ? a.n0 + (size_t)s * a.N : a.dens + ((size_t)s * a.T + (k - 1)) * a.N;
}

// --------------------------------------------------------------------------------------------
// k_init: reset per-system state (a kernel rather than a memset, so it can sit in a graph)
// --------------------------------------------------------------------------------------------
__global__ void k_init(Dev1 a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.B) return;
  a.st[4 * s + 0] = 0;
  a.st[4 * s + 1] = -1;
  a.st[4 * s + 2] = 0;
  a.st[4 * s + 3] = 0;
}

// --------------------------------------------------------------------------------------------
// k_mlp: windowed per-point MLP.  Workgroup = NT / LPP points x LPP lanes (LPP = 8, or 64 for
// networks too wide for 32 points' activations in LDS); lane q of a point owns units q, q+LPP, ...
// of every layer.  Pass 1 (input density, unless frozen): e and the input-VJP
// G[i][o] = m_i de_i/dx_i[o].  Pass 2 (pending output density): e only.
// --------------------------------------------------------------------------------------------
template <int LPP>
__device__ void mlp_point_pass(const Dev1& a, const double* __restrict__ dens_lds, int lo, int hi, int i, int q,
                               int pl, double* actA, double* actB, double* dact, bool backward,
                               double* e_out, double* G_out) {
  // dens_lds holds the (already reflected when needed) density on the whole grid
  constexpr int P = NT / LPP;
  const int N = a.N, W = a.W, h = W / 2, stride = a.wmax + 1;
  const bool valid = i < N;
  double* in = actA + pl * stride;
  // window input / norm
  for (int o = q; o < W; o += LPP) {
    const int j = i + o - h;
    in[o] = (valid && j >= 0 && j < N) ? dens_lds[j] / a.norm : 0.0;
  }
  __syncthreads();
  double* cur = actA;
  double* nxt = actB;
  const int L = a.nlay;
  // forward through hidden layers
  for (int l = 0; l < L - 1; ++l) {
    const int fi = a.widths[l], fo = a.widths[l + 1];
    const double* Wl = a.mlp + a.woff[l];
    const double* bl = a.mlp + a.boff[l];
    const double* x = cur + pl * stride;
    for (int u = q; u < fo; u += LPP) {
      double z = bl[u];
      for (int kk = 0; kk < fi; ++kk) z = fma(x[kk], Wl[kk * fo + u], z);
      double dz;
      const double y = act_fwd(a.act, z, &dz);
      nxt[pl * stride + u] = y;
      if (backward) dact[((size_t)l * P + pl) * stride + u] = dz;
    }
    __syncthreads();
    double* t = cur;
    cur = nxt;
    nxt = t;
  }
  // output layer (fan_out 1): lane-partial dot products reduced over the 8 lanes of the point
  const int fi = a.widths[L - 1];
  const double* WL = a.mlp + a.woff[L - 1];
  double part = 0.0;
  for (int kk = q; kk < fi; kk += LPP) part = fma(cur[pl * stride + kk], WL[kk], part);
  // LPP-lane reduction (lanes of a point are contiguous: pl*LPP + q)
#pragma unroll
  for (int o = 1; o < LPP; o <<= 1) part += __shfl_xor(part, o, 64);
  double de;
  const double e = out_transform(a.mlpout, part + a.mlp[a.boff[L - 1]], &de);
  if (q == 0 && valid) *e_out = e;
  if (!backward) {
    __syncthreads();
    return;
  }
  // backward with cotangent m_i on e_i (through the output transform)
  const double m = valid ? dens_lds[i] * de : 0.0;
  // gradient w.r.t. last hidden activations
  double* g = nxt;  // free buffer
  for (int kk = q; kk < fi; kk += LPP)
    g[pl * stride + kk] = m * WL[kk] * dact[((size_t)(L - 2) * P + pl) * stride + kk];
  __syncthreads();
  double* gcur = g;
  double* gnxt = cur;
  for (int l = L - 2; l >= 1; --l) {
    // layer l maps widths[l] -> widths[l+1]; gradient w.r.t. its input (widths[l] units)
    const int fi2 = a.widths[l], fo2 = a.widths[l + 1];
    const double* Wl = a.mlp + a.woff[l];
    const double* gz = gcur + pl * stride;
    for (int kk = q; kk < fi2; kk += LPP) {
      double acc = 0.0;
      for (int u = 0; u < fo2; ++u) acc = fma(gz[u], Wl[kk * fo2 + u], acc);
      gnxt[pl * stride + kk] = acc * dact[((size_t)(l - 1) * P + pl) * stride + kk];
    }
    __syncthreads();
    double* t = gcur;
    gcur = gnxt;
    gnxt = t;
  }
  // first layer: gradient w.r.t. the window inputs
  {
    const int fo1 = a.widths[1];
    const double* W0 = a.mlp + a.woff[0];
    const double* gz = gcur + pl * stride;
    for (int o = q; o < W; o += LPP) {
      double acc = 0.0;
      for (int u = 0; u < fo1; ++u) acc = fma(gz[u], W0[o * fo1 + u], acc);
      if (valid) G_out[o] = acc;
    }
  }
  __syncthreads();
}

template <int LPP>
__global__ __launch_bounds__(NT) void k_mlp(Dev1 a, int k_launch) {
  constexpr int P = NT / LPP;
  extern __shared__ double smem[];
  const int s = blockIdx.y;
  const int k = iter_of(a, s, k_launch);
  const int frozen = a.st[4 * s + 0];
  const int pending = a.st[4 * s + 1];
  const bool doPot = k >= 0 && k < a.T && !frozen;
  const bool doE = pending >= 0 && pending < a.T;
  if (!doPot && !doE) return;
  const int N = a.N, stride = a.wmax + 1;
  double* dens = smem;                      // N
  double* actA = dens + N;              // P * stride
  double* actB = actA + P * stride;     // P * stride
  double* dact = actB + P * stride;     // (L-1) * P * stride
  const int pl = threadIdx.x / LPP, q = threadIdx.x % LPP;
  const int i = blockIdx.x * P + pl;
  int lo = 0, hi = N - 1;
  const bool sym = a.reflect && xc_is_local(a.xc);
  if (a.reflect) {
    lo = a.win[2 * s];
    hi = a.win[2 * s + 1];
  }
  if (doPot) {
    const double* nin = input_density(a, s, k);
    for (int t = threadIdx.x; t < N; t += NT) {
      double v = nin[t];
      if (sym) v = 0.5 * (v + nin[mirror(t, lo, hi)]);
      dens[t] = v;
    }
    __syncthreads();
    double eo;
    mlp_point_pass<LPP>(a, dens, lo, hi, i, q, pl, actA, actB, dact, true, &eo,
                   a.mlp_G + ((size_t)s * N + (i < N ? i : 0)) * a.W);
    if (q == 0 && i < N) a.mlp_e[(size_t)s * N + i] = eo;
    __syncthreads();
  }
  if (doE) {
    const double* ox = a.o_xc + (size_t)s * N;
    for (int t = threadIdx.x; t < N; t += NT) dens[t] = ox[t];
    __syncthreads();
    double eo;
    mlp_point_pass<LPP>(a, dens, lo, hi, i, q, pl, actA, actB, dact, false, &eo, nullptr);
    if (q == 0 && i < N) a.mlp_eout[(size_t)s * N + i] = eo;
  }
}

// --------------------------------------------------------------------------------------------
// k_mlp_mfma: the same windowed MLP on FP64 matrix cores (v_mfma_f64_16x16x4f64) for hidden
// widths <= 64 and <= 4 hidden layers.  A workgroup takes MP = 48 consecutive points of the
// flattened [B][N] grid; each layer is a 48 x H tile GEMM (3 row blocks), wave w owning output
// columns 16w..16w+15, written back in place into the one activation tile once every wave has
// read it (25 KB of LDS, so three workgroups share a CU: the 64 x 513 points of a c2 batch are
// 684 tiles, one round on 256 CUs with at most 144 points per CU).  Activation derivatives stay
// in registers for the backward GEMMs (W^T).
// MFMA operand layout (16x16x4, one wave): lane l supplies A[l%16][l/16] and B[l/16][l%16];
// it receives D[4v + l/16][l%16] in element v.
// --------------------------------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));
constexpr int MP = 48;         // points per workgroup (3 MFMA row blocks)
constexpr int MLP_RB = MP / 16;  // row blocks of 16 points
constexpr int MLP_HMAX = 64;   // widest hidden layer on the MFMA path
constexpr int MLP_LH = 4;      // hidden layers on the MFMA path
constexpr int MLP_S = 65;      // LDS row stride (doubles) of the activation tiles

__device__ __forceinline__ d4_t mfma64(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int rup(int x, int m) { return (x + m - 1) / m * m; }

// acc[rb] (rows 16rb.., columns of this wave's block cb) = in[MP x K] . B[K x ncols], where
// B[kk][col] = W[kk * ldw + col] (or W[col * ldw + kk] when trans); entries outside
// kk < K, col < ncols are zero.  `in` is an LDS tile with row stride MLP_S, zero beyond K.
__device__ __forceinline__ void tile_gemm(const double* in, int K, const double* W, int ldw, int ncols, bool trans,
                                         int cb, d4_t (&acc)[MLP_RB]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int rb = 0; rb < MLP_RB; ++rb) acc[rb] = (d4_t){0.0, 0.0, 0.0, 0.0};
  const int col = cb * 16 + (l & 15);
  const int kp = rup(K, 4);
  for (int k0 = 0; k0 < kp; k0 += 4) {
    const int kk = k0 + (l >> 4);
    double bv = 0.0;
    if (kk < K && col < ncols) bv = trans ? W[(size_t)col * ldw + kk] : W[(size_t)kk * ldw + col];
#pragma unroll
    for (int rb = 0; rb < MLP_RB; ++rb) {
      const double av = in[(rb * 16 + (l & 15)) * MLP_S + kk];
      acc[rb] = mfma64(av, bv, acc[rb]);
    }
  }
}

// One hidden layer of the tile MLP: nxt = act(cur W_lh + b_lh); keeps act' in dzl.
__device__ __forceinline__ void mlp_hidden_fwd(const Dev1& a, int lh, const double* cur, double* nxt,
                                               double (&dzl)[MLP_RB][4], int wv, int l) {
  const int fi = a.widths[lh], fo = a.widths[lh + 1];
  const double* Wl = a.mlp + a.woff[lh];
  const double* bl = a.mlp + a.boff[lh];
  d4_t acc[MLP_RB];
  tile_gemm(cur, fi, Wl, fo, fo, false, wv, acc);
  const int col = wv * 16 + (l & 15);
  const double bias = col < fo ? bl[col] : 0.0;
  __syncthreads();  // every wave has read cur (nxt may be the same tile)
#pragma unroll
  for (int rb = 0; rb < MLP_RB; ++rb) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rb * 16 + 4 * v + (l >> 4);
      double d;
      const double y = act_fwd(a.act, acc[rb][v] + bias, &d);
      nxt[row * MLP_S + col] = col < fo ? y : 0.0;
      dzl[rb][v] = col < fo ? d : 0.0;
    }
  }
  __syncthreads();
}

// Backward through hidden layer lh: nxt = (cur W_lh^T) * act'(z_{lh-1}) (dzp holds act' of lh-1).
__device__ __forceinline__ void mlp_hidden_bwd(const Dev1& a, int lh, const double* cur, double* nxt,
                                               const double (&dzp)[MLP_RB][4], int wv, int l) {
  const int fi = a.widths[lh], fo = a.widths[lh + 1];
  const double* Wl = a.mlp + a.woff[lh];
  d4_t acc[MLP_RB];
  tile_gemm(cur, fo, Wl, fo, fi, true, wv, acc);
  const int col = wv * 16 + (l & 15);
  __syncthreads();  // every wave has read cur (nxt may be the same tile)
#pragma unroll
  for (int rb = 0; rb < MLP_RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rb * 16 + 4 * v + (l >> 4);
      nxt[row * MLP_S + col] = col < fi ? acc[rb][v] * dzp[rb][v] : 0.0;
    }
  __syncthreads();
}

// Gradient at the last hidden layer's pre-activations: cot_row * w_out[col] * act'.
__device__ __forceinline__ void mlp_out_grad(double* nxt, const double* cot, double wo, const double (&dzl)[MLP_RB][4],
                                             int col, int l) {
#pragma unroll
  for (int rb = 0; rb < MLP_RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rb * 16 + 4 * v + (l >> 4);
      nxt[row * MLP_S + col] = cot[row] * wo * dzl[rb][v];
    }
}

template <int NH>
__global__ __launch_bounds__(NT, NH <= 2 ? 3 : 2) void k_mlp_mfma(Dev1 a, int k_launch) {
  extern __shared__ double smem[];
  // one activation tile: each layer's GEMM keeps its result in registers until every wave has
  // read the tile, then writes it back in place (34 KB of LDS: three workgroups per CU)
  double* bufB = smem;                 // MP x MLP_S
  double* cot = bufB + MP * MLP_S;     // MP cotangents
  __shared__ int sh_any[2];
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
  const int N = a.N, W = a.W, h = W / 2, L = a.nlay;
  const long total = (long)a.B * N;
  const long g0 = (long)blockIdx.x * MP;
  const bool sym = a.reflect && xc_is_local(a.xc);
  if (tid < 2) sh_any[tid] = 0;
  __syncthreads();
  if (tid < MP) {
    const long g = g0 + tid;
    if (g < total) {
      const int s = (int)(g / N);
      const int k = iter_of(a, s, k_launch);
      if (k >= 0 && k < a.T && !a.st[4 * s + 0]) sh_any[0] = 1;
      if (a.st[4 * s + 1] >= 0 && a.st[4 * s + 1] < a.T) sh_any[1] = 1;
    }
  }
  __syncthreads();
  const bool any_pot = sh_any[0] != 0, any_e = sh_any[1] != 0;
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 0 && !any_pot) continue;
    if (pass == 1 && !any_e) continue;
    const bool backward = pass == 0;
    // ---- window inputs (x / norm) into bufB, zero padded to a multiple of 4 columns ----
    // 4 threads per point: the system index and status are read once per point, and a thread's
    // window loads are independent, so they are in flight together
    static_assert(MP * 4 <= NT, "4 threads per point");
    const int wp = rup(W, 4);
    if (tid < MP * 4) {
      const int p = tid >> 2;
      const long g = g0 + p;
      const bool inb = g < total;
      const int s = inb ? (int)(g / N) : 0, i = inb ? (int)(g % N) : 0;
      const int k = iter_of(a, s, k_launch);
      const bool act =
          inb && (pass == 0 ? (k >= 0 && k < a.T && !a.st[4 * s + 0]) : (a.st[4 * s + 1] >= 0 && a.st[4 * s + 1] < a.T));
      const double* src = !act ? nullptr : pass == 0 ? input_density(a, s, k) : a.o_xc + (size_t)s * N;
      int lo = 0, hi = -1;
      if (pass == 0 && sym && act) {
        lo = a.win[2 * s];
        hi = a.win[2 * s + 1];
      }
#pragma unroll 4
      for (int o = tid & 3; o < wp; o += 4) {
        const int j = i + o - h;
        double x = 0.0;
        if (act && o < W && j >= 0 && j < N) {
          x = src[j];
          if (pass == 0 && sym) x = 0.5 * (x + src[mirror(j, lo, hi)]);
          x /= a.norm;
        }
        if (o == h) cot[p] = act ? x * a.norm : 0.0;  // the point's own density (cotangent of e)
        bufB[p * MLP_S + o] = x;
      }
    }
    __syncthreads();
    // ---- hidden layers (one named register array of activation derivatives per layer) ----
    // NH (compile time) = L - 1 hidden layers; only the arrays in use stay live
    double dz0[MLP_RB][4], dz1[NH > 1 ? MLP_RB : 1][4], dz2[NH > 2 ? MLP_RB : 1][4], dz3[NH > 3 ? MLP_RB : 1][4];
    double* cur = bufB;
    double* nxt = bufB;
    auto swap_bufs = [&]() {};
    if constexpr (NH > 0) { mlp_hidden_fwd(a, 0, cur, nxt, dz0, wv, l); swap_bufs(); }
    if constexpr (NH > 1) { mlp_hidden_fwd(a, 1, cur, nxt, dz1, wv, l); swap_bufs(); }
    if constexpr (NH > 2) { mlp_hidden_fwd(a, 2, cur, nxt, dz2, wv, l); swap_bufs(); }
    if constexpr (NH > 3) { mlp_hidden_fwd(a, 3, cur, nxt, dz3, wv, l); swap_bufs(); }
    // ---- output layer (fan_out 1): 4 lanes per point ----
    const int hl = a.widths[L - 1];
    const double* WL = a.mlp + a.woff[L - 1];
    {
      const int p = tid >> 2, q = tid & 3;
      double part = 0.0;
      if (p < MP)
        for (int u = q; u < hl; u += 4) part = fma(cur[p * MLP_S + u], WL[u], part);
      part += __shfl_xor(part, 1, 64);
      part += __shfl_xor(part, 2, 64);
      const long g = g0 + p;
      if (q == 0 && p < MP) {
        double de;
        const double e = out_transform(a.mlpout, part + a.mlp[a.boff[L - 1]], &de);
        if (g < total) {
          if (pass == 0)
            a.mlp_e[g] = e;
          else
            a.mlp_eout[g] = e;
        }
        if (pass == 0) cot[p] *= de;  // cotangent of the raw output (read after the next barrier)
      }
    }
    if (!backward) {
      __syncthreads();
      continue;
    }
    // ---- backward: g_{L-2} = cot * w_out * act'(z_{L-2}) ----
    __syncthreads();
    {
      const int col = wv * 16 + (l & 15);
      const double wo = col < hl ? WL[col] : 0.0;
      if constexpr (NH == 1) mlp_out_grad(nxt, cot, wo, dz0, col, l);
      if constexpr (NH == 2) mlp_out_grad(nxt, cot, wo, dz1, col, l);
      if constexpr (NH == 3) mlp_out_grad(nxt, cot, wo, dz2, col, l);
      if constexpr (NH == 4) mlp_out_grad(nxt, cot, wo, dz3, col, l);
    }
    __syncthreads();
    swap_bufs();
    // g_{l-1} = (g_l W_l^T) * act'(z_{l-1}), l = L-2 .. 1
    if constexpr (NH > 3) { mlp_hidden_bwd(a, 3, cur, nxt, dz2, wv, l); swap_bufs(); }
    if constexpr (NH > 2) { mlp_hidden_bwd(a, 2, cur, nxt, dz1, wv, l); swap_bufs(); }
    if constexpr (NH > 1) { mlp_hidden_bwd(a, 1, cur, nxt, dz0, wv, l); swap_bufs(); }
    // gX = g_0 W_0^T  (window columns): waves 0..ceil(W/16)-1
    {
      const int f1 = a.widths[1];
      const double* W0 = a.mlp + a.woff[0];
      if (wv * 16 < W) {
        d4_t acc[MLP_RB];
        tile_gemm(cur, f1, W0, f1, W, true, wv, acc);
        const int col = wv * 16 + (l & 15);
#pragma unroll
        for (int rb = 0; rb < MLP_RB; ++rb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = rb * 16 + 4 * v + (l >> 4);
            const long g = g0 + row;
            if (col < W && g < total) a.mlp_G[g * W + col] = acc[rb][v];
          }
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------------
// k_potential
// --------------------------------------------------------------------------------------------
__device__ double mlp_v_unsym(const Dev1& a, int s, int i) {
  const int N = a.N, W = a.W, h = W / 2;
  const double* G = a.mlp_G + (size_t)s * N * W;
  double v = a.mlp_e[(size_t)s * N + i];
  double acc = 0.0;
  for (int o = 0; o < W; ++o) {
    const int p = i - o + h;  // point whose window slot o reads density i
    if (p >= 0 && p < N) acc += G[(size_t)p * W + o];
  }
  return v + acc / a.norm;
}

// Global MLP on the whole density, one workgroup per system: f(n_in) with the input gradient
// (potential dx df/dn into mlp_e) and f at the pending output density; scalars in
// mlp_eout[2 s] (f(n_in)) and mlp_eout[2 s + 1] (f(n_out)).  Layer activations and derivatives
// in LDS, hmax doubles per layer (HB).
__device__ double gmlp_forward(const Dev1& a, const double* x, double* acts, double* dzs, double* part, double* red) {
  const int N = a.N, tid = threadIdx.x, HB = a.hmax;
  const int H1 = a.widths[1];
  if (H1 <= NT) {
    // layer 1: z[u] = b[u] + sum_i x_i W[i][u], split over NT / H1 interleaved row groups
    const int G = NT / H1, u = tid % H1, g = tid / H1;
    double acc = 0.0;
    if (g < G)
      for (int i = g; i < N; i += G) acc = fma(x[i], a.mlp[a.woff[0] + (size_t)i * H1 + u], acc);
    part[tid] = acc;
    __syncthreads();
    if (tid < H1) {
      double z = a.mlp[a.boff[0] + tid];
      for (int q = 0; q < G; ++q) z += part[q * H1 + tid];
      double dz;
      acts[tid] = act_fwd(a.act, z, &dz);
      dzs[tid] = dz;
    }
  } else {
    // wide first layer: every thread takes units tid, tid + NT, ... over the whole input
    for (int u = tid; u < H1; u += NT) {
      double z = a.mlp[a.boff[0] + u];
      for (int i = 0; i < N; ++i) z = fma(x[i], a.mlp[a.woff[0] + (size_t)i * H1 + u], z);
      double dz;
      acts[u] = act_fwd(a.act, z, &dz);
      dzs[u] = dz;
    }
  }
  __syncthreads();
  // hidden layers 2 .. L-1 (activations of layer l at acts + (l-1) * HB)
  for (int l = 1; l < a.nlay - 1; ++l) {
    const int fi = a.widths[l], fo = a.widths[l + 1];
    const double* prev = acts + (l - 1) * HB;
    for (int u = tid; u < fo; u += NT) {
      double z = a.mlp[a.boff[l] + u];
      for (int v = 0; v < fi; ++v) z = fma(prev[v], a.mlp[a.woff[l] + v * fo + u], z);
      double dz;
      acts[l * HB + u] = act_fwd(a.act, z, &dz);
      dzs[l * HB + u] = dz;
    }
    __syncthreads();
  }
  const int L = a.nlay - 1, fl = a.widths[L];
  double o = 0.0;
  for (int v = tid; v < fl; v += NT) o = fma(acts[(L - 1) * HB + v], a.mlp[a.woff[L] + v], o);
  double r[1] = {o};
  block_sum<1>(r, red);
  return r[0] + a.mlp[a.boff[L]];
}

__global__ __launch_bounds__(NT) void k_gmlp(Dev1 a, int k_launch) {
  extern __shared__ double smem[];
  __shared__ double red[NWAVE * 8];
  const int s = blockIdx.x, N = a.N, tid = threadIdx.x, HB = a.hmax;
  const int k = iter_of(a, s, k_launch);
  const bool doPot = k >= 0 && k < a.T && a.st[4 * s + 0] == 0;
  const bool doE = a.st[4 * s + 1] >= 0 && a.st[4 * s + 1] < a.T;
  double* x = smem;                       // N
  double* acts = x + N;                   // (nlay - 1) x HB
  double* dzs = acts + (a.nlay - 1) * HB;  // (nlay - 1) x HB
  double* part = dzs + (a.nlay - 1) * HB;  // NT
  double* dl = part + NT;                 // 2 x HB backward buffers
  const double in = 1.0 / a.norm;
  if (doPot) {
    const double* n = input_density(a, s, k);
    for (int t = tid; t < N; t += NT) x[t] = n[t] * in;
    __syncthreads();
    double df;
    const double f = out_transform(a.mlpout, gmlp_forward(a, x, acts, dzs, part, red), &df);
    // backward: d f / d acts of the last hidden layer = df W_L, then through the hidden layers
    const int L = a.nlay - 1;
    double* cur = dl;
    double* nxt = dl + HB;
    for (int v = tid; v < a.widths[L]; v += NT) cur[v] = df * a.mlp[a.woff[L] + v] * dzs[(L - 1) * HB + v];
    __syncthreads();
    for (int l = L - 1; l >= 1; --l) {
      const int fi = a.widths[l], fo = a.widths[l + 1];
      for (int v = tid; v < fi; v += NT) {
        double g = 0.0;
        for (int u = 0; u < fo; ++u) g = fma(a.mlp[a.woff[l] + v * fo + u], cur[u], g);
        nxt[v] = g * dzs[(l - 1) * HB + v];
      }
      __syncthreads();
      double* t2 = cur;
      cur = nxt;
      nxt = t2;
    }
    // v_i = dx df/dn_i = dx / norm sum_u W1[i][u] delta_u
    const int H1 = a.widths[1];
    for (int i = tid; i < N; i += NT) {
      double g = 0.0;
      for (int u = 0; u < H1; ++u) g = fma(a.mlp[a.woff[0] + (size_t)i * H1 + u], cur[u], g);
      a.mlp_e[(size_t)s * N + i] = a.dx * in * g;
    }
    if (tid == 0) a.mlp_eout[2 * s] = f;
    __syncthreads();
  }
  if (doE) {
    for (int t = tid; t < N; t += NT) x[t] = a.o_xc[(size_t)s * N + t] * in;
    __syncthreads();
    double df;
    const double f = out_transform(a.mlpout, gmlp_forward(a, x, acts, dzs, part, red), &df);
    if (tid == 0) a.mlp_eout[2 * s + 1] = f;
  }
}

__global__ __launch_bounds__(NT) void k_potential(Dev1 a, int k_launch) {
  extern __shared__ double smem[];
  const int s = blockIdx.y, tile = blockIdx.x;
  const int k = iter_of(a, s, k_launch);
  const int frozen = a.st[4 * s + 0];
  const int pending = a.st[4 * s + 1];
  const bool doPot = k >= 0 && k < a.T && !frozen;
  const bool doE = pending >= 0 && pending < a.T;
  if (!doPot && !doE) return;
  const int N = a.N;
  double* tab = smem;        // N
  double* nin = tab + N;     // N
  double* oraw = nin + N;    // N
  double* red = oraw + N;    // 2 * NT
  double* red2 = red + 2 * NT;  // 3 * NWAVE
  const double* nsrc = doPot ? input_density(a, s, k) : nullptr;
  for (int t = threadIdx.x; t < N; t += NT) {
    tab[t] = a.table[t];
    nin[t] = doPot ? nsrc[t] : 0.0;
    oraw[t] = doE ? a.o_raw[(size_t)s * N + t] : 0.0;
  }
  __syncthreads();
  const int li = threadIdx.x & (TILE - 1), qd = threadIdx.x / TILE;  // 4 j-chunks
  const int i = tile * TILE + li;
  const int chunk = (N + 3) / 4;
  const int j0 = qd * chunk, j1 = min(N, j0 + chunk);
  double accV = 0.0, accE = 0.0;
  if (i < N) {
    for (int j = j0; j < j1; ++j) {
      const double t = tab[abs(i - j)];
      accV = fma(t, nin[j], accV);
      accE = fma(t, oraw[j], accE);
    }
  }
  red[threadIdx.x] = accV;
  red[NT + threadIdx.x] = accE;
  __syncthreads();
  double eH = 0.0, eX = 0.0, fsum = 0.0;
  if (qd == 0 && i < N) {
    const double KV = ((red[li] + red[TILE + li]) + red[2 * TILE + li]) + red[3 * TILE + li];
    const double KE = ((red[NT + li] + red[NT + TILE + li]) + red[NT + 2 * TILE + li]) + red[NT + 3 * TILE + li];
    int lo = 0, hi = N - 1;
    if (a.reflect) {
      lo = a.win[2 * s];
      hi = a.win[2 * s + 1];
    }
    const bool sym = a.reflect && xc_is_local(a.xc);
    if (doPot) {
      const size_t o = ((size_t)s * a.T + k) * N + i;
      const double vH = a.dx * KV;
      double e = 0.0, v = 0.0;
      if (a.xc == QEX1D_XC_GLOBAL_MLP) {
        // potential dx df/dn from k_gmlp; the scalar f(n_in) is carried by point 0
        v = a.mlp_e[(size_t)s * N + i];
        e = i == 0 ? a.mlp_eout[2 * s] : 0.0;
      } else if (a.xc == QEX1D_XC_EXTERNAL_GLOBAL) {
        // the caller's exchange buffer: [0] f(n_in) at point 0, [1] dE/dn
        const size_t BN = (size_t)a.B * N;
        v = a.mlp[BN + (size_t)s * N + i];
        e = i == 0 ? a.mlp[(size_t)s * N] : 0.0;
      } else if (a.xc == QEX1D_XC_EXTERNAL_LOCAL) {
        // [0] e at the (flipped and averaged) input density, [1] d(sum_j e_j n_j)/dn_i there;
        // symmetrised like the network's values
        const size_t BN = (size_t)a.B * N;
        const int im = sym ? mirror(i, lo, hi) : i;
        const double* ex = a.mlp + (size_t)s * N;
        const double* vx = a.mlp + BN + (size_t)s * N;
        if (im != i) {
          v = 0.5 * (vx[i] + vx[im]);
          e = 0.5 * (ex[i] + ex[im]);
        } else {
          v = vx[i];
          e = ex[i];
        }
      } else if (a.xc == QEX1D_XC_LOCAL_MLP) {
        const int im = sym ? mirror(i, lo, hi) : i;
        const double vu = mlp_v_unsym(a, s, i);
        const double eu = a.mlp_e[(size_t)s * N + i];
        if (im != i) {
          v = 0.5 * (vu + mlp_v_unsym(a, s, im));
          e = 0.5 * (eu + a.mlp_e[(size_t)s * N + im]);
        } else {
          v = vu;
          e = eu;
        }
      } else {
        double x = nin[i];
        if (sym) x = 0.5 * (x + nin[mirror(i, lo, hi)]);
        xc_point(a, x, &e, &v);
      }
      a.vh[o] = vH;
      a.vxc[o] = v;
      if (xc_is_local(a.xc))
        a.exc[o] = e;
      else
        fsum = e;
      a.vks[(size_t)s * N + i] = vH + v + a.vext[(size_t)s * N + i];
    }
    if (doE) {
      const double ox = a.o_xc[(size_t)s * N + i];
      eH = 0.5 * a.dx * a.dx * oraw[i] * KE;
      if (a.xc == QEX1D_XC_GLOBAL_MLP) {
        eX = i == 0 ? a.dx * a.mlp_eout[2 * s + 1] : 0.0;
      } else if (a.xc == QEX1D_XC_EXTERNAL_GLOBAL) {
        eX = i == 0 ? a.dx * a.mlp[2 * (size_t)a.B * N + (size_t)s * N] : 0.0;
      } else if (a.xc == QEX1D_XC_EXTERNAL_LOCAL) {
        eX = a.dx * a.mlp[2 * (size_t)a.B * N + (size_t)s * N + i] * ox;
      } else if (a.xc == QEX1D_XC_LOCAL_MLP) {
        eX = a.dx * a.mlp_eout[(size_t)s * N + i] * ox;
      } else if (a.xc != QEX1D_XC_NONE) {
        double e, v;
        xc_point(a, ox, &e, &v);
        eX = xc_is_local(a.xc) ? a.dx * e * ox : a.dx * e;
      }
    }
  }
  double sums[3] = {eH, eX, fsum};
  block_sum<3>(sums, red2);
  if (threadIdx.x == 0) {
    if (doE) {
      a.partH[((size_t)s * a.T + pending) * a.nTiles + tile] = sums[0];
      a.partX[((size_t)s * a.T + pending) * a.nTiles + tile] = sums[1];
    }
    if (doPot) a.partF[((size_t)s * a.T + k) * a.nTiles + tile] = sums[2];
  }
}

// --------------------------------------------------------------------------------------------
// Eigensolver for the lowest eigenpairs of the pentadiagonal KS Hamiltonian
// --------------------------------------------------------------------------------------------
struct EigArgs {
  int B, N, K;
  double Td, Tb, Tc;
  const double* pot;      // [B][N] potential on the diagonal
  const int* ne;          // [B]
  const int* st;          // [B][4] (frozen flag at +0, warm flag at +2) or null
  double* vec;            // [B][K][N]
  double* eig;            // [B][K]
  const double* eig_prev; // [B][K] or null
  const double* eig_delta;
  int* pairs;             // [K*B] (s << 8) | j for every wanted eigenpair, then -1
  unsigned long long* stats;  // optional: [0] Sturm rounds executed, [1] eigenpairs solved
};

// Compacts the wanted (system, eigenpair) list so the eigen launch has no idle leading blocks.
// One workgroup of 256 threads: per chunk of 256 systems an exclusive scan of the pair counts.  In
// slot mode (kslot) only slots running an SCF iteration (0 <= kslot < T) take part.
constexpr int PAIR_NT = 256;
__global__ __launch_bounds__(PAIR_NT) void k_pairs(int B, int K, const int* ne, int* pairs, const int* kslot, int T) {
  __shared__ int scan[PAIR_NT];
  __shared__ int base;
  const int tid = threadIdx.x;
  if (tid == 0) base = 0;
  __syncthreads();
  for (int s0 = 0; s0 < B; s0 += PAIR_NT) {
    const int s = s0 + tid;
    int c = 0;
    if (s < B && (!kslot || (kslot[s] >= 0 && kslot[s] < T))) c = min(ne[s] / 2 + 1, K);
    scan[tid] = c;
    __syncthreads();
    for (int off = 1; off < PAIR_NT; off <<= 1) {  // inclusive Hillis-Steele scan
      const int v = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    const int start = base + scan[tid] - c;
    for (int j = 0; j < c; ++j) pairs[start + j] = (s << 8) | j;
    __syncthreads();
    if (tid == PAIR_NT - 1) base += scan[tid];
    __syncthreads();
  }
  for (int n = base + tid; n < K * B; n += PAIR_NT) pairs[n] = -1;
}

__device__ __forceinline__ double start_value(int i, int j) {
  // deterministic pseudo-random start vector for inverse iteration
  unsigned h = (unsigned)i * 2654435761u ^ ((unsigned)j * 40503u + 0x9e3779b9u);
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  return ((double)(h & 0xffffff) / 16777216.0) - 0.5;
}

__global__ __launch_bounds__(NT) void k_eigen(EigArgs e) {
  extern __shared__ double smem[];
  const int pair = e.pairs[blockIdx.x];  // (system << 8) | eigen index; -1 past the last pair
  if (pair < 0) return;
  const int j = pair & 255, s = pair >> 8, tid = threadIdx.x;
  if (e.st && e.st[4 * s + 0]) return;
  const int N = e.N;
  double* a = smem;         // N    diagonal T_ii + v_i
  double* y = a + N;        // N    inverse-iteration vector
  double* F = y + N;        // 4N   LDL^T factors and forward-solve values
  double* red = F + 4 * N;  // 16
  __shared__ double sh_lo, sh_hi;
  __shared__ int sh_first[NWAVE];
  const double* pot = e.pot + (size_t)s * N;
  double amin = 1e300, amax = -1e300, aabs = 0.0;
  for (int t = tid; t < N; t += NT) {
    const double v = e.Td + pot[t];
    a[t] = v;
    amin = fmin(amin, v);
    amax = fmax(amax, v);
    aabs = fmax(aabs, fabs(v));
  }
  const double off = 2.0 * (fabs(e.Tb) + fabs(e.Tc));
  const double gl = -block_max(-amin, red) - off;
  const double gu = block_max(amax, red) + off;
  const double hnorm = block_max(aabs, red) + off;
  // Cold start: the wanted states are the lowest few, so the first bracket is the bottom 40 Ha
  // of the Gershgorin interval (a miss widens it to the top bound after one round).
  double lo = gl, hi = fmin(gu, gl + 40.0);
  const bool warm = e.st && e.st[4 * s + 2] && e.eig_prev;
  if (warm) {
    const double p = e.eig_prev[(size_t)s * e.K + j];
    const double w = 3.0 * fabs(e.eig_delta[(size_t)s * e.K + j]) + 1e-9 * fmax(1.0, fabs(p));
    lo = fmax(gl, p - w);
    hi = fmin(gu, p + w);
    if (!(hi > lo)) {
      lo = gl;
      hi = gu;
    }
  }
  int nrounds = 0;
  // Multisection: NS shifts per round (both ends included) until the bracket of the j-th
  // eigenvalue is 1e-9 relative; two inverse-iteration solves and the Rayleigh quotient then
  // give the eigenpair to ~1e-13.  For N >= 4 each shift is counted by a pair of lanes, the even
  // lane eliminating the top rows 0..kt-1 and the odd lane the bottom rows N-1..kt+2 (half the
  // recurrence each), joined through the 2x2 block of rows kt, kt+1.
  const bool two = N >= 4;
  const int kt = (N - 2) / 2, kb = N - 2 - kt;
  const int NS = two ? NT / 2 : NT, me = two ? tid >> 1 : tid;
  for (int round = 0; round < 80; ++round) {
    const double mid = 0.5 * (lo + hi);
    const double tol = 1e-9 * fmax(1.0, fabs(mid)) + 64.0 * EPS * hnorm;
    if (hi - lo <= tol) break;
    ++nrounds;
    const double step = (hi - lo) / (double)(NS - 1);
    const double sig = me == NS - 1 ? hi : fma((double)me, step, lo);
    int cnt;
    if (two) {
      const bool bot = tid & 1;
      const Side h = sturm_side(bot ? a + N - 1 : a, bot ? -1 : 1, bot ? kb : kt, e.Tb, e.Tc, sig);
      const Side o{__shfl_xor(h.r1, 1, 64), __shfl_xor(h.r2, 1, 64), __shfl_xor(h.u1, 1, 64), __shfl_xor(h.neg, 1, 64)};
      cnt = bot ? sturm_join(o, h, a[kt], a[kt + 1], e.Tb, e.Tc, sig) : sturm_join(h, o, a[kt], a[kt + 1], e.Tb, e.Tc, sig);
    } else {
      cnt = sturm_count(a, N, e.Tb, e.Tc, sig);
    }
    const unsigned long long m = __ballot(cnt > j);
    if ((tid & 63) == 0) sh_first[tid >> 6] = m ? (((__ffsll((long long)m) - 1) + (tid & ~63)) >> (two ? 1 : 0)) : NS;
    __syncthreads();
    int first = NS;
#pragma unroll
    for (int q = 0; q < NWAVE; ++q) first = min(first, sh_first[q]);
    if (tid == 0) {
      // the eigenvalue lies outside the bracket: the next one is the adjacent interval 256 times
      // as wide (clipped to the Gershgorin bounds), so a missed warm start costs about one round
      const double wide = 256.0 * (hi - lo);
      if (first == 0) {
        sh_hi = lo;
        sh_lo = fmax(gl, lo - wide);
        if (!(sh_lo < lo)) sh_lo = lo > gl ? gl : lo;
      } else if (first == NS) {
        sh_lo = hi;
        sh_hi = fmin(gu, hi + wide);
        if (!(sh_hi > hi)) sh_hi = hi < gu ? gu : hi;
      } else {
        sh_lo = first - 1 == 0 ? lo : fma((double)(first - 1), step, lo);
        sh_hi = first == NS - 1 ? hi : fma((double)first, step, lo);
      }
    }
    __syncthreads();
    const double nlo = sh_lo, nhi = sh_hi;
    __syncthreads();
    if (nlo == lo && nhi == hi) break;  // no progress possible (rounding floor)
    lo = nlo;
    hi = nhi;
  }
  const double sigma = 0.5 * (lo + hi);
  if (e.stats && tid == 0) {
    atomicAdd(&e.stats[0], (unsigned long long)nrounds);
    atomicAdd(&e.stats[1], 1ull);
  }
  // inverse iteration: one LDL^T factorisation, two solves, normalised in between
  double* v = e.vec + ((size_t)s * e.K + j) * N;
  for (int t = tid; t < N; t += NT) y[t] = warm ? v[t] + 1e-3 * start_value(t, j) : start_value(t, j);
  __syncthreads();
  // For N >= 4 lanes 0 and 1 of wave 0 solve the top and bottom halves (twisted at rows kt, kt+1):
  // factorisation and forward solve of each side, the middle pair, then each side's back
  // substitution outward.  F rows of the bottom side are stored at their own row indices.
  Side fs{};
  for (int it = 0; it < 2; ++it) {
    if (two) {
      if (tid < 2) {
        const bool bot = tid == 1;
        const int st = bot ? -1 : 1, rows = bot ? kb : kt;
        double* Fs = bot ? F + 4 * (N - 1) : F;
        double* ys = bot ? y + N - 1 : y;
        const SideZ z = it == 0 ? ldlt_side_factor_forward(bot ? a + N - 1 : a, ys, st, rows, e.Tb, e.Tc, sigma,
                                                           EPS * hnorm, Fs)
                                : ldlt_side_forward(ys, st, rows, Fs);
        if (it == 0) fs = z.s;
        const Side os{__shfl_xor(fs.r1, 1, 64), __shfl_xor(fs.r2, 1, 64), __shfl_xor(fs.u1, 1, 64), 0};
        const SideZ oz{os, __shfl_xor(z.l1, 1, 64), __shfl_xor(z.m1, 1, 64), __shfl_xor(z.m2, 1, 64),
                       __shfl_xor(z.z1, 1, 64), __shfl_xor(z.z2, 1, 64)};
        double xk, xk1;
        if (bot)
          middle_solve(os, fs, oz, z, a[kt], a[kt + 1], y[kt], y[kt + 1], e.Tb, e.Tc, sigma, EPS * hnorm, xk, xk1);
        else
          middle_solve(fs, os, z, oz, a[kt], a[kt + 1], y[kt], y[kt + 1], e.Tb, e.Tc, sigma, EPS * hnorm, xk, xk1);
        ldlt_side_backward(Fs, st, rows, bot ? xk1 : xk, bot ? xk : xk1, ys);
        if (tid == 0) {
          y[kt] = xk;
          y[kt + 1] = xk1;
        }
      }
    } else if (tid == 0) {
      if (it == 0)
        ldlt_factor_forward(a, y, N, e.Tb, e.Tc, sigma, EPS * hnorm, F);
      else
        ldlt_forward(y, N, F);
      ldlt_backward(F, N, y);
    }
    __syncthreads();
    double nrm[1] = {0.0};
    for (int t = tid; t < N; t += NT) nrm[0] = fma(y[t], y[t], nrm[0]);
    block_sum<1>(nrm, red);
    const double sc = 1.0 / sqrt(nrm[0]);
    for (int t = tid; t < N; t += NT) y[t] *= sc;
    __syncthreads();
  }
  for (int t = tid; t < N; t += NT) v[t] = y[t];
  if (tid == 0) e.eig[(size_t)s * e.K + j] = sigma;
}

// --------------------------------------------------------------------------------------------
// Orthonormalise the computed eigenvectors (classical Gram-Schmidt in eigenvalue order, 2-norm),
// replace the bisection eigenvalues by Rayleigh quotients, and accumulate the density of the
// first N_e spin-orbitals (each orbital doubly occupied, the last one singly when N_e is odd).
// a (LDS) holds the Hamiltonian diagonal.  PPT = points per thread.
// --------------------------------------------------------------------------------------------
constexpr int PPT = (QEX1D_MAX_GRIDS + NT - 1) / NT;

// Cholesky QR of keig <= CQR_K orbitals, which come out of inverse iteration nearly
// orthonormal: one reduction gives the Gram matrix G = V^T V, every thread factors G = R^T R
// (the same values everywhere, so no barrier), and Q = V R^-1 point by point -- the Q of
// Gram-Schmidt in column order.  The Rayleigh quotients of all orbitals then take one more
// reduction.  Returns false, having written nothing, when a pivot of R is not clearly positive
// (nearly dependent orbitals), for the Gram-Schmidt path to handle.
constexpr int CQR_KMAX = 8;
template <int CQR_K>
__device__ bool orbitals_cqr(double* vecs, int N, int keig, int ne, double dx, const double* a, double b, double c,
                             double* lam, double* dens_reg) {
  constexpr int CQR_M = CQR_K * (CQR_K + 1) / 2;
  __shared__ double redg[NWAVE * CQR_M];
  double x[CQR_K][PPT];
#pragma unroll
  for (int j = 0; j < CQR_K; ++j)
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int t = threadIdx.x + p * NT;
      x[j][p] = (j < keig && t < N) ? vecs[(size_t)j * N + t] : 0.0;
    }
  // G (upper triangle, row by row) and then R in place
  double g[CQR_M];
#pragma unroll
  for (int i = 0, m = 0; i < CQR_K; ++i)
#pragma unroll
    for (int j = i; j < CQR_K; ++j, ++m) {
      double acc = 0.0;
#pragma unroll
      for (int p = 0; p < PPT; ++p) acc = fma(x[i][p], x[j][p], acc);
      g[m] = acc;
    }
  block_sum<CQR_M>(g, redg);
  auto at = [](int i, int j) { return i * CQR_K - i * (i - 1) / 2 + (j - i); };  // upper (i <= j)
  bool ok = true;
#pragma unroll
  for (int j = 0; j < CQR_K; ++j) {
    if (j < keig) {
#pragma unroll
      for (int i = 0; i < j; ++i) {
        double v = g[at(i, j)];
#pragma unroll
        for (int k = 0; k < i; ++k) v = fma(-g[at(k, i)], g[at(k, j)], v);
        g[at(i, j)] = v / g[at(i, i)];
      }
      double d = g[at(j, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) d = fma(-g[at(k, j)], g[at(k, j)], d);
      if (!(d > 1e-12 * g[at(j, j)])) ok = false;
      g[at(j, j)] = ok ? sqrt(d) : 1.0;
    }
  }
  if (!ok) return false;
  // Q = V R^-1, the density of the occupied orbitals, Q back to memory
  const int nocc = (ne + 1) / 2;
#pragma unroll
  for (int j = 0; j < CQR_K; ++j) {
    if (j < keig) {
      const double inv = 1.0 / g[at(j, j)];
      const double occ = j < nocc ? (double)min(2, ne - 2 * j) : 0.0;
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        double v = x[j][p];
#pragma unroll
        for (int i = 0; i < j; ++i) v = fma(-g[at(i, j)], x[i][p], v);
        v *= inv;
        x[j][p] = v;
        const int t = threadIdx.x + p * NT;
        if (t < N) {
          vecs[(size_t)j * N + t] = v;
          dens_reg[p] = fma(occ, v * v / dx, dens_reg[p]);
        }
      }
    }
  }
  __syncthreads();
  // Rayleigh quotients x^T H x with H = diag(a) + b (|i-j| = 1) + c (|i-j| = 2)
  double rq[CQR_K];
#pragma unroll
  for (int j = 0; j < CQR_K; ++j) {
    rq[j] = 0.0;
    if (j < keig) {
      const double* vj = vecs + (size_t)j * N;
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int t = threadIdx.x + p * NT;
        if (t < N) {
          double hx = a[t] * x[j][p];
          if (t >= 1) hx = fma(b, vj[t - 1], hx);
          if (t + 1 < N) hx = fma(b, vj[t + 1], hx);
          if (t >= 2) hx = fma(c, vj[t - 2], hx);
          if (t + 2 < N) hx = fma(c, vj[t + 2], hx);
          rq[j] = fma(x[j][p], hx, rq[j]);
        }
      }
    }
  }
  block_sum<CQR_K>(rq, redg);
  if (threadIdx.x == 0)
    for (int j = 0; j < keig; ++j) lam[j] = rq[j];
  return true;
}

__device__ void orbitals_to_density(double* vecs, int N, int keig, int ne, double dx, const double* a, double b,
                                    double c, double* lam, double* dens_reg, double* red) {
#pragma unroll
  for (int p = 0; p < PPT; ++p) dens_reg[p] = 0.0;
  const int nocc = (ne + 1) / 2;
  // the Gram matrix is sized to the orbital count: 6, 15 or 36 entries in one reduction
  const bool cqr = keig <= 3   ? orbitals_cqr<3>(vecs, N, keig, ne, dx, a, b, c, lam, dens_reg)
                   : keig <= 5 ? orbitals_cqr<5>(vecs, N, keig, ne, dx, a, b, c, lam, dens_reg)
                   : keig <= CQR_KMAX ? orbitals_cqr<CQR_KMAX>(vecs, N, keig, ne, dx, a, b, c, lam, dens_reg)
                                      : false;
  if (!cqr)
  for (int jj = 0; jj < keig; ++jj) {
    double x[PPT];
    double* vj = vecs + (size_t)jj * N;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int t = threadIdx.x + p * NT;
      x[p] = t < N ? vj[t] : 0.0;
    }
    for (int i0 = 0; i0 < jj; i0 += 4) {
      double d[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (i0 + q < jj) {
          const double* vi = vecs + (size_t)(i0 + q) * N;
#pragma unroll
          for (int p = 0; p < PPT; ++p) {
            const int t = threadIdx.x + p * NT;
            if (t < N) d[q] = fma(x[p], vi[t], d[q]);
          }
        }
      }
      block_sum<4>(d, red);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (i0 + q < jj) {
          const double* vi = vecs + (size_t)(i0 + q) * N;
#pragma unroll
          for (int p = 0; p < PPT; ++p) {
            const int t = threadIdx.x + p * NT;
            if (t < N) x[p] = fma(-d[q], vi[t], x[p]);
          }
        }
      }
    }
    double nrm[1] = {0.0};
#pragma unroll
    for (int p = 0; p < PPT; ++p) nrm[0] = fma(x[p], x[p], nrm[0]);
    block_sum<1>(nrm, red);
    const double sc = 1.0 / sqrt(nrm[0]);
    const double occ = jj < nocc ? (double)min(2, ne - 2 * jj) : 0.0;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int t = threadIdx.x + p * NT;
      x[p] *= sc;
      if (t < N) {
        vj[t] = x[p];
        dens_reg[p] = fma(occ, x[p] * x[p] / dx, dens_reg[p]);
      }
    }
    __syncthreads();
    // Rayleigh quotient x^T H x with H = diag(a) + b (|i-j| = 1) + c (|i-j| = 2)
    double rq[1] = {0.0};
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int t = threadIdx.x + p * NT;
      if (t < N) {
        double hx = a[t] * x[p];
        if (t >= 1) hx = fma(b, vj[t - 1], hx);
        if (t + 1 < N) hx = fma(b, vj[t + 1], hx);
        if (t >= 2) hx = fma(c, vj[t - 2], hx);
        if (t + 2 < N) hx = fma(c, vj[t + 2], hx);
        rq[0] = fma(x[p], hx, rq[0]);
      }
    }
    block_sum<1>(rq, red);
    if (threadIdx.x == 0) lam[jj] = rq[0];
  }
  // Rayleigh quotients of (nearly) degenerate pairs can come out of order by rounding
  if (threadIdx.x == 0) {
    for (int x = 1; x < keig; ++x) {
      const double t = lam[x];
      int y = x - 1;
      while (y >= 0 && lam[y] > t) {
        lam[y + 1] = lam[y];
        --y;
      }
      lam[y + 1] = t;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(NT) void k_density(Dev1 a, int k_launch, double alpha_launch) {
  extern __shared__ double smem[];
  __shared__ double red[NWAVE * 8];
  __shared__ double lam[QEX1D_MAX_ELECTRONS / 2 + 1];
  const int s = blockIdx.x, tid = threadIdx.x, N = a.N, T = a.T;
  const int k = iter_of(a, s, k_launch);
  if (k < 0 || k >= T) return;  // idle slot, or a slot in its final energy-only pass
  double alpha_k = alpha_launch;  // (set below in slot mode)
  if (a.kslot) {  // the same product the batch loop forms on the host, so slot and batch results agree bitwise
    alpha_k = a.alpha0;
    for (int i = 0; i < k; ++i) alpha_k *= a.decay;
  }
  const bool frozen = k > 0 && a.st[4 * s + 0] != 0;  // a frozen state repeats iteration k - 1
  const bool warm_before = a.st[4 * s + 2] != 0;
  const size_t cur = ((size_t)s * T + k) * N;
  if (frozen) {
    const size_t prv = cur - N;
    for (int t = tid; t < N; t += NT) {
      a.dens[cur + t] = a.dens[prv + t];
      a.vh[cur + t] = a.vh[prv + t];
      a.vxc[cur + t] = a.vxc[prv + t];
      if (xc_is_local(a.xc)) a.exc[cur + t] = a.exc[prv + t];
    }
    if (tid == 0) {
      a.gap[(size_t)s * T + k] = a.gap[(size_t)s * T + k - 1];
      a.conv[(size_t)s * T + k] = 1;
      a.src[(size_t)s * T + k] = a.src[(size_t)s * T + k - 1];
      a.st[4 * s + 1] = -1;
    }
    if (a.eigs_out && tid < a.K)
      a.eigs_out[((size_t)s * T + k) * a.K + tid] = a.eigs_out[((size_t)s * T + k - 1) * a.K + tid];
    if (a.orb_out) {
      const int kk = min(a.ne[s] / 2 + 1, a.K);
      for (int t = tid; t < kk * N; t += NT)
        a.orb_out[((size_t)s * T + k) * a.K * N + t] = a.orb_out[((size_t)s * T + k - 1) * a.K * N + t];
    }
    return;
  }
  const int ne = max(1, min(a.ne[s], 2 * a.K - 1));  // contract: N_e <= max_electrons (clamped as a guard)
  const int nocc = (ne + 1) / 2;
  const int keig = ne / 2 + 1;
  double* hd = smem;  // Hamiltonian diagonal of this iteration
  for (int t = tid; t < N; t += NT) hd[t] = a.Td + a.vks[(size_t)s * N + t];
  __syncthreads();
  double dn[PPT];
  orbitals_to_density(a.vec + (size_t)s * a.K * N, N, keig, ne, a.dx, hd, a.Tb, a.Tc, lam, dn, red);
  if (a.orb_out)
    for (int t = tid; t < keig * N; t += NT)
      a.orb_out[((size_t)s * T + k) * a.K * N + t] = a.vec[(size_t)s * a.K * N + t];
  double eeig = 0.0;
  for (int jj = 0; jj < nocc; ++jj) eeig += (double)min(2, ne - 2 * jj) * lam[jj];
  const double gap = lam[ne / 2] - lam[(ne - 1) / 2];
  // energies of the unmixed output density
  double ed[2] = {0.0, 0.0};
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int t = tid + p * NT;
    if (t < N) {
      ed[0] = fma(a.vks[(size_t)s * N + t], dn[p], ed[0]);
      ed[1] = fma(a.vext[(size_t)s * N + t], dn[p], ed[1]);
    }
  }
  block_sum<2>(ed, red);
  // reflection, differences, mixing
  int lo = 0, hi = N - 1;
  if (a.reflect) {
    lo = a.win[2 * s];
    hi = a.win[2 * s + 1];
  }
  double* oraw = a.o_raw + (size_t)s * N;
  double* oxc = a.o_xc + (size_t)s * N;
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int t = tid + p * NT;
    if (t < N) oraw[t] = dn[p];
  }
  __syncthreads();
  const double* nin = input_density(a, s, k);
  const int M = a.M;
  const int slot = k % M;
  const int cnt = min(M, k + 1);
  double* hist = a.hist + (size_t)s * M * N;
  double mse[2] = {0.0, 0.0};
  const bool symxc = a.reflect && xc_is_local(a.xc);
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int t = tid + p * NT;
    if (t < N) {
      double of = dn[p];
      if (a.reflect) {
        const int tm = mirror(t, lo, hi);
        of = 0.5 * (oraw[t] + oraw[tm]);
      }
      oxc[t] = symxc ? of : dn[p];
      const double d = of - nin[t];
      hist[(size_t)slot * N + t] = d;
      double acc = 0.0;
      for (int q = 0; q < cnt; ++q) {
        const int sl = ((k - cnt + 1 + q) % M + M) % M;
        acc += q == cnt - 1 ? d : hist[(size_t)sl * N + t];
      }
      const double mixed = nin[t] + alpha_k * (acc / (double)cnt);
      a.dens[cur + t] = mixed;
      mse[0] = fma(d, d, mse[0]);
      const double dm = mixed - nin[t];
      mse[1] = fma(dm, dm, mse[1]);
    }
  }
  block_sum<2>(mse, red);
  if (tid == 0) {
    const double tol = a.tol;
    bool conv_out, frozen_next;
    if (a.loop == QEX1D_LOOP_PYTHON) {
      conv_out = mse[0] / (double)N < tol;
      frozen_next = conv_out;
    } else {
      conv_out = false;
      frozen_next = mse[1] / (double)N < tol;
    }
    const size_t sk = (size_t)s * T + k;
    a.conv[sk] = conv_out ? 1 : 0;
    a.gap[sk] = gap;
    a.src[sk] = k;
    a.terms[sk * 4 + 0] = eeig;
    a.terms[sk * 4 + 1] = a.dx * ed[0];
    a.terms[sk * 4 + 2] = a.dx * ed[1];
    a.st[4 * s + 0] = frozen_next ? 1 : 0;
    a.st[4 * s + 1] = k;
    a.st[4 * s + 2] = 1;
  }
  if (tid < keig) {
    const size_t ix = (size_t)s * a.K + tid;
    const double l = lam[tid];
    a.eig[ix] = l;
    a.eig_delta[ix] = warm_before ? l - a.eig_prev[ix] : 1.0;
    a.eig_prev[ix] = l;
    if (a.eigs_out) a.eigs_out[((size_t)s * T + k) * a.K + tid] = l;
  } else if (a.eigs_out && tid < a.K) {
    a.eigs_out[((size_t)s * T + k) * a.K + tid] = NAN;
  }
}

// total energies of system s's T iterations from the per-tile partials (and the global XC energy)
__device__ void finalize_system(const Dev1& a, int s) {
  for (int k = threadIdx.x; k < a.T; k += blockDim.x) {
    const size_t sk = (size_t)s * a.T + k;
    const int src = a.src[sk];
    const size_t ss = (size_t)s * a.T + src;
    double eh = 0.0, ex = 0.0, fs = 0.0;
    for (int t = 0; t < a.nTiles; ++t) {
      eh += a.partH[ss * a.nTiles + t];
      ex += a.partX[ss * a.nTiles + t];
      fs += a.partF[ss * a.nTiles + t];
    }
    const double* tm = a.terms + ss * 4;
    a.etot[sk] = (((tm[0] - tm[1]) + eh) + ex) + tm[2];
    if (!xc_is_local(a.xc)) a.exc[sk] = fs;
  }
}

__global__ void k_finalize(Dev1 a) { finalize_system(a, blockIdx.x); }

// slot mode: the slots a new request was placed in start at iteration 0 with a cold state; idle
// slots (never admitted, or finished) are held with no pending energy and no frozen flag, whatever
// the workspace held before (the caller's workspace is not cleared)
__global__ void k_slot_begin(Dev1 a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.B) return;
  if (a.admit[s]) {
    a.st[4 * s + 0] = 0;
    a.st[4 * s + 1] = -1;
    a.st[4 * s + 2] = 0;
    a.st[4 * s + 3] = 0;
    a.kslot[s] = 0;
  } else if (a.kslot[s] < 0 || a.kslot[s] > a.T) {
    a.st[4 * s + 0] = 0;
    a.st[4 * s + 1] = -1;
  }
}

// slot mode, end of a step: a slot that ran its energy-only pass (iteration T) is finalised and
// marked done (-2, no pending energy); the others move on to their next iteration
__global__ void k_slot_end(Dev1 a) {
  const int s = blockIdx.x;
  const int k = a.kslot[s];
  if (k == a.T) {
    finalize_system(a, s);
    __syncthreads();
    if (threadIdx.x == 0) {
      a.st[4 * s + 1] = -1;
      a.kslot[s] = -2;
    }
  } else if (k >= 0 && k < a.T && threadIdx.x == 0) {
    a.kslot[s] = k + 1;
  }
}

// --------------------------------------------------------------------------------------------
// non-interacting solve: density of the first N_e spin-orbitals
// --------------------------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_nonint_density(int N, int K, double dx, double Td, double Tb, double Tc,
                                                       const double* pot, const int* nes, double* vec,
                                                       double* density, double* etot, double* gap, double* eig_out,
                                                       double* wf_out) {
  extern __shared__ double smem[];
  __shared__ double red[NWAVE * 8];
  __shared__ double lam[QEX1D_MAX_ELECTRONS / 2 + 1];
  const int s = blockIdx.x, tid = threadIdx.x;
  // the caller guarantees 1 <= N_e <= max_electrons (K = max_electrons / 2 + 1); the clamp keeps a
  // violated contract inside this system's K orbital slots
  const int ne = max(1, min(nes[s], 2 * K - 1)), nocc = (ne + 1) / 2, keig = ne / 2 + 1;
  double* hd = smem;
  for (int t = tid; t < N; t += NT) hd[t] = Td + pot[(size_t)s * N + t];
  __syncthreads();
  double dn[PPT];
  orbitals_to_density(vec + (size_t)s * K * N, N, keig, ne, dx, hd, Tb, Tc, lam, dn, red);
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int t = tid + p * NT;
    if (t < N) density[(size_t)s * N + t] = dn[p];
  }
  if (tid == 0) {
    double ee = 0.0;
    for (int jj = 0; jj < nocc; ++jj) ee += (double)min(2, ne - 2 * jj) * lam[jj];
    if (etot) etot[s] = ee;
    if (gap) gap[s] = lam[ne / 2] - lam[(ne - 1) / 2];
  }
  if (eig_out && tid < K) eig_out[(size_t)s * K + tid] = tid < keig ? lam[tid] : NAN;
  if (wf_out) {
    const double sc = 1.0 / sqrt(dx);
    for (int jj = 0; jj < K; ++jj) {
      for (int t = tid; t < N; t += NT)
        wf_out[((size_t)s * K + jj) * N + t] = jj < keig ? vec[((size_t)s * K + jj) * N + t] * sc : 0.0;
    }
  }
}

// --------------------------------------------------------------------------------------------
// external potential and Hartree potential
// --------------------------------------------------------------------------------------------
__global__ void k_external(int B, int N, const double* grids, int A, const double* loc, const double* z, int kind,
                           double p0, double p1, double* out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * N) return;
  const int s = idx / N, i = idx % N;
  const double x = grids[i];
  double v = 0.0;
  for (int q = 0; q < A; ++q) {
    const double r = x - loc[(size_t)s * A + q];
    const double u = kind == QEX_INTERACTION_EXPONENTIAL ? p0 * exp(-fabs(r) * p1) : 1.0 / sqrt(r * r + p0);
    v = fma(-z[(size_t)s * A + q], u, v);
  }
  out[idx] = v;
}

__global__ __launch_bounds__(NT) void k_hartree(int N, double dx, const double* table, const double* dens,
                                                double* out) {
  extern __shared__ double smem[];
  const int s = blockIdx.y;
  double* tab = smem;
  double* n = smem + N;
  for (int t = threadIdx.x; t < N; t += NT) {
    tab[t] = table[t];
    n[t] = dens[(size_t)s * N + t];
  }
  __syncthreads();
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  for (int j = 0; j < N; ++j) acc = fma(tab[abs(i - j)], n[j], acc);
  out[(size_t)s * N + i] = dx * acc;
}

// k_xc_gather: the densities a caller-evaluated (EXTERNAL) functional needs before pass k: the
// iteration's input density (flipped and averaged for LOCAL kinds under reflection, as the
// network kernels read it) and the pending output density o_xc (the input density at k = 0).
__global__ __launch_bounds__(NT) void k_xc_gather(Dev1 a, int k, double* din, double* dout) {
  const int s = blockIdx.x, N = a.N;
  const double* nin = input_density(a, s, k);
  const bool sym = a.reflect && xc_is_local(a.xc);
  int lo = 0, hi = N - 1;
  if (sym) {
    lo = a.win[2 * s];
    hi = a.win[2 * s + 1];
  }
  for (int t = threadIdx.x; t < N; t += NT) {
    double x = nin[t];
    if (sym) x = 0.5 * (x + nin[mirror(t, lo, hi)]);
    din[(size_t)s * N + t] = x;
    dout[(size_t)s * N + t] = k == 0 ? x : a.o_xc[(size_t)s * N + t];
  }
}

// k_xc_eval: the XC functional on given densities (a.n0), one workgroup per system, after the
// network kernel has run on them as iteration 0 of a one-iteration batch (its outputs in mlp_e,
// mlp_G, mlp_eout).  LOCAL kinds: e_i and v_i = e_i + sum_j n_j de_j/dn_i per point (the
// derivative of sum_j e_j n_j), E = dx sum_i e_i n_i.  GLOBAL kinds: v_i = dE/dn_i, the value
// (sum_i f(n_i), or the network's f(n)) and E = dx * value.  eps, vxc, exc may each be null.
__global__ __launch_bounds__(NT) void k_xc_eval(Dev1 a, double* eps, double* vxc, double* exc) {
  __shared__ double red[NWAVE];
  const int s = blockIdx.x, N = a.N;
  const double* n = a.n0 + (size_t)s * N;
  const bool local = xc_is_local(a.xc);
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < N; i += NT) {
    double e = 0.0, v = 0.0;
    if (a.xc == QEX1D_XC_GLOBAL_MLP) {
      v = a.mlp_e[(size_t)s * N + i];
    } else if (a.xc == QEX1D_XC_LOCAL_MLP) {
      e = a.mlp_e[(size_t)s * N + i];
      v = mlp_v_unsym(a, s, i);
    } else {
      xc_point(a, n[i], &e, &v);
    }
    if (local) {
      if (eps) eps[(size_t)s * N + i] = e;
      acc[0] = fma(e, n[i], acc[0]);
    } else {
      acc[0] += e;
    }
    if (vxc) vxc[(size_t)s * N + i] = v;
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) {
    const double val = a.xc == QEX1D_XC_GLOBAL_MLP ? a.mlp_eout[2 * s] : acc[0];
    if (!local && eps) eps[s] = val;
    if (exc) exc[s] = a.dx * val;
  }
}

// --------------------------------------------------------------------------------------------
// host-side layout
// --------------------------------------------------------------------------------------------
struct Layout {
  size_t o_raw, o_xc, vks, hist, vec, eig, eig_prev, eig_delta, terms, partH, partX, partF, src, st, mlp_e, mlp_G,
      mlp_eout, pairs, total;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

Layout layout1d(const qex1d_problem* p) {
  const size_t B = p->batch, N = p->num_grids, T = p->num_iterations;
  const size_t K = p->max_electrons / 2 + 1, M = p->num_mixing_iterations;
  const size_t nT = (N + TILE - 1) / TILE;
  const bool mlp = p->xc_kind == QEX1D_XC_LOCAL_MLP || p->xc_kind == QEX1D_XC_GLOBAL_MLP;
  const size_t W = p->xc_kind == QEX1D_XC_LOCAL_MLP ? p->mlp_widths[0] : 0;
  Layout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  L.o_raw = take(B * N * 8);
  L.o_xc = take(B * N * 8);
  L.vks = take(B * N * 8);
  L.hist = take(B * M * N * 8);
  L.vec = take(B * K * N * 8);
  L.eig = take(B * K * 8);
  L.eig_prev = take(B * K * 8);
  L.eig_delta = take(B * K * 8);
  L.terms = take(B * T * 4 * 8);
  L.partH = take(B * T * nT * 8);
  L.partX = take(B * T * nT * 8);
  L.partF = take(B * T * nT * 8);
  L.src = take(B * T * 4);
  L.st = take(B * 4 * 4);
  L.mlp_e = take(mlp ? B * N * 8 : 0);
  L.mlp_G = take(mlp ? B * N * W * 8 : 0);
  L.mlp_eout = take(mlp ? B * (N > 2 ? N : 2) * 8 : 0);
  L.pairs = take(B * K * 4);
  L.total = off;
  return L;
}

int validate1d(const qex1d_problem* p) {
  if (!p) return fail(QEX_E_BADARG, "null problem");
  if (p->batch < 1) return fail(QEX_E_BADARG, "batch must be >= 1 (got %d)", p->batch);
  if (p->num_grids < 5 || p->num_grids > QEX1D_MAX_GRIDS)
    return fail(QEX_E_BADARG, "num_grids must be in [5, %d] (got %d)", QEX1D_MAX_GRIDS, p->num_grids);
  if (!(p->dx > 0.0)) return fail(QEX_E_BADARG, "dx must be positive");
  if (p->num_iterations < 1) return fail(QEX_E_BADARG, "num_iterations must be >= 1");
  if (p->max_electrons < 1 || p->max_electrons > QEX1D_MAX_ELECTRONS)
    return fail(QEX_E_BADARG, "max_electrons must be in [1, %d]", QEX1D_MAX_ELECTRONS);
  if (p->max_electrons / 2 + 1 > p->num_grids) return fail(QEX_E_BADARG, "too many electrons for the grid");
  if (p->num_mixing_iterations < 1) return fail(QEX_E_BADARG, "num_mixing_iterations must be >= 1");
  if (p->loop != QEX1D_LOOP_PYTHON && p->loop != QEX1D_LOOP_SCAN) return fail(QEX_E_BADARG, "bad loop kind");
  if (p->xc_kind < 0 || p->xc_kind > QEX1D_XC_EXTERNAL_GLOBAL) return fail(QEX_E_BADARG, "bad xc_kind");
  if ((p->xc_kind == QEX1D_XC_GLOBAL_MLP || p->xc_kind == QEX1D_XC_LOCAL_MLP) &&
      p->mlp_output != QEX1D_MLP_OUTPUT_LINEAR && p->mlp_output != QEX1D_MLP_OUTPUT_NEGATIVITY)
    return fail(QEX_E_BADARG, "bad mlp_output");
  if (p->xc_kind == QEX1D_XC_GLOBAL_MLP) {
    if (p->mlp_layers < 2 || p->mlp_layers > QEX1D_MAX_MLP_LAYERS)
      return fail(QEX_E_UNSUPPORTED, "GLOBAL_MLP needs 2..%d Dense layers", QEX1D_MAX_MLP_LAYERS);
    if (p->mlp_widths[0] != p->num_grids) return fail(QEX_E_BADARG, "GLOBAL_MLP input width must be num_grids");
    if (p->mlp_widths[p->mlp_layers] != 1) return fail(QEX_E_UNSUPPORTED, "MLP output width must be 1");
    for (int l = 1; l < p->mlp_layers; ++l)
      if (p->mlp_widths[l] < 1 || p->mlp_widths[l] > QEX1D_MAX_MLP_WIDTH)
        return fail(QEX_E_UNSUPPORTED, "GLOBAL_MLP hidden widths must be in [1, %d]", QEX1D_MAX_MLP_WIDTH);
    if (p->mlp_activation < 0 || p->mlp_activation > QEX_ACT_GELU_TANH) return fail(QEX_E_BADARG, "bad activation");
    if (!(p->mlp_norm != 0.0)) return fail(QEX_E_BADARG, "mlp_norm must be nonzero");
  }
  if (p->xc_kind == QEX1D_XC_LOCAL_MLP) {
    if (p->mlp_layers < 2 || p->mlp_layers > QEX1D_MAX_MLP_LAYERS)
      return fail(QEX_E_UNSUPPORTED, "LOCAL_MLP needs 2..%d Dense layers", QEX1D_MAX_MLP_LAYERS);
    if (p->mlp_widths[0] < 1 || p->mlp_widths[0] > QEX1D_MAX_MLP_WINDOW || (p->mlp_widths[0] % 2) == 0)
      return fail(QEX_E_UNSUPPORTED, "MLP window must be odd and <= %d", QEX1D_MAX_MLP_WINDOW);
    if (p->mlp_widths[p->mlp_layers] != 1) return fail(QEX_E_UNSUPPORTED, "MLP output width must be 1");
    for (int l = 1; l < p->mlp_layers; ++l)
      if (p->mlp_widths[l] < 1 || p->mlp_widths[l] > QEX1D_MAX_MLP_WIDTH)
        return fail(QEX_E_UNSUPPORTED, "MLP hidden width must be in [1, %d]", QEX1D_MAX_MLP_WIDTH);
    if (p->mlp_activation < 0 || p->mlp_activation > QEX_ACT_GELU_TANH) return fail(QEX_E_BADARG, "bad activation");
    if (!(p->mlp_norm != 0.0)) return fail(QEX_E_BADARG, "mlp_norm must be nonzero");
  }
  return QEX_OK;
}

size_t mlp_lds_bytes(const Dev1& a, int lpp) {
  const size_t stride = a.wmax + 1, P = NT / lpp;
  return ((size_t)a.N + 2 * P * stride + (size_t)(a.nlay - 1) * P * stride) * 8;
}
constexpr size_t LDS_MAX = 160 * 1024;
size_t gmlp_lds_bytes(const Dev1& a) {
  return ((size_t)a.N + 2 * (size_t)(a.nlay - 1) * a.hmax + NT + 2 * (size_t)a.hmax) * 8;
}

size_t eigen_lds_bytes(int N) { return (size_t)(6 * N + 16) * 8; }

void kinetic(double dx, double* Td, double* Tb, double* Tc) {
  // T = -1/2 L / dx^2 with L = (-1/12, 4/3, -5/2, 4/3, -1/12)
  const double s = -0.5 / (dx * dx);
  *Td = s * -2.5;
  *Tb = s * (4.0 / 3.0);
  *Tc = s * (-1.0 / 12.0);
}

}  // namespace

// ============================================================================================
// C ABI
// ============================================================================================
extern "C" {

const char* qex_last_error(void) { return g_err; }
int32_t qex_abi_version(void) { return 1; }

int qex_profile_enable(int32_t enable, int32_t max_launches) {
  using namespace qexprof;
  if (enable) {
    if (max_launches < 1) return fail(QEX_E_BADARG, "max_launches must be >= 1");
    if (max_launches > g_cap) {
      Rec* np = (Rec*)realloc(g_rec, sizeof(Rec) * max_launches);
      if (!np) return fail(QEX_E_BADARG, "out of host memory");
      g_rec = np;
      g_cap = max_launches;
    }
    const int need = 2 * max_launches;
    if (need > g_pool_n) {
      hipEvent_t* nq = (hipEvent_t*)realloc(g_pool, sizeof(hipEvent_t) * need);
      if (!nq) return fail(QEX_E_BADARG, "out of host memory");
      g_pool = nq;
      for (int i = g_pool_n; i < need; ++i) {
        if (hipEventCreate(&g_pool[i]) != hipSuccess) return fail(QEX_E_LAUNCH, "hipEventCreate failed");
        g_pool_n = i + 1;
      }
    }
    g_n = 0;
    g_pool_used = 0;
  }
  g_on = enable != 0;
  return QEX_OK;
}

int qex_profile_kinds(uint32_t mask) {
  qexprof::g_mask = mask;
  return QEX_OK;
}

int qex_profile_read(double* ms_per_kind, int32_t* launches_per_kind, int32_t nkinds) {
  using namespace qexprof;
  if (!ms_per_kind || !launches_per_kind || nkinds < 1) return fail(QEX_E_BADARG, "bad arguments");
  for (int k = 0; k < nkinds; ++k) {
    ms_per_kind[k] = 0.0;
    launches_per_kind[k] = 0;
  }
  for (int i = 0; i < g_n; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_rec[i].b) != hipSuccess || hipEventElapsedTime(&ms, g_rec[i].a, g_rec[i].b) != hipSuccess)
      return fail(QEX_E_LAUNCH, "event timing failed");
    const int k = g_rec[i].kind;
    if (k < nkinds) {
      ms_per_kind[k] += ms;
      launches_per_kind[k] += 1;
    }
  }
  g_n = 0;
  g_pool_used = 0;
  return QEX_OK;
}

size_t qex1d_workspace_bytes(const qex1d_problem* p) {
  if (validate1d(p) != QEX_OK) return 0;
  return layout1d(p).total;
}

}  // extern "C"

namespace {
// Everything one SCF iteration's launches need: the device view, the eigen arguments, launch shapes.
struct Plan1d {
  Dev1 a;
  EigArgs e;
  size_t lds_pot, lds_eig, lds_mlp, lds_gmlp;
  int mlp_lpp;  // lanes per point of the generic local-MLP kernel
  bool mlp_mfma;
  dim3 gmlp_mfma, gpot, geig, gmlp;
};

void fill_network(const qex1d_problem* p, Dev1& a);
int plan_network(Plan1d* P);

int plan1d(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out, void* workspace,
           size_t workspace_bytes, Plan1d* P) {
  int rc = validate1d(p);
  if (rc) return rc;
  if (!in || !out) return fail(QEX_E_BADARG, "null inputs/outputs");
  if (!in->interaction_table || !in->external_potential || !in->initial_density || !in->num_electrons)
    return fail(QEX_E_BADARG, "missing input array");
  if (p->reflect && !in->reflect_window) return fail(QEX_E_BADARG, "reflect needs reflect_window");
  if ((p->xc_kind == QEX1D_XC_LOCAL_MLP || p->xc_kind == QEX1D_XC_GLOBAL_MLP) && !in->xc_params)
    return fail(QEX_E_BADARG, "MLP functionals need xc_params");
  if (xc_is_external(p->xc_kind) && !in->xc_params)
    return fail(QEX_E_BADARG, "EXTERNAL functionals need the exchange buffer in xc_params");
  if (!out->density || !out->total_energy || !out->hartree_potential || !out->xc_potential ||
      !out->xc_energy_density || !out->gap || !out->converged)
    return fail(QEX_E_BADARG, "missing output array");
  const Layout L = layout1d(p);
  if (!workspace || workspace_bytes < L.total)
    return fail(QEX_E_WORKSPACE, "workspace too small: need %zu bytes, got %zu", L.total, workspace_bytes);
  char* ws = (char*)workspace;
  Dev1& a = P->a;
  a = Dev1{};
  a.B = p->batch;
  a.N = p->num_grids;
  a.T = p->num_iterations;
  a.K = p->max_electrons / 2 + 1;
  a.M = p->num_mixing_iterations;
  a.nTiles = (a.N + TILE - 1) / TILE;
  a.loop = p->loop;
  a.reflect = p->reflect ? 1 : 0;
  a.xc = p->xc_kind;
  a.dx = p->dx;
  a.tol = p->density_mse_converge_tolerance;
  a.xc_c = p->xc_c;
  a.xc_p = p->xc_p;
  a.xc_eps = p->xc_eps;
  kinetic(p->dx, &a.Td, &a.Tb, &a.Tc);
  fill_network(p, a);
  a.table = in->interaction_table;
  a.vext = in->external_potential;
  a.n0 = in->initial_density;
  a.ne = in->num_electrons;
  a.win = in->reflect_window;
  a.mlp = in->xc_params;
  a.dens = out->density;
  a.etot = out->total_energy;
  a.vh = out->hartree_potential;
  a.vxc = out->xc_potential;
  a.exc = out->xc_energy_density;
  a.gap = out->gap;
  a.conv = out->converged;
  a.eigs_out = out->eigenvalues;
  a.orb_out = out->orbitals;
  a.o_raw = (double*)(ws + L.o_raw);
  a.o_xc = (double*)(ws + L.o_xc);
  a.vks = (double*)(ws + L.vks);
  a.hist = (double*)(ws + L.hist);
  a.vec = (double*)(ws + L.vec);
  a.eig = (double*)(ws + L.eig);
  a.eig_prev = (double*)(ws + L.eig_prev);
  a.eig_delta = (double*)(ws + L.eig_delta);
  a.terms = (double*)(ws + L.terms);
  a.partH = (double*)(ws + L.partH);
  a.partX = (double*)(ws + L.partX);
  a.partF = (double*)(ws + L.partF);
  a.src = (int*)(ws + L.src);
  a.st = (int*)(ws + L.st);
  a.mlp_e = (double*)(ws + L.mlp_e);
  a.mlp_G = (double*)(ws + L.mlp_G);
  a.mlp_eout = (double*)(ws + L.mlp_eout);

  EigArgs& e = P->e;
  e = EigArgs{};
  e.B = a.B;
  e.N = a.N;
  e.K = a.K;
  e.Td = a.Td;
  e.Tb = a.Tb;
  e.Tc = a.Tc;
  e.pot = a.vks;
  e.ne = a.ne;
  e.st = a.st;
  e.vec = a.vec;
  e.eig = a.eig;
  e.eig_prev = a.eig_prev;
  e.eig_delta = a.eig_delta;
  e.pairs = (int*)(ws + L.pairs);
  e.stats = (unsigned long long*)out->stats;

  P->lds_pot = (size_t)(3 * a.N + 2 * NT + 3 * NWAVE) * 8;
  P->lds_eig = eigen_lds_bytes(a.N);
  P->geig = dim3(a.K * a.B);
  if (P->lds_eig > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_eigen, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->lds_eig);
  return plan_network(P);
}

// The network's fields of the device view (widths, packed offsets, widest layers), from p.
void fill_network(const qex1d_problem* p, Dev1& a) {
  if (a.xc != QEX1D_XC_LOCAL_MLP && a.xc != QEX1D_XC_GLOBAL_MLP) return;
  a.nlay = p->mlp_layers;
  a.act = p->mlp_activation;
  a.norm = p->mlp_norm;
  a.mlpout = p->mlp_output;
  a.W = p->mlp_widths[0];
  int off = 0;
  a.wmax = a.W;
  for (int l = 0; l <= a.nlay; ++l) a.widths[l] = p->mlp_widths[l];
  for (int l = 0; l < a.nlay; ++l) {
    a.woff[l] = off;
    off += a.widths[l] * a.widths[l + 1];
    a.boff[l] = off;
    off += a.widths[l + 1];
    if (l + 1 < a.nlay) a.wmax = a.wmax > a.widths[l + 1] ? a.wmax : a.widths[l + 1];
  }
  a.hmax = 1;
  for (int l = 1; l < a.nlay; ++l) a.hmax = a.hmax > a.widths[l] ? a.hmax : a.widths[l];
}

// Launch shapes and LDS of the network kernels and of the potential kernel's grid.
int plan_network(Plan1d* P) {
  const Dev1& a = P->a;
  bool mlp_mfma = a.xc == QEX1D_XC_LOCAL_MLP && a.nlay - 1 <= MLP_LH && a.W <= MLP_S - 1;
  for (int l = 1; l < a.nlay && mlp_mfma; ++l) mlp_mfma = a.widths[l] <= MLP_HMAX;
  P->mlp_mfma = mlp_mfma;
  // the generic path: 8 lanes per point (32 points per workgroup) when their activations fit in
  // LDS, else 64 lanes per point (4 points)
  P->mlp_lpp = mlp_lds_bytes(a, 8) <= LDS_MAX ? 8 : 64;
  P->lds_mlp = a.xc != QEX1D_XC_LOCAL_MLP ? 0 : mlp_mfma ? (size_t)(MP * MLP_S + MP) * 8 : mlp_lds_bytes(a, P->mlp_lpp);
  if (P->lds_mlp > LDS_MAX) return fail(QEX_E_UNSUPPORTED, "MLP too wide for LDS (%zu bytes)", P->lds_mlp);
  P->lds_gmlp = a.xc != QEX1D_XC_GLOBAL_MLP ? 0 : gmlp_lds_bytes(a);
  if (P->lds_gmlp > LDS_MAX) return fail(QEX_E_UNSUPPORTED, "global MLP too wide for LDS (%zu bytes)", P->lds_gmlp);
  P->gmlp_mfma = dim3((unsigned)(((long)a.B * a.N + MP - 1) / MP));
  P->gpot = dim3(a.nTiles, a.B);
  P->gmlp = dim3((a.N + NT / P->mlp_lpp - 1) / (NT / P->mlp_lpp), a.B);
  const void* mlp_fn = P->mlp_lpp == 8 ? (const void*)k_mlp<8> : (const void*)k_mlp<64>;
  if (mlp_mfma) {
    const void* fns[4] = {(const void*)k_mlp_mfma<1>, (const void*)k_mlp_mfma<2>, (const void*)k_mlp_mfma<3>,
                          (const void*)k_mlp_mfma<4>};
    mlp_fn = fns[a.nlay - 2];
  }
  if (P->lds_mlp > 64 * 1024)
    (void)hipFuncSetAttribute(mlp_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->lds_mlp);
  if (P->lds_gmlp > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_gmlp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->lds_gmlp);
  return QEX_OK;
}

// The XC network's launch of iteration k (none for the closed-form functionals).
void launch_network(const Plan1d& P, int k, hipStream_t stream) {
  const Dev1& a = P.a;
  if (a.xc == QEX1D_XC_LOCAL_MLP) {
    Scope ps(qexprof::K_MLP, stream);
    if (P.mlp_mfma) {
      switch (a.nlay - 1) {
        case 1: hipLaunchKernelGGL(k_mlp_mfma<1>, P.gmlp_mfma, dim3(NT), P.lds_mlp, stream, a, k); break;
        case 2: hipLaunchKernelGGL(k_mlp_mfma<2>, P.gmlp_mfma, dim3(NT), P.lds_mlp, stream, a, k); break;
        case 3: hipLaunchKernelGGL(k_mlp_mfma<3>, P.gmlp_mfma, dim3(NT), P.lds_mlp, stream, a, k); break;
        default: hipLaunchKernelGGL(k_mlp_mfma<4>, P.gmlp_mfma, dim3(NT), P.lds_mlp, stream, a, k); break;
      }
    } else if (P.mlp_lpp == 8) {
      hipLaunchKernelGGL(k_mlp<8>, P.gmlp, dim3(NT), P.lds_mlp, stream, a, k);
    } else {
      hipLaunchKernelGGL(k_mlp<64>, P.gmlp, dim3(NT), P.lds_mlp, stream, a, k);
    }
  }
  if (a.xc == QEX1D_XC_GLOBAL_MLP) {
    Scope ps(qexprof::K_MLP, stream);
    hipLaunchKernelGGL(k_gmlp, dim3(a.B), dim3(NT), P.lds_gmlp, stream, a, k);
  }
}

// One SCF iteration's launches (the XC network, potentials and deferred energies, then unless
// energy_only the eigensolve and density/mixing); in slot mode every slot runs its own iteration.
void launch_iteration(const Plan1d& P, int k, double alpha_k, bool energy_only, hipStream_t stream) {
  const Dev1& a = P.a;
  launch_network(P, k, stream);
  {
    Scope ps(qexprof::K_POTENTIAL, stream);
    hipLaunchKernelGGL(k_potential, P.gpot, dim3(NT), P.lds_pot, stream, a, k);
  }
  if (energy_only) return;
  {
    Scope ps(qexprof::K_EIGEN, stream);
    hipLaunchKernelGGL(k_eigen, P.geig, dim3(NT), P.lds_eig, stream, P.e);
  }
  {
    Scope ps(qexprof::K_DENSITY, stream);
    hipLaunchKernelGGL(k_density, dim3(a.B), dim3(NT), (size_t)a.N * 8, stream, a, k, alpha_k);
  }
}
}  // namespace

extern "C" {

int qex1d_kohn_sham(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out, void* workspace,
                    size_t workspace_bytes, void* stream_) {
  Plan1d P;
  const int rc = plan1d(p, in, out, workspace, workspace_bytes, &P);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  const Dev1& a = P.a;
  hipLaunchKernelGGL(k_init, dim3((a.B + 255) / 256), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(k_pairs, dim3(1), dim3(PAIR_NT), 0, stream, a.B, a.K, a.ne, P.e.pairs, nullptr, a.T);
  double alpha_k = p->alpha;
  for (int k = 0; k <= a.T; ++k) {
    launch_iteration(P, k, alpha_k, k == a.T, stream);  // the last pass only evaluates the final energies
    alpha_k *= p->alpha_decay;
  }
  {
    Scope ps(qexprof::K_FINALIZE, stream);
    hipLaunchKernelGGL(k_finalize, dim3(a.B), dim3(64), 0, stream, a);
  }
  return check_launch("qex1d_kohn_sham");
}

int qex1d_slots_step(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out, int32_t* slot_iteration,
                     const int32_t* admit, void* workspace, size_t workspace_bytes, void* stream_) {
  if (!slot_iteration || !admit) return fail(QEX_E_BADARG, "slot mode needs slot_iteration and admit arrays");
  Plan1d P;
  const int rc = plan1d(p, in, out, workspace, workspace_bytes, &P);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  Dev1& a = P.a;
  a.kslot = slot_iteration;
  a.admit = admit;
  a.alpha0 = p->alpha;
  a.decay = p->alpha_decay;
  hipLaunchKernelGGL(k_slot_begin, dim3((a.B + 255) / 256), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(k_pairs, dim3(1), dim3(PAIR_NT), 0, stream, a.B, a.K, a.ne, P.e.pairs, a.kslot, a.T);
  launch_iteration(P, 0, 0.0, false, stream);
  {
    Scope ps(qexprof::K_FINALIZE, stream);
    hipLaunchKernelGGL(k_slot_end, dim3(a.B), dim3(64), 0, stream, a);
  }
  return check_launch("qex1d_slots_step");
}

int qex1d_kohn_sham_step(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out, int32_t iteration,
                         void* workspace, size_t workspace_bytes, void* stream_) {
  Plan1d P;
  const int rc = plan1d(p, in, out, workspace, workspace_bytes, &P);
  if (rc) return rc;
  const Dev1& a = P.a;
  if (iteration < 0 || iteration > a.T) return fail(QEX_E_BADARG, "iteration must be in [0, %d]", a.T);
  hipStream_t stream = (hipStream_t)stream_;
  if (iteration == 0) {
    hipLaunchKernelGGL(k_init, dim3((a.B + 255) / 256), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(k_pairs, dim3(1), dim3(PAIR_NT), 0, stream, a.B, a.K, a.ne, P.e.pairs, nullptr, a.T);
  }
  double alpha_k = p->alpha;  // the product qex1d_kohn_sham's loop forms
  for (int i = 0; i < iteration; ++i) alpha_k *= p->alpha_decay;
  launch_iteration(P, iteration, alpha_k, iteration == a.T, stream);
  if (iteration == a.T) {
    Scope ps(qexprof::K_FINALIZE, stream);
    hipLaunchKernelGGL(k_finalize, dim3(a.B), dim3(64), 0, stream, a);
  }
  return check_launch("qex1d_kohn_sham_step");
}

int qex1d_external_xc_densities(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out,
                                int32_t iteration, void* workspace, size_t workspace_bytes, double* density_in,
                                double* density_out, void* stream_) {
  Plan1d P;
  const int rc = plan1d(p, in, out, workspace, workspace_bytes, &P);
  if (rc) return rc;
  const Dev1& a = P.a;
  if (iteration < 0 || iteration > a.T) return fail(QEX_E_BADARG, "iteration must be in [0, %d]", a.T);
  if (!density_in || !density_out) return fail(QEX_E_BADARG, "null density_in or density_out");
  hipLaunchKernelGGL(k_xc_gather, dim3(a.B), dim3(NT), 0, (hipStream_t)stream_, a, (int)iteration, density_in,
                     density_out);
  return check_launch("qex1d_external_xc_densities");
}

size_t qex1d_noninteracting_workspace_bytes(int32_t batch, int32_t num_grids, int32_t max_electrons) {
  if (batch < 1 || num_grids < 5 || num_grids > QEX1D_MAX_GRIDS || max_electrons < 1 ||
      max_electrons > QEX1D_MAX_ELECTRONS)
    return 0;
  const size_t K = max_electrons / 2 + 1;
  return align256((size_t)batch * K * num_grids * 8) + align256((size_t)batch * K * 8) +
         align256((size_t)batch * K * 4);
}

int qex1d_solve_noninteracting(int32_t batch, int32_t num_grids, double dx, int32_t max_electrons,
                               const double* potential, const int32_t* num_electrons, double* density,
                               double* total_eigen_energies, double* gap, double* eigenvalues, double* wavefunctions,
                               void* workspace, size_t workspace_bytes, void* stream_) {
  const size_t need = qex1d_noninteracting_workspace_bytes(batch, num_grids, max_electrons);
  if (need == 0) return fail(QEX_E_BADARG, "bad sizes (batch %d, num_grids %d, max_electrons %d)", batch, num_grids,
                             max_electrons);
  if (!(dx > 0.0)) return fail(QEX_E_BADARG, "dx must be positive");
  if (max_electrons / 2 + 1 > num_grids) return fail(QEX_E_BADARG, "too many electrons for the grid");
  if (!potential || !num_electrons || !density) return fail(QEX_E_BADARG, "missing array");
  if (!workspace || workspace_bytes < need)
    return fail(QEX_E_WORKSPACE, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
  hipStream_t stream = (hipStream_t)stream_;
  const int K = max_electrons / 2 + 1;
  EigArgs e{};
  e.B = batch;
  e.N = num_grids;
  e.K = K;
  kinetic(dx, &e.Td, &e.Tb, &e.Tc);
  e.pot = potential;
  e.ne = num_electrons;
  e.st = nullptr;
  e.vec = (double*)workspace;
  e.eig = (double*)((char*)workspace + align256((size_t)batch * K * num_grids * 8));
  e.pairs = (int*)((char*)e.eig + align256((size_t)batch * K * 8));
  hipLaunchKernelGGL(k_pairs, dim3(1), dim3(PAIR_NT), 0, stream, batch, K, num_electrons, e.pairs, nullptr, 1);
  if (eigen_lds_bytes(num_grids) > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_eigen, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)eigen_lds_bytes(num_grids));
  hipLaunchKernelGGL(k_eigen, dim3(K * batch), dim3(NT), eigen_lds_bytes(num_grids), stream, e);
  hipLaunchKernelGGL(k_nonint_density, dim3(batch), dim3(NT), (size_t)num_grids * 8, stream, num_grids, K, dx, e.Td,
                     e.Tb, e.Tc, potential, num_electrons, e.vec, density, total_eigen_energies, gap, eigenvalues,
                     wavefunctions);
  return check_launch("qex1d_solve_noninteracting");
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Allocation-free non-interacting solve: the same kernels as qex1d_solve_noninteracting, with the
// workspace in a fixed device array of QEX1D_NOALLOC_SCRATCH_BYTES.  Each GPU holds its own copy;
// its address is looked up once per device (hipGetSymbolAddress) and kept here.  Calls that overlap
// on one GPU share the array and race; the kernels overwrite every scratch byte they read
// (cold-start inverse iteration, no state carried between calls), so nothing is cleared per call.
// ---------------------------------------------------------------------------------------------
namespace {
__device__ __align__(256) unsigned char scratch[QEX1D_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// device address of this GPU's scratch array (the host-side symbol is only a stand-in)
int scratch_for_current_device(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(QEX_E_LAUNCH, "no current HIP device (or device index >= %d)", kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr)
      return fail(QEX_E_LAUNCH, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(err));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return QEX_OK;
}
}  // namespace

extern "C" {

int qex1d_solve_noninteracting_noalloc(int32_t batch, int32_t num_grids, double dx, int32_t max_electrons,
                                       const double* potential, const int32_t* num_electrons, double* density,
                                       double* total_eigen_energies, double* gap, double* eigenvalues,
                                       double* wavefunctions, void* stream) {
  const size_t need = qex1d_noninteracting_workspace_bytes(batch, num_grids, max_electrons);
  if (need == 0) return fail(QEX_E_BADARG, "bad sizes (batch %d, num_grids %d, max_electrons %d)", batch, num_grids,
                             max_electrons);
  if (need > (size_t)QEX1D_NOALLOC_SCRATCH_BYTES)
    return fail(QEX_E_WORKSPACE,
                "batch %d x %d grid points x %d orbitals needs %zu scratch bytes; the allocation-free entry point "
                "holds %zu (use qex1d_solve_noninteracting with a workspace, or split the batch)",
                batch, num_grids, max_electrons / 2 + 1, need, (size_t)QEX1D_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != QEX_OK) return rc;
  return qex1d_solve_noninteracting(batch, num_grids, dx, max_electrons, potential, num_electrons, density,
                                    total_eigen_energies, gap, eigenvalues, wavefunctions, ws, need, stream);
}

int qex1d_external_potential(int32_t batch, int32_t num_grids, const double* grids, int32_t num_nuclei,
                             const double* locations, const double* nuclear_charges, int32_t interaction, double p0,
                             double p1, double* out, void* stream_) {
  if (batch < 1 || num_grids < 1 || num_nuclei < 0) return fail(QEX_E_BADARG, "bad sizes");
  if (!grids || !out || (num_nuclei > 0 && (!locations || !nuclear_charges))) return fail(QEX_E_BADARG, "null array");
  if (interaction != QEX_INTERACTION_EXPONENTIAL && interaction != QEX_INTERACTION_SOFT)
    return fail(QEX_E_BADARG, "bad interaction kind");
  const int total = batch * num_grids;
  hipLaunchKernelGGL(k_external, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream_, batch, num_grids,
                     grids, num_nuclei, locations, nuclear_charges, interaction, p0, p1, out);
  return check_launch("qex1d_external_potential");
}

int qex1d_hartree_potential(int32_t batch, int32_t num_grids, double dx, const double* table, const double* density,
                            double* out, void* stream_) {
  if (batch < 1 || num_grids < 1 || num_grids > QEX1D_MAX_GRIDS) return fail(QEX_E_BADARG, "bad sizes");
  if (!table || !density || !out) return fail(QEX_E_BADARG, "null array");
  hipLaunchKernelGGL(k_hartree, dim3((num_grids + NT - 1) / NT, batch), dim3(NT), (size_t)2 * num_grids * 8,
                     (hipStream_t)stream_, num_grids, dx, table, density, out);
  return check_launch("qex1d_hartree_potential");
}

}  // extern "C"

namespace {
// Workspace of qex1d_xc_evaluate: the per-system status words the network kernels read and the
// network outputs (per-point values, input gradients, scalars).
struct XcLayout {
  size_t st, mlp_e, mlp_G, mlp_eout, total;
};

// p with the SCF-only fields set to accepted values: the evaluation reads the functional and sizes.
qex1d_problem xc_problem(const qex1d_problem* p) {
  qex1d_problem q = *p;
  q.num_iterations = 1;
  q.max_electrons = 2;
  q.num_mixing_iterations = 1;
  q.loop = QEX1D_LOOP_PYTHON;
  q.reflect = 0;
  return q;
}

XcLayout layout_xc(const qex1d_problem* p) {
  const size_t B = p->batch, N = p->num_grids;
  const bool mlp = p->xc_kind == QEX1D_XC_LOCAL_MLP || p->xc_kind == QEX1D_XC_GLOBAL_MLP;
  const size_t W = p->xc_kind == QEX1D_XC_LOCAL_MLP ? p->mlp_widths[0] : 0;
  XcLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  L.st = take(B * 4 * 4);
  L.mlp_e = take(mlp ? B * N * 8 : 0);
  L.mlp_G = take(mlp ? B * N * W * 8 : 0);
  L.mlp_eout = take(mlp ? B * (N > 2 ? N : 2) * 8 : 0);
  L.total = off;
  return L;
}
}  // namespace

extern "C" {

size_t qex1d_xc_workspace_bytes(const qex1d_problem* p) {
  if (!p) return 0;
  const qex1d_problem q = xc_problem(p);
  if (validate1d(&q) != QEX_OK) return 0;
  return layout_xc(&q).total;
}

int qex1d_xc_evaluate(const qex1d_problem* p, const double* xc_params, const double* density, double* energy_density,
                      double* xc_energy, double* xc_potential, void* workspace, size_t workspace_bytes,
                      void* stream_) {
  if (!p) return fail(QEX_E_BADARG, "null problem");
  const qex1d_problem q = xc_problem(p);
  int rc = validate1d(&q);
  if (rc) return rc;
  if (!density) return fail(QEX_E_BADARG, "null density");
  if (xc_is_external(q.xc_kind)) return fail(QEX_E_UNSUPPORTED, "EXTERNAL functionals are evaluated by the caller");
  const bool mlp = q.xc_kind == QEX1D_XC_LOCAL_MLP || q.xc_kind == QEX1D_XC_GLOBAL_MLP;
  if (mlp && !xc_params) return fail(QEX_E_BADARG, "MLP functionals need xc_params");
  const XcLayout L = layout_xc(&q);
  if (!workspace || workspace_bytes < L.total)
    return fail(QEX_E_WORKSPACE, "workspace too small: need %zu bytes, got %zu", L.total, workspace_bytes);
  char* ws = (char*)workspace;
  Plan1d P{};
  Dev1& a = P.a;
  a.B = q.batch;
  a.N = q.num_grids;
  a.T = 1;
  a.K = 2;
  a.M = 1;
  a.nTiles = (a.N + TILE - 1) / TILE;
  a.xc = q.xc_kind;
  a.dx = q.dx;
  a.xc_c = q.xc_c;
  a.xc_p = q.xc_p;
  a.xc_eps = q.xc_eps;
  fill_network(&q, a);
  a.n0 = density;
  a.mlp = xc_params;
  a.st = (int*)(ws + L.st);
  a.mlp_e = (double*)(ws + L.mlp_e);
  a.mlp_G = (double*)(ws + L.mlp_G);
  a.mlp_eout = (double*)(ws + L.mlp_eout);
  rc = plan_network(&P);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_;
  // status words: not frozen, no deferred energy, so the network kernels evaluate iteration 0 on n0
  hipLaunchKernelGGL(k_init, dim3((a.B + 255) / 256), dim3(256), 0, stream, a);
  launch_network(P, 0, stream);
  hipLaunchKernelGGL(k_xc_eval, dim3(a.B), dim3(NT), 0, stream, a, energy_density, xc_potential, xc_energy);
  return check_launch("qex1d_xc_evaluate");
}

}  // extern "C"
