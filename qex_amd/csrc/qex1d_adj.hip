// qex1d_adj.hip -- reverse pass of the batched 1D Kohn-Sham scan loop with respect to the
// parameters of a windowed LocalMLP XC functional (the gradient that jax.grad takes through
// jit_scf.kohn_sham in qedft/train/od/train.py:250-323).
//
// Given the forward outputs of qex1d_kohn_sham (stored densities, v_H, v_xc and the orbitals of
// every iteration) and cotangents of the stored energies and densities, iteration k is reversed
// as (n_k: input density, n_out: unmixed output, psi_j: occupied orbitals, 2-norm 1):
//   k_adj_mix        nbar_k, dbar (the mixing n_k+1 = n_k + alpha_k mean(last m differences))
//   energy           nob += c (-v_KS + v_H(n_out) + v_xc(n_out) + v_ext), thetabar += c dF(n_out)/dtheta,
//                    c = dE_k dx (the eigenvalue sum and -dx v_KS.n_out cancel at fixed n_out)
//   k_adj_eig        vbar = sum_j occ_j (2/dx) psi_j * y_j with y_j = -(H - eps_j)^-1 Q (nob psi_j)
//                    (Q projects out the occupied orbitals; a banded LU with partial pivoting in
//                    LDS, then Q again) minus the pairs of occupied orbitals with unequal occupations
//   Hartree + XC     nbar += dx U vbar + H_F vbar, thetabar += d(vbar . grad F)/dtheta, F = sum e(n) n
// The MLP terms come from reverse mode over the network's forward and forward-tangent passes, one
// wavefront per point with one lane per hidden unit; each wavefront accumulates its parameter
// partials in registers, the waves of a workgroup combine them in LDS in a fixed order and add the
// sum to the workgroup's own slot, and the slots are summed in a fixed order.
// A GLOBAL_MLP functional (amplitude encoded: E_xc = dx f(n), v_xc = dx grad f) takes the same
// steps with F = f in the energy and F = dx f in the potential, through k_adj_gmlp (one workgroup
// per system); under reflection its functional sees the unflipped densities (the forward flips only
// the density passed on), so only the mixing cotangents are flipped.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "prof.h"
#include "qex.h"

namespace {

constexpr int NT = 256;
constexpr int ADJ_SLOTS = 1024;  // wavefronts of the MLP adjoint launch
constexpr int ADJ_WG = ADJ_SLOTS / (NT / 64);  // its workgroups = parameter-partial slots
constexpr int HMAX = 64;        // hidden width (one lane per unit)
constexpr int WMAXW = QEX1D_MAX_MLP_WINDOW;
thread_local char g_erra[512] = "";

int faila(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_erra, sizeof(g_erra), fmt, ap);
  va_end(ap);
  return code;
}

__device__ __forceinline__ double wsum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

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

// activation and its first two derivatives
__device__ void act3(int act, double z, double* a, double* d1, double* d2) {
  switch (act) {
    case QEX_ACT_SILU: {
      const double s = 1.0 / (1.0 + exp(-z));
      *a = z * s;
      *d1 = s * (1.0 + z * (1.0 - s));
      *d2 = s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s));
      break;
    }
    case QEX_ACT_TANH: {
      const double t = tanh(z);
      *a = t;
      *d1 = 1.0 - t * t;
      *d2 = -2.0 * t * (1.0 - t * t);
      break;
    }
    case QEX_ACT_SOFTPLUS: {
      const double s = 1.0 / (1.0 + exp(-z));
      *a = z > 0.0 ? z + log1p(exp(-z)) : log1p(exp(z));
      *d1 = s;
      *d2 = s * (1.0 - s);
      break;
    }
    case QEX_ACT_SIGMOID: {
      const double s = 1.0 / (1.0 + exp(-z));
      *a = s;
      *d1 = s * (1.0 - s);
      *d2 = s * (1.0 - s) * (1.0 - 2.0 * s);
      break;
    }
    case QEX_ACT_RELU:
      *a = z > 0.0 ? z : 0.0;
      *d1 = z > 0.0 ? 1.0 : 0.0;
      *d2 = 0.0;
      break;
    default: {  // GELU, tanh form
      const double k = 0.7978845608028654, c = 0.044715;
      const double t = tanh(k * (z + c * z * z * z));
      const double du = k * (1.0 + 3.0 * c * z * z);
      const double s2 = 1.0 - t * t;
      *a = 0.5 * z * (1.0 + t);
      *d1 = 0.5 * (1.0 + t) + 0.5 * z * s2 * du;
      *d2 = s2 * du + 0.5 * z * (-2.0 * t * s2 * du * du + s2 * k * 6.0 * c * z);
      break;
    }
  }
}

// output transform of the network (QEX1D_MLP_OUTPUT_*) and its first two derivatives
__device__ __forceinline__ double out3(int kind, double y, double* d1, double* d2) {
  if (kind == QEX1D_MLP_OUTPUT_NEGATIVITY) {  // -swish(y)
    const double s = 1.0 / (1.0 + exp(-y));
    *d1 = -(s * (1.0 + y * (1.0 - s)));
    *d2 = -(s * (1.0 - s) * (2.0 + y * (1.0 - 2.0 * s)));
    return -y * s;
  }
  *d1 = 1.0;
  *d2 = 0.0;
  return y;
}

struct Adj {
  int B, N, T, K, M, W, H1, H2, NL, act, P, mlpout, reflect, sgs;
  const int* win;  // [B][2] reflection window (reflect)
  double dx, norm;
  int w1, b1, w2, b2, w3, b3;  // parameter offsets (w2/b2 unused when NL == 1)
  const double* table;
  const double* vext;
  const double* n0;
  const int* ne;
  const double* mlp;
  const double* dens;  // [B][T][N] stored (mixed) densities
  const double* vh;    // [B][T][N] v_H(n_k)
  const double* vxc;   // [B][T][N] v_xc(n_k)
  const double* orb;   // [B][T][K][N]
  const double* dE;    // [B][T] or null
  const double* dn;    // [B][T][N] or null
  // workspace
  double* nbar;  // [B][N]
  double* nob;   // [B][N]
  double* nout;  // [B][N]
  double* vbar;  // [B][N]
  double* tmp;   // [B][N]
  double* dbar;  // [B][T][N]
  double* xbar;  // [B][N][W]
  double* fv;    // [B][N]
  double* fdv;   // [B][N]
  double* thp;   // [ADJ_WG][P]
};

__device__ __forceinline__ const double* input_density(const Adj& a, int s, int k) {
  return k == 0 ? a.n0 + (size_t)s * a.N : a.dens + ((size_t)s * a.T + k - 1) * a.N;
}

__device__ __forceinline__ double occupation(int ne, int j) { return j < (ne + 1) / 2 ? (double)min(2, ne - 2 * j) : 0.0; }

__global__ void k_zero_a(double* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

// mixing and difference cotangents; one workgroup per system
__global__ __launch_bounds__(NT) void k_adj_mix(Adj a, int k, double alpha_k) {
  const int s = blockIdx.x, N = a.N, T = a.T;
  const int cnt = min(a.M, k + 1);
  const double wgt = alpha_k / (double)cnt;
  const bool stopped = k <= a.sgs;  // the state of this iteration is a constant: no cotangent reaches it
  for (int t = threadIdx.x; t < N; t += NT) {
    double g = stopped ? 0.0 : a.nbar[(size_t)s * N + t];
    if (a.dn && !stopped) g += a.dn[((size_t)s * T + k) * N + t];
    for (int j = k - cnt + 1; j <= k; ++j) a.dbar[((size_t)s * T + j) * N + t] += wgt * g;
    const double dk = a.dbar[((size_t)s * T + k) * N + t];
    a.nob[(size_t)s * N + t] = dk;
    a.nbar[(size_t)s * N + t] = g - dk;
    // unmixed output density of iteration k from its orbitals
    double no = 0.0;
    const int ne = a.ne[s];
    for (int j = 0; j < (ne + 1) / 2; ++j) {
      const double x = a.orb[(((size_t)s * T + k) * a.K + j) * N + t];
      no = fma(occupation(ne, j), x * x, no);
    }
    a.nout[(size_t)s * N + t] = no / a.dx;
  }
}

// x[s] <- S x[s] in place: the flip-and-average map inside the system's reflection window (S is
// symmetric and idempotent, so it is also its own adjoint)
__global__ void k_sym_inplace(Adj a, double* x) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.B * a.N) return;
  const int s = (int)(idx / a.N), i = (int)(idx % a.N);
  const int lo = a.win[2 * s], hi = a.win[2 * s + 1];
  const int m = (i >= lo && i <= hi) ? lo + hi - i : i;
  if (i < m) {
    double* r = x + (size_t)s * a.N;
    const double v = 0.5 * (r[i] + r[m]);
    r[i] = v;
    r[m] = v;
  }
}

// dst[s] += S c[s] + d[s] (d optional)
__global__ void k_sym_add(Adj a, double* dst, const double* c, const double* d) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.B * a.N) return;
  const int s = (int)(idx / a.N), i = (int)(idx % a.N);
  const int lo = a.win[2 * s], hi = a.win[2 * s + 1];
  const int m = (i >= lo && i <= hi) ? lo + hi - i : i;
  dst[idx] += 0.5 * (c[idx] + c[(size_t)s * a.N + m]) + (d ? d[idx] : 0.0);
}

// dst[s] = S src[s]
__global__ void k_sym_copy(Adj a, double* dst, const double* src) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.B * a.N) return;
  const int s = (int)(idx / a.N), i = (int)(idx % a.N);
  const int lo = a.win[2 * s], hi = a.win[2 * s + 1];
  const int m = (i >= lo && i <= hi) ? lo + hi - i : i;
  dst[idx] = 0.5 * (src[idx] + src[(size_t)s * a.N + m]);
}

// out[s] = dx * U in[s] (Toeplitz interaction matrix from its first row)
__global__ __launch_bounds__(NT) void k_toeplitz(int N, double dx, const double* table, const double* in,
                                                 double* out) {
  extern __shared__ double sm[];
  const int s = blockIdx.y;
  double* tab = sm;
  double* x = sm + N;
  for (int t = threadIdx.x; t < N; t += NT) {
    tab[t] = table[t];
    x[t] = in[(size_t)s * N + t];
  }
  __syncthreads();
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  for (int j = 0; j < N; ++j) acc = fma(tab[abs(i - j)], x[j], acc);
  out[(size_t)s * N + i] = dx * acc;
}

// MLP second-order adjoint.  Per point p = (s, i): x = window(xin[s])_i / norm, tangent
// xt = window(tin[s])_i / norm (zero without tin), fbar = sf[s] * F1[s][i], fdbar = F2 ? F2[s][i] : 0.
// Outputs per point: xbar (dPhi/dx), f, fdot; parameter gradient of Phi = fbar f + fdbar fdot
// accumulated into the wavefront's slot.
template <int NL>
__global__ __launch_bounds__(NT, 1) void k_adj_mlp(Adj a, const double* xin, const double* tin, const double* F1,
                                                   const double* sf, const double* F2) {
  extern __shared__ double sm[];
  const int W = a.W, H1 = a.H1, H2 = a.H2, HL = NL == 2 ? H2 : H1;
  // weights in LDS
  double* sW1 = sm;                       // W x H1
  double* sb1 = sW1 + W * H1;             // H1
  double* sW2 = sb1 + H1;                 // H1 x H2 (NL == 2)
  double* sb2 = sW2 + (NL == 2 ? H1 * H2 : 0);
  double* sW3 = sb2 + (NL == 2 ? H2 : 0);  // HL
  double* sW1t = sW3 + HL;                 // H1 x W (transposed W1)
  double* sW2t = sW1t + W * H1;            // H2 x H1 (transposed W2, NL == 2)
  double* wbuf = sW2t + (NL == 2 ? H1 * H2 : 0);  // per wave: 6 x 64
  for (int t = threadIdx.x; t < W * H1; t += NT) {
    sW1[t] = a.mlp[a.w1 + t];
    sW1t[(t % H1) * W + t / H1] = a.mlp[a.w1 + t];
  }
  for (int t = threadIdx.x; t < H1; t += NT) sb1[t] = a.mlp[a.b1 + t];
  if (NL == 2) {
    for (int t = threadIdx.x; t < H1 * H2; t += NT) {
      sW2[t] = a.mlp[a.w2 + t];
      sW2t[(t % H2) * H1 + t / H2] = a.mlp[a.w2 + t];
    }
    for (int t = threadIdx.x; t < H2; t += NT) sb2[t] = a.mlp[a.b2 + t];
  }
  for (int t = threadIdx.x; t < HL; t += NT) sW3[t] = a.mlp[a.w3 + t];
  __syncthreads();
  const double b3 = a.mlp[a.b3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int slot = blockIdx.x * (NT / 64) + wv;
  double* sx = wbuf + wv * 6 * 64;  // window
  double* sxt = sx + 64;            // tangent window
  double* sa = sxt + 64;            // a1 / zb (reused)
  double* sat = sa + 64;            // at1 / zdb
  double* sz = sat + 64;            // zb1
  const int h = W / 2, N = a.N;
  const double inorm = 1.0 / a.norm;
  // parameter partials of this lane (column `lane` of each weight matrix)
  double gW1[WMAXW], gW2[HMAX];
#pragma unroll
  for (int o = 0; o < WMAXW; ++o) gW1[o] = 0.0;
#pragma unroll
  for (int v = 0; v < HMAX; ++v) gW2[v] = 0.0;
  double gb1 = 0.0, gb2 = 0.0, gW3 = 0.0, gb3 = 0.0;
  const long npts = (long)a.B * N;
  for (long p = slot; p < npts; p += ADJ_SLOTS) {
    const int s = (int)(p / N), i = (int)(p % N);
    const double* xs = xin + (size_t)s * N;
    if (lane < W) {
      const int j = i + lane - h;
      sx[lane] = (j >= 0 && j < N) ? xs[j] * inorm : 0.0;
      sxt[lane] = (tin && j >= 0 && j < N) ? tin[(size_t)s * N + j] * inorm : 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    const double fbar = sf[s] * F1[(size_t)s * N + i];
    const double fdbar = F2 ? F2[(size_t)s * N + i] : 0.0;
    // layer 1 (lane u)
    double z1 = 0.0, zt1 = 0.0, a1 = 0.0, d11 = 0.0, d21 = 0.0, at1 = 0.0;
    if (lane < H1) {
      z1 = sb1[lane];
      for (int o = 0; o < W; ++o) {
        z1 = fma(sx[o], sW1[o * H1 + lane], z1);
        zt1 = fma(sxt[o], sW1[o * H1 + lane], zt1);
      }
      act3(a.act, z1, &a1, &d11, &d21);
      at1 = d11 * zt1;
    }
    double aL = a1, atL = at1, d1L = d11, d2L = d21, ztL = zt1;
    double zt2 = 0.0, a2 = 0.0, d12 = 0.0, d22 = 0.0, at2 = 0.0;
    if (NL == 2) {
      if (lane < H1) {
        sa[lane] = a1;
        sat[lane] = at1;
      }
      __builtin_amdgcn_wave_barrier();
      if (lane < H2) {
        double z2 = sb2[lane];
        for (int v = 0; v < H1; ++v) {
          z2 = fma(sa[v], sW2[v * H2 + lane], z2);
          zt2 = fma(sat[v], sW2[v * H2 + lane], zt2);
        }
        act3(a.act, z2, &a2, &d12, &d22);
        at2 = d12 * zt2;
      }
      aL = a2;
      atL = at2;
      d1L = d12;
      d2L = d22;
      ztL = zt2;
    }
    const double w3 = lane < HL ? sW3[lane] : 0.0;
    // raw output y and tangent yd, then the output transform f = g(y), fd = g'(y) yd
    const double y = wsum(aL * w3) + b3;
    const double yd = wsum(atL * w3);
    double g1, g2;
    const double f = out3(a.mlpout, y, &g1, &g2);
    const double fd = g1 * yd;
    // reverse through the transform: ybar = fbar g' + fdbar g'' yd, ydbar = fdbar g'
    const double ybar = fma(fbar, g1, fdbar * g2 * yd), ydbar = fdbar * g1;
    // reverse: output layer
    if (lane < HL) gW3 = fma(ybar, aL, fma(ydbar, atL, gW3));
    if (lane == 0) gb3 += ybar;
    const double hb = ybar * w3, hdb = ydbar * w3;
    double zdb = hdb * d1L;
    double zb = fma(hb, d1L, hdb * d2L * ztL);
    if (lane >= HL) {
      zdb = 0.0;
      zb = 0.0;
    }
    double zb1 = zb, zdb1 = zdb;
    if (NL == 2) {
      // W2 column `lane` (unit u = lane of layer 2): += a1_v zb_u + at1_v zdb_u
#pragma unroll
      for (int v = 0; v < HMAX; ++v)
        if (v < H1) gW2[v] = fma(sa[v], zb, fma(sat[v], zdb, gW2[v]));
      gb2 += zb;
      __builtin_amdgcn_wave_barrier();
      sa[lane] = zb;  // a1 / at1 no longer needed
      sat[lane] = zdb;
      __builtin_amdgcn_wave_barrier();
      double hb1 = 0.0, hdb1 = 0.0;
      if (lane < H1) {
        for (int u = 0; u < H2; ++u) {
          hb1 = fma(sa[u], sW2t[u * H1 + lane], hb1);
          hdb1 = fma(sat[u], sW2t[u * H1 + lane], hdb1);
        }
      }
      zdb1 = hdb1 * d11;
      zb1 = fma(hb1, d11, hdb1 * d21 * zt1);
      if (lane >= H1) {
        zdb1 = 0.0;
        zb1 = 0.0;
      }
    }
    // W1 column `lane`: += x_o zb1 + xt_o zdb1
#pragma unroll
    for (int o = 0; o < WMAXW; ++o)
      if (o < W) gW1[o] = fma(sx[o], zb1, fma(sxt[o], zdb1, gW1[o]));
    gb1 += zb1;
    sz[lane] = zb1;
    __builtin_amdgcn_wave_barrier();
    if (lane < W) {
      double xb = 0.0;
      for (int u = 0; u < H1; ++u) xb = fma(sz[u], sW1t[u * W + lane], xb);
      a.xbar[(size_t)p * W + lane] = xb;
    }
    if (lane == 0) {
      a.fv[p] = f;
      a.fdv[p] = fd;
    }
    __builtin_amdgcn_wave_barrier();
  }
  // the workgroup's waves add their partials into LDS one after another (the weights are no longer
  // needed, and their area holds the P parameters), then the sum goes to the workgroup's slot (slots
  // are owned: no races; a fixed order)
  __syncthreads();
  double* r = sm;
  for (int t = threadIdx.x; t < a.P; t += NT) r[t] = 0.0;
  __syncthreads();
  for (int w = 0; w < NT / 64; ++w) {
    if (wv == w) {
      if (lane < H1) {
#pragma unroll
        for (int o = 0; o < WMAXW; ++o)
          if (o < W) r[a.w1 + o * H1 + lane] += gW1[o];
        r[a.b1 + lane] += gb1;
      }
      if (NL == 2 && lane < H2) {
#pragma unroll
        for (int v = 0; v < HMAX; ++v)
          if (v < H1) r[a.w2 + v * H2 + lane] += gW2[v];
        r[a.b2 + lane] += gb2;
      }
      if (lane < HL) r[a.w3 + lane] += gW3;
      if (lane == 0) r[a.b3] += gb3;
    }
    __syncthreads();
  }
  double* g = a.thp + (size_t)blockIdx.x * a.P;
  for (int t = threadIdx.x; t < a.P; t += NT) g[t] += r[t];
}

// Global-network second-order adjoint (GLOBAL_MLP: f(n) = g(MLP(n / norm)), one scalar per
// system).  Per system s: x = xin[s] / norm, tangent xt = tin[s] / norm (zero without tin),
// Phi = sf[s] f + fd fdot (sf null: 0; fdot = grad f . tin).  dst[s][i] += dPhi/dn_i (+ A2[s][i]);
// the parameter gradient of Phi is added to slot blockIdx.x.  A workgroup takes systems blockIdx.x,
// blockIdx.x + gridDim.x, ... in order, so each slot is written by one workgroup in a fixed order.
// Hidden widths <= 64: unit u of a layer is lane u of a wave; the input layer's sums over the N
// grid points are split over the 4 waves and combined in a fixed order.
template <int NL>
__global__ __launch_bounds__(NT) void k_adj_gmlp(Adj a, const double* xin, const double* tin, const double* sf,
                                                 double fd, const double* A2, double* dst) {
  extern __shared__ double sm[];
  constexpr int NW = NT / 64;
  const int N = a.N, H1 = a.H1, H2 = a.H2, HL = NL == 2 ? H2 : H1;
  const int t = threadIdx.x, u = t & 63, q = t >> 6;
  double* x = sm;                 // N
  double* xd = x + N;             // N
  double* part = xd + N;          // 2 x NW x 64
  double* h1 = part + 2 * NW * 64;  // activations, tangents, act', act'' and the two adjoints, 6 x 64
  double* h2 = h1 + 6 * 64;         // the same for layer 2
  double* sc = h2 + 6 * 64;         // scalars: fbar', fdbar'
  const double* W1 = a.mlp + a.w1;
  const double* W2 = a.mlp + a.w2;
  const double* W3 = a.mlp + a.w3;
  const double in = 1.0 / a.norm;
  double* g = a.thp + (size_t)blockIdx.x * a.P;
  for (int s = blockIdx.x; s < a.B; s += gridDim.x) {
    for (int i = t; i < N; i += NT) {
      x[i] = xin[(size_t)s * N + i] * in;
      xd[i] = tin ? tin[(size_t)s * N + i] * in : 0.0;
    }
    __syncthreads();
    // layer 1: z = b1 + W1^T x, zd = W1^T xd, the grid points split over the waves
    {
      double acc = 0.0, accd = 0.0;
      if (u < H1)
        for (int i = q; i < N; i += NW) {
          const double w = W1[(size_t)i * H1 + u];
          acc = fma(x[i], w, acc);
          accd = fma(xd[i], w, accd);
        }
      part[q * 64 + u] = acc;
      part[(NW + q) * 64 + u] = accd;
    }
    __syncthreads();
    if (t < H1) {
      double z = a.mlp[a.b1 + t], zd = 0.0;
      for (int w = 0; w < NW; ++w) {
        z += part[w * 64 + t];
        zd += part[(NW + w) * 64 + t];
      }
      double av, d1, d2;
      act3(a.act, z, &av, &d1, &d2);
      h1[t] = av;
      h1[64 + t] = d1 * zd;
      h1[128 + t] = d1;
      h1[192 + t] = d2;
      h1[256 + t] = zd;
    }
    __syncthreads();
    if (NL == 2 && t < H2) {
      double z = a.mlp[a.b2 + t], zd = 0.0;
      for (int v = 0; v < H1; ++v) {
        const double w = W2[v * H2 + t];
        z = fma(h1[v], w, z);
        zd = fma(h1[64 + v], w, zd);
      }
      double av, d1, d2;
      act3(a.act, z, &av, &d1, &d2);
      h2[t] = av;
      h2[64 + t] = d1 * zd;
      h2[128 + t] = d1;
      h2[192 + t] = d2;
      h2[256 + t] = zd;
    }
    __syncthreads();
    const double* hl = NL == 2 ? h2 : h1;
    if (t == 0) {
      double y = a.mlp[a.b3], yd = 0.0;
      for (int v = 0; v < HL; ++v) {
        y = fma(hl[v], W3[v], y);
        yd = fma(hl[64 + v], W3[v], yd);
      }
      double g1, g2;
      (void)out3(a.mlpout, y, &g1, &g2);
      const double fb = sf ? sf[s] : 0.0;
      sc[0] = fb * g1 + fd * g2 * yd;  // cotangent of y
      sc[1] = fd * g1;                 // cotangent of yd
      g[a.b3] += sc[0];
    }
    __syncthreads();
    const double yb = sc[0], ydb = sc[1];
    // last hidden layer: its adjoints (zb, zdb) from the output layer
    double* hL = NL == 2 ? h2 : h1;
    if (t < HL) {
      const double w = W3[t];
      g[a.w3 + t] += hL[t] * yb + hL[64 + t] * ydb;
      const double hb = yb * w, hdb = ydb * w;
      hL[320 + t] = hb * hL[128 + t] + hdb * hL[192 + t] * hL[256 + t];  // zb
      hL[256 + t] = hdb * hL[128 + t];                                 // zdb (zd no longer needed)
    }
    __syncthreads();
    if (NL == 2) {
      // W2 (H1 x H2): lane u = row, waves over the columns
      if (u < H1)
        for (int v = q; v < H2; v += NW)
          g[a.w2 + u * H2 + v] += h1[u] * h2[320 + v] + h1[64 + u] * h2[256 + v];
      if (t < H2) g[a.b2 + t] += h2[320 + t];
      __syncthreads();
      if (t < H1) {
        double hb = 0.0, hdb = 0.0;
        for (int v = 0; v < H2; ++v) {
          const double w = W2[t * H2 + v];
          hb = fma(w, h2[320 + v], hb);
          hdb = fma(w, h2[256 + v], hdb);
        }
        const double zb = hb * h1[128 + t] + hdb * h1[192 + t] * h1[256 + t];
        h1[256 + t] = hdb * h1[128 + t];
        h1[320 + t] = zb;
      }
      __syncthreads();
    }
    // layer 1: W1 (N x H1), b1, and the input cotangent
    if (u < H1)
      for (int i = q; i < N; i += NW) g[a.w1 + (size_t)i * H1 + u] += x[i] * h1[320 + u] + xd[i] * h1[256 + u];
    if (t < H1) g[a.b1 + t] += h1[320 + t];
    for (int i = t; i < N; i += NT) {
      double xb = 0.0;
      for (int v = 0; v < H1; ++v) xb = fma(W1[(size_t)i * H1 + v], h1[320 + v], xb);
      double r = xb * in;
      if (A2) r += A2[(size_t)s * N + i];
      dst[(size_t)s * N + i] += r;
    }
    __syncthreads();
  }
}

// dst[s][i] += sa[s] * (A1[s][i] + F[s][i]) + A2[s][i] + window^T(xbar)_i / norm (A1, A2 optional)
__global__ void k_adj_gather(Adj a, double* dst, const double* A1, const double* A2, const double* F,
                             const double* sa) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.B * a.N) return;
  const int s = (int)(idx / a.N), i = (int)(idx % a.N), N = a.N, W = a.W, h = W / 2;
  double acc = 0.0;
  for (int o = 0; o < W; ++o) {
    const int q = i - o + h;  // point whose window slot o reads n_i
    if (q >= 0 && q < N) acc += a.xbar[((size_t)s * N + q) * W + o];
  }
  double v = acc / a.norm;
  const double c = sa ? sa[s] : 1.0;
  double direct = F[idx];
  if (A1) direct += A1[idx];
  v = fma(c, direct, v);
  if (A2) v += A2[idx];
  dst[idx] += v;
}

// energy term: nob += c (-(v_H(n_k) + v_xc(n_k)) + v_H(n_out)) with c = dE dx (the v_xc(n_out) part
// is added by k_adj_gather from the MLP pass at n_out)
__global__ void k_adj_energy(Adj a, int k, const double* cs) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.B * a.N) return;
  const int s = (int)(idx / a.N), i = (int)(idx % a.N);
  const size_t kk = ((size_t)s * a.T + k) * a.N + i;
  a.nob[idx] += cs[s] * (a.tmp[idx] - a.vh[kk] - a.vxc[kk]);
}

__global__ void k_adj_scales(Adj a, int k, double* cs, double* ones) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.B) return;
  cs[s] = (a.dE && k > a.sgs) ? a.dE[(size_t)s * a.T + k] * a.dx : 0.0;
  ones[s] = 1.0;
}

// eigen adjoint, one workgroup per system.  LDS: H diagonal, r, y, vbar and the rows of U of the
// banded LU factorisation (row i holds columns i .. i + 4).
__global__ __launch_bounds__(NT) void k_adj_eig(Adj a, int k, double Td, double Tb, double Tc) {
  extern __shared__ double sm[];
  __shared__ double red[NT / 64];
  __shared__ double eps[QEX1D_MAX_ELECTRONS / 2 + 1];
  const int s = blockIdx.x, N = a.N, T = a.T;
  double* hd = sm;       // N
  double* r = hd + N;    // N
  double* y = r + N;     // N
  double* vb = y + N;    // N
  double* R = vb + N;    // N x 5 rows of U
  const int ne = a.ne[s], nocc = (ne + 1) / 2;
  const double* vxc = a.vxc + ((size_t)s * T + k) * N;
  const double* vh = a.vh + ((size_t)s * T + k) * N;
  const double* ve = a.vext + (size_t)s * N;
  const double* psi = a.orb + ((size_t)s * T + k) * a.K * N;
  const double* nob = a.nob + (size_t)s * N;
  for (int t = threadIdx.x; t < N; t += NT) {
    hd[t] = Td + ve[t] + vh[t] + vxc[t];
    vb[t] = 0.0;
  }
  __syncthreads();
  // eigenvalues as Rayleigh quotients of the stored orbitals
  for (int j = 0; j < nocc; ++j) {
    const double* pj = psi + (size_t)j * N;
    double q = 0.0;
    for (int t = threadIdx.x; t < N; t += NT) {
      double hx = hd[t] * pj[t];
      if (t >= 1) hx = fma(Tb, pj[t - 1], hx);
      if (t + 1 < N) hx = fma(Tb, pj[t + 1], hx);
      if (t >= 2) hx = fma(Tc, pj[t - 2], hx);
      if (t + 2 < N) hx = fma(Tc, pj[t + 2], hx);
      q = fma(pj[t], hx, q);
    }
    q = bsum(q, red);
    if (threadIdx.x == 0) eps[j] = q;
  }
  __syncthreads();
  __shared__ double hmax;
  if (threadIdx.x == 0) {
    double hn = 0.0;
    for (int t = 0; t < N; ++t) hn = fmax(hn, fabs(hd[t]));
    hmax = hn;
  }
  __syncthreads();
  // zero pivots are replaced by +-eps |H| (the solution's component along psi_j is projected out)
  const double tiny = 2.220446049250313e-16 * (hmax + 2.0 * fabs(Tb) + 2.0 * fabs(Tc));
  for (int j = 0; j < nocc; ++j) {
    const double* pj = psi + (size_t)j * N;
    const double oj = occupation(ne, j);
    for (int t = threadIdx.x; t < N; t += NT) r[t] = nob[t] * pj[t];
    __syncthreads();
    // rhs = Q r
    for (int t = threadIdx.x; t < N; t += NT) y[t] = r[t];
    __syncthreads();
    for (int b = 0; b < nocc; ++b) {
      const double* pb = psi + (size_t)b * N;
      double d = 0.0;
      for (int t = threadIdx.x; t < N; t += NT) d = fma(pb[t], r[t], d);
      d = bsum(d, red);
      for (int t = threadIdx.x; t < N; t += NT) y[t] = fma(-d, pb[t], y[t]);
      __syncthreads();
    }
    // (H - eps_j) x = y by Gaussian elimination with partial pivoting (bandwidth 2, fill to 4):
    // one lane keeps the three active rows (columns i .. i + 4) in registers, stores each finished
    // row of U (5 entries) and its right-hand side in LDS, then back-substitutes.
    const double ej = eps[j];
    if (threadIdx.x == 0) {
      auto band = [&](int row, int col) -> double {  // entry of H - eps_j (zero outside the matrix)
        if (row >= N || col < 0 || col >= N) return 0.0;
        const int d = col - row;
        return d == 0 ? hd[row] - ej : (d == 1 || d == -1) ? Tb : (d == 2 || d == -2) ? Tc : 0.0;
      };
      double r0[5], r1[5], r2[5];
      for (int c = 0; c < 5; ++c) {
        r0[c] = band(0, c);
        r1[c] = band(1, c);
        r2[c] = band(2, c);
      }
      double y0 = y[0], y1 = N > 1 ? y[1] : 0.0, y2 = N > 2 ? y[2] : 0.0;
      for (int i = 0; i < N; ++i) {
        // pivot among the (existing) rows i, i+1, i+2 in column i
        int pr = 0;
        double best = fabs(r0[0]);
        if (i + 1 < N && fabs(r1[0]) > best) {
          best = fabs(r1[0]);
          pr = 1;
        }
        if (i + 2 < N && fabs(r2[0]) > best) pr = 2;
        if (pr == 1) {
          for (int c = 0; c < 5; ++c) {
            const double t1 = r0[c];
            r0[c] = r1[c];
            r1[c] = t1;
          }
          const double ty = y0;
          y0 = y1;
          y1 = ty;
        } else if (pr == 2) {
          for (int c = 0; c < 5; ++c) {
            const double t1 = r0[c];
            r0[c] = r2[c];
            r2[c] = t1;
          }
          const double ty = y0;
          y0 = y2;
          y2 = ty;
        }
        if (fabs(r0[0]) < tiny) r0[0] = r0[0] < 0.0 ? -tiny : tiny;
        const double ip = 1.0 / r0[0];
        const double l1 = r1[0] * ip, l2 = r2[0] * ip;
        for (int c = 1; c < 5; ++c) {
          r1[c] = fma(-l1, r0[c], r1[c]);
          r2[c] = fma(-l2, r0[c], r2[c]);
        }
        y1 = fma(-l1, y0, y1);
        y2 = fma(-l2, y0, y2);
        for (int c = 0; c < 5; ++c) R[(size_t)i * 5 + c] = r0[c];
        y[i] = y0;
        // advance: rows i+1, i+2 shift one column left; row i+3 enters
        for (int c = 0; c < 4; ++c) {
          r0[c] = r1[c + 1];
          r1[c] = r2[c + 1];
        }
        r0[4] = 0.0;
        r1[4] = 0.0;
        for (int c = 0; c < 5; ++c) r2[c] = band(i + 3, i + 1 + c);
        y0 = y1;
        y1 = y2;
        y2 = i + 3 < N ? y[i + 3] : 0.0;
      }
      for (int i = N - 1; i >= 0; --i) {
        double v = y[i];
        for (int c = 1; c < 5 && i + c < N; ++c) v = fma(-R[(size_t)i * 5 + c], y[i + c], v);
        y[i] = v / R[(size_t)i * 5];
      }
    }
    __syncthreads();
    // x <- Q x; y_j = -x - sum over occupied b with occ_b != occ_j of psi_b (psi_b . r) / (eps_b - eps_j)
    for (int b = 0; b < nocc; ++b) {
      const double* pb = psi + (size_t)b * N;
      double d = 0.0, dr = 0.0;
      for (int t = threadIdx.x; t < N; t += NT) {
        d = fma(pb[t], y[t], d);
        dr = fma(pb[t], r[t], dr);
      }
      d = bsum(d, red);
      dr = bsum(dr, red);
      const double ob = occupation(ne, b);
      const double extra = (b != j && ob != oj) ? dr / (eps[b] - ej) : 0.0;
      // y holds x; the final y_j = -(Q x) - sum extra psi_b  ->  accumulate as x - d psi_b + extra psi_b, negated below
      for (int t = threadIdx.x; t < N; t += NT) y[t] = fma(extra - d, pb[t], y[t]);
      __syncthreads();
    }
    const double sc = -oj * 2.0 / a.dx;
    for (int t = threadIdx.x; t < N; t += NT) vb[t] = fma(sc * pj[t], y[t], vb[t]);
    __syncthreads();
  }
  for (int t = threadIdx.x; t < N; t += NT) a.vbar[(size_t)s * N + t] = vb[t];
}

// the input densities n_k of all systems, contiguous [B][N]
__global__ void k_adj_gather_input(Adj a, int k, double* dst) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.B * a.N) return;
  const int s = (int)(idx / a.N), i = (int)(idx % a.N);
  dst[idx] = input_density(a, s, k)[i];
}

// grad[q] = sum over slots of thp[slot][q]: 32 parameters per workgroup, 8 interleaved partial
// sums each, combined in a fixed order
__global__ __launch_bounds__(256) void k_adj_reduce(int P, const double* thp, double* grad) {
  __shared__ double red[8][33];
  const int t = threadIdx.x, q = blockIdx.x * 32 + (t & 31), part = t >> 5;
  double g = 0.0;
  if (q < P)
    for (int sl = part; sl < ADJ_WG; sl += 8) g += thp[(size_t)sl * P + q];
  red[part][t & 31] = g;
  __syncthreads();
  if (t < 32 && q < P) {
    double v = 0.0;
    for (int k = 0; k < 8; ++k) v += red[k][t];
    grad[q] = v;
  }
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct LayA {
  size_t nbar, nob, nout, vbar, tmp, dbar, xbar, fv, fdv, thp, cs, ones, total;
};

int adj_check(const qex1d_problem* p) {
  if (!p) return faila(QEX_E_BADARG, "null problem");
  if (p->batch < 1 || p->num_grids < 5 || p->num_grids > QEX1D_MAX_GRIDS || !(p->dx > 0.0) || p->num_iterations < 1 ||
      p->max_electrons < 1 || p->max_electrons > QEX1D_MAX_ELECTRONS || p->num_mixing_iterations < 1)
    return faila(QEX_E_BADARG, "bad problem sizes");
  // with density_mse_converge_tolerance < 0 (checked below) the two loop kinds differ only in the
  // reflection window, which the reverse pass takes per system
  if (p->loop != QEX1D_LOOP_SCAN && p->loop != QEX1D_LOOP_PYTHON) return faila(QEX_E_BADARG, "bad loop kind");
  if (!(p->density_mse_converge_tolerance < 0.0))
    return faila(QEX_E_UNSUPPORTED, "the reverse pass needs density_mse_converge_tolerance < 0 (no frozen states)");
  const bool glob = p->xc_kind == QEX1D_XC_GLOBAL_MLP;
  if (p->xc_kind != QEX1D_XC_LOCAL_MLP && !glob)
    return faila(QEX_E_UNSUPPORTED, "the reverse pass needs a LOCAL_MLP or GLOBAL_MLP functional");
  if (p->mlp_layers != 2 && p->mlp_layers != 3)
    return faila(QEX_E_UNSUPPORTED, "the reverse pass supports 1 or 2 hidden layers");
  if (glob) {
    if (p->mlp_widths[0] != p->num_grids) return faila(QEX_E_BADARG, "GLOBAL_MLP input width must be num_grids");
  } else if (p->mlp_widths[0] < 1 || p->mlp_widths[0] > WMAXW || p->mlp_widths[0] % 2 == 0) {
    return faila(QEX_E_UNSUPPORTED, "window must be odd and <= %d", WMAXW);
  }
  for (int l = 1; l < p->mlp_layers; ++l)
    if (p->mlp_widths[l] < 1 || p->mlp_widths[l] > HMAX)
      return faila(QEX_E_UNSUPPORTED, "hidden widths must be in [1, %d] for the reverse pass", HMAX);
  if (p->mlp_widths[p->mlp_layers] != 1) return faila(QEX_E_UNSUPPORTED, "MLP output width must be 1");
  if (p->mlp_activation < 0 || p->mlp_activation > QEX_ACT_GELU_TANH || !(p->mlp_norm != 0.0))
    return faila(QEX_E_BADARG, "bad activation or norm");
  if (p->mlp_output != QEX1D_MLP_OUTPUT_LINEAR && p->mlp_output != QEX1D_MLP_OUTPUT_NEGATIVITY)
    return faila(QEX_E_BADARG, "bad mlp_output");
  if (p->stop_gradient_iterations < 0) return faila(QEX_E_BADARG, "stop_gradient_iterations must be >= 0");
  return QEX_OK;
}

int param_count(const qex1d_problem* p) {
  int n = 0;
  for (int l = 0; l < p->mlp_layers; ++l) n += p->mlp_widths[l] * p->mlp_widths[l + 1] + p->mlp_widths[l + 1];
  return n;
}

LayA layout_adj(const qex1d_problem* p) {
  LayA L{};
  size_t off = 0;
  const size_t B = p->batch, N = p->num_grids, T = p->num_iterations;
  const size_t W = p->xc_kind == QEX1D_XC_GLOBAL_MLP ? 0 : p->mlp_widths[0];  // window cotangents (local only)
  auto take = [&](size_t b) {
    const size_t o = off;
    off = al(off + b);
    return o;
  };
  L.nbar = take(B * N * 8);
  L.nob = take(B * N * 8);
  L.nout = take(B * N * 8);
  L.vbar = take(B * N * 8);
  L.tmp = take(B * N * 8);
  L.dbar = take(B * T * N * 8);
  L.xbar = take(B * N * W * 8);
  L.fv = take(B * N * 8);
  L.fdv = take(B * N * 8);
  L.thp = take((size_t)ADJ_WG * param_count(p) * 8);
  L.cs = take(B * 8);
  L.ones = take(B * 8);
  L.total = off;
  return L;
}

}  // namespace

extern "C" {

const char* qex1d_vjp_last_error(void) { return g_erra; }

size_t qex1d_kohn_sham_vjp_workspace_bytes(const qex1d_problem* p) {
  if (adj_check(p) != QEX_OK) return 0;
  return layout_adj(p).total;
}

int qex1d_kohn_sham_vjp(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* fwd,
                        const double* d_total_energy, const double* d_density, double* d_xc_params, void* workspace,
                        size_t workspace_bytes, void* stream_) {
  int rc = adj_check(p);
  if (rc) return rc;
  if (!in || !fwd || !d_xc_params) return faila(QEX_E_BADARG, "null inputs, forward outputs or gradient");
  if (!in->interaction_table || !in->external_potential || !in->initial_density || !in->num_electrons ||
      !in->xc_params)
    return faila(QEX_E_BADARG, "missing forward inputs");
  if (!fwd->density || !fwd->hartree_potential || !fwd->xc_potential || !fwd->orbitals)
    return faila(QEX_E_BADARG, "the forward outputs density, hartree_potential, xc_potential and orbitals are required");
  if (p->reflect && !in->reflect_window) return faila(QEX_E_BADARG, "reflection needs reflect_window");
  const size_t need = layout_adj(p).total;
  if (!workspace || workspace_bytes < need) return faila(QEX_E_WORKSPACE, "workspace too small: need %zu", need);
  hipStream_t st = (hipStream_t)stream_;
  const LayA L = layout_adj(p);
  char* ws = (char*)workspace;
  Adj a{};
  a.B = p->batch;
  a.N = p->num_grids;
  a.T = p->num_iterations;
  a.K = p->max_electrons / 2 + 1;
  a.M = p->num_mixing_iterations;
  a.W = p->mlp_widths[0];
  a.NL = p->mlp_layers - 1;
  a.H1 = p->mlp_widths[1];
  a.H2 = a.NL == 2 ? p->mlp_widths[2] : 0;
  a.act = p->mlp_activation;
  a.mlpout = p->mlp_output;
  a.reflect = p->reflect ? 1 : 0;
  a.win = in->reflect_window;
  a.sgs = p->stop_gradient_iterations - 1;  // last stopped iteration (-1: none)
  a.P = param_count(p);
  a.dx = p->dx;
  a.norm = p->mlp_norm;
  int off = 0;
  a.w1 = off;
  off += a.W * a.H1;
  a.b1 = off;
  off += a.H1;
  if (a.NL == 2) {
    a.w2 = off;
    off += a.H1 * a.H2;
    a.b2 = off;
    off += a.H2;
  }
  a.w3 = off;
  off += a.NL == 2 ? a.H2 : a.H1;
  a.b3 = off;
  a.table = in->interaction_table;
  a.vext = in->external_potential;
  a.n0 = in->initial_density;
  a.ne = in->num_electrons;
  a.mlp = in->xc_params;
  a.dens = fwd->density;
  a.vh = fwd->hartree_potential;
  a.vxc = fwd->xc_potential;
  a.orb = fwd->orbitals;
  a.dE = d_total_energy;
  a.dn = d_density;
  a.nbar = (double*)(ws + L.nbar);
  a.nob = (double*)(ws + L.nob);
  a.nout = (double*)(ws + L.nout);
  a.vbar = (double*)(ws + L.vbar);
  a.tmp = (double*)(ws + L.tmp);
  a.dbar = (double*)(ws + L.dbar);
  a.xbar = (double*)(ws + L.xbar);
  a.fv = (double*)(ws + L.fv);
  a.fdv = (double*)(ws + L.fdv);
  a.thp = (double*)(ws + L.thp);
  double* cs = (double*)(ws + L.cs);
  double* ones = (double*)(ws + L.ones);
  const long BN = (long)a.B * a.N;
  const double s2 = -0.5 / (a.dx * a.dx);
  const double Td = s2 * -2.5, Tb = s2 * (4.0 / 3.0), Tc = s2 * (-1.0 / 12.0);
  const int HL = a.NL == 2 ? a.H2 : a.H1;
  const size_t mlp_lds = (2 * (size_t)a.W * a.H1 + a.H1 + (a.NL == 2 ? 2 * (size_t)a.H1 * a.H2 + a.H2 : 0) + HL +
                          (size_t)(NT / 64) * 6 * 64) * 8;
  const size_t eig_lds = (size_t)(4 + 5) * a.N * 8;
  const bool glob = p->xc_kind == QEX1D_XC_GLOBAL_MLP;
  const size_t gmlp_lds = (2 * (size_t)a.N + 2 * (NT / 64) * 64 + 12 * 64 + 8) * 8;
  if ((!glob && mlp_lds > 150 * 1024) || eig_lds > 150 * 1024)
    return faila(QEX_E_UNSUPPORTED, "problem too large for LDS");
  if (eig_lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)k_adj_eig, hipFuncAttributeMaxDynamicSharedMemorySize, (int)eig_lds);
  if (!glob && mlp_lds > 64 * 1024) {
    (void)hipFuncSetAttribute((const void*)k_adj_mlp<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlp_lds);
    (void)hipFuncSetAttribute((const void*)k_adj_mlp<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlp_lds);
  }
  auto mlp = [&](const double* xin, const double* tin, const double* F1, const double* sf, const double* F2) {
    qexprof::Scope ps(qexprof::K_ADJ_MLP, st);
    if (a.NL == 2)
      hipLaunchKernelGGL(k_adj_mlp<2>, dim3(ADJ_WG), dim3(NT), mlp_lds, st, a, xin, tin, F1, sf, F2);
    else
      hipLaunchKernelGGL(k_adj_mlp<1>, dim3(ADJ_WG), dim3(NT), mlp_lds, st, a, xin, tin, F1, sf, F2);
  };
  // the global network: one workgroup per system (at most ADJ_WG, each then taking several in turn)
  const unsigned gG = (unsigned)(a.B < ADJ_WG ? a.B : ADJ_WG);
  auto gmlp = [&](const double* xin, const double* tin, const double* sf, double fd, const double* A2, double* dst) {
    qexprof::Scope ps(qexprof::K_ADJ_MLP, st);
    if (a.NL == 2)
      hipLaunchKernelGGL(k_adj_gmlp<2>, dim3(gG), dim3(NT), gmlp_lds, st, a, xin, tin, sf, fd, A2, dst);
    else
      hipLaunchKernelGGL(k_adj_gmlp<1>, dim3(gG), dim3(NT), gmlp_lds, st, a, xin, tin, sf, fd, A2, dst);
  };
  const unsigned gBN = (unsigned)((BN + NT - 1) / NT);
  // alpha_k as the forward computes it
  double alphas[4096];
  if (a.T > 4096) return faila(QEX_E_UNSUPPORTED, "too many iterations");
  double al_k = p->alpha;
  for (int k = 0; k < a.T; ++k) {
    alphas[k] = al_k;
    al_k *= p->alpha_decay;
  }
  {
    hipLaunchKernelGGL(k_zero_a, dim3(256), dim3(NT), 0, st, a.nbar, BN);
    hipLaunchKernelGGL(k_zero_a, dim3(256), dim3(NT), 0, st, a.dbar, BN * a.T);
    hipLaunchKernelGGL(k_zero_a, dim3(1024), dim3(NT), 0, st, a.thp, (long)ADJ_WG * a.P);
    // with one mixing difference nothing reaches the stopped iterations; with more,
    // the differences of those iterations still enter later mixes and are reversed
    const int k_end = a.M == 1 ? (a.sgs >= 0 ? a.sgs + 1 : 0) : 0;
    for (int k = a.T - 1; k >= k_end; --k) {
      hipLaunchKernelGGL(k_adj_mix, dim3(a.B), dim3(NT), 0, st, a, k, alphas[k]);
      hipLaunchKernelGGL(k_adj_scales, dim3((a.B + 63) / 64), dim3(64), 0, st, a, k, cs, ones);
      // with reflection the density passed on is S n_raw: its cotangent maps to S dbar_k on n_raw
      if (a.reflect) hipLaunchKernelGGL(k_sym_inplace, dim3(gBN), dim3(NT), 0, st, a, a.nob);
      // energy: v_H(n_raw) and the MLP at the XC density of the energy (n_raw, or S n_raw with reflection)
      hipLaunchKernelGGL(k_toeplitz, dim3((a.N + NT - 1) / NT, a.B), dim3(NT), (size_t)2 * a.N * 8, st, a.N, a.dx,
                         a.table, a.nout, a.tmp);
      hipLaunchKernelGGL(k_adj_energy, dim3(gBN), dim3(NT), 0, st, a, k, cs);
      if (glob) {
        // dx f(n_raw): nob += c grad f(n_raw), thetabar += c df/dtheta (no flip of the XC input)
        gmlp(a.nout, nullptr, cs, 0.0, nullptr, a.nob);
      } else if (a.reflect) {
        // dx F(S n_raw): nob += c S grad F(S n_raw) (vbar and tmp are free until the eigen adjoint)
        hipLaunchKernelGGL(k_sym_copy, dim3(gBN), dim3(NT), 0, st, a, a.vbar, a.nout);
        mlp(a.vbar, nullptr, a.vbar, cs, nullptr);
        hipLaunchKernelGGL(k_zero_a, dim3(256), dim3(NT), 0, st, a.tmp, BN);
        hipLaunchKernelGGL(k_adj_gather, dim3(gBN), dim3(NT), 0, st, a, a.tmp, nullptr, nullptr, a.fv, cs);
        hipLaunchKernelGGL(k_sym_add, dim3(gBN), dim3(NT), 0, st, a, a.nob, a.tmp, nullptr);
      } else {
        mlp(a.nout, nullptr, a.nout, cs, nullptr);
        hipLaunchKernelGGL(k_adj_gather, dim3(gBN), dim3(NT), 0, st, a, a.nob, nullptr, nullptr, a.fv, cs);
      }
      // eigen adjoint: nob -> vbar
      {
        qexprof::Scope ps(qexprof::K_ADJ_EIG, st);
        hipLaunchKernelGGL(k_adj_eig, dim3(a.B), dim3(NT), eig_lds, st, a, k, Td, Tb, Tc);
      }
      // Hartree and XC: nbar += dx U vbar + H_F vbar
      hipLaunchKernelGGL(k_toeplitz, dim3((a.N + NT - 1) / NT, a.B), dim3(NT), (size_t)2 * a.N * 8, st, a.N, a.dx,
                         a.table, a.vbar, a.tmp);
      // the input densities n_k of all systems, gathered contiguously into nout (free again)
      hipLaunchKernelGGL(k_adj_gather_input, dim3(gBN), dim3(NT), 0, st, a, k, a.nout);
      if (glob) {
        // v_xc = dx grad f: nbar += dx U vbar + dx H_f vbar, thetabar += dx d(vbar . grad f)/dtheta
        gmlp(a.nout, a.vbar, nullptr, a.dx, a.tmp, a.nbar);
      } else if (a.reflect) {
        // v_xc(n) = S grad F(S n): the second-order adjoint at (S n, S vbar), its n-cotangent mapped by S
        hipLaunchKernelGGL(k_sym_inplace, dim3(gBN), dim3(NT), 0, st, a, a.nout);
        hipLaunchKernelGGL(k_sym_inplace, dim3(gBN), dim3(NT), 0, st, a, a.vbar);
        mlp(a.nout, a.vbar, a.vbar, ones, a.nout);
        hipLaunchKernelGGL(k_zero_a, dim3(256), dim3(NT), 0, st, a.nob, BN);
        hipLaunchKernelGGL(k_adj_gather, dim3(gBN), dim3(NT), 0, st, a, a.nob, nullptr, nullptr, a.fdv, ones);
        hipLaunchKernelGGL(k_sym_add, dim3(gBN), dim3(NT), 0, st, a, a.nbar, a.nob, a.tmp);
      } else {
        mlp(a.nout, a.vbar, a.vbar, ones, a.nout);
        hipLaunchKernelGGL(k_adj_gather, dim3(gBN), dim3(NT), 0, st, a, a.nbar, nullptr, a.tmp, a.fdv, ones);
      }
    }
    hipLaunchKernelGGL(k_adj_reduce, dim3((a.P + 31) / 32), dim3(256), 0, st, a.P, a.thp, d_xc_params);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return faila(QEX_E_LAUNCH, "qex1d_kohn_sham_vjp: %s", hipGetErrorString(e));
  return QEX_OK;
}

}  // extern "C"
