// Device code of the uiuc_7sp chemistry beyond the pointwise source: the
// constant-pressure reactor and the learned-surrogate network (training
// GEMMs, Adam, fused gated inference).  Included by mg_chem.hip (the entry
// points) and mg_ensemble.hip (the batched inference).
#pragma once

#include "mgdg_host.h"

namespace mgchem {

using mgdg::c_mech;
using mgdg::MIX_NSP;

constexpr int kIn = 32, kH = 256, kOut = 32;
constexpr int kNIn = 10, kNOut = 7;
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + int64_t(kIn) * kH;
constexpr int64_t kOffW2 = kOffB1 + kH;
constexpr int64_t kOffB2 = kOffW2 + int64_t(kH) * kH;
constexpr int64_t kOffW3 = kOffB2 + kH;
constexpr int64_t kOffB3 = kOffW3 + int64_t(kH) * kH;
constexpr int64_t kOffW4 = kOffB3 + kH;
constexpr int64_t kOffB4 = kOffW4 + int64_t(kH) * kOut;
constexpr int64_t kNParam = kOffB4 + kOut;
// BF16 copies: Wt1..Wt4 then W1..W4 (same element counts as the W_l)
constexpr int64_t kBfW1 = 0, kBfW2 = kBfW1 + int64_t(kIn) * kH,
                  kBfW3 = kBfW2 + int64_t(kH) * kH, kBfW4 = kBfW3 + int64_t(kH) * kH,
                  kBfHalf = kBfW4 + int64_t(kH) * kOut;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// GELU in its tanh form (Hendrycks & Gimpel 2016, eq. for the fast
// approximation): gelu(z) = z (1 + tanh u) / 2, u = sqrt(2/pi) (z + 0.044715 z^3).
// With E = exp(2u) and r = 1 / (1 + E): tanh u = 1 - 2r, so
//   gelu(z)  = z (1 - r)
//   gelu'(z) = (1 - r) + 2 z r (1 - r) sqrt(2/pi) (1 + 3 * 0.044715 z^2)
// evaluated with the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32),
// about 1e-7 relative: far below the BF16 rounding of the activations.  The
// erf form costs several times the VALU work and, at 128 activations per
// lane and layer, outweighed the matrix-core time of the fused inference.
constexpr float kGeluC0 = 0.7978845608028654f;   // sqrt(2 / pi)
constexpr float kGeluC1 = 0.044715f;
constexpr float kLog2e2 = 2.0f * 1.4426950408889634f;

__device__ __forceinline__ float gelu_r(float z) {
  const float u = kGeluC0 * (z + kGeluC1 * z * z * z);
  const float e = __builtin_amdgcn_exp2f(kLog2e2 * u);
  return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float z) { return z * (1.0f - gelu_r(z)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float r = gelu_r(z);
  return (1.0f - r) + 2.0f * z * r * (1.0f - r) * kGeluC0 * (1.0f + 3.0f * kGeluC1 * z * z);
}

// {{{ constant-pressure reactor (FP64, one reactor per thread)
//
// State (T, Y); the pressure p and the mixture enthalpy h0 are constant
// (adiabatic, constant pressure).  dY/dt = W omega(rho, T, Y) / rho with
// rho = p / (R(Y) T); after every RK4 stage T is recovered from h(T, Y) = h0
// by kRxNewton Newton steps from the previous stage's T.

constexpr int kRxNewton = 4;

__device__ __forceinline__ double rx_temperature(double h0, double t, const double* y) {
#pragma unroll 1
  for (int it = 0; it < kRxNewton; ++it) {
    double h, cp;
    mgdg::mix_h_cp(t, y, h, cp);
    t = t + (h0 - h) / cp;
  }
  return t;
}

__device__ __forceinline__ void rx_rhs(double p, double t, const double* y, double* dydt) {
  const double rho = p / (mgdg::mix_gas_const(y) * t);
  double src[MIX_NSP];
  mgdg::mix_species_source(rho, t, y, src);
  const double irho = 1.0 / rho;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) dydt[k] = src[k] * irho;
}

// one classical RK4 step of size dt
__device__ __forceinline__ void rx_step(double p, double h0, double dt, double& t,
                                        double* y) {
  double k1[MIX_NSP], k2[MIX_NSP], k3[MIX_NSP], k4[MIX_NSP], ys[MIX_NSP];
  rx_rhs(p, t, y, k1);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) ys[k] = y[k] + 0.5 * dt * k1[k];
  const double t1 = rx_temperature(h0, t, ys);
  rx_rhs(p, t1, ys, k2);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) ys[k] = y[k] + 0.5 * dt * k2[k];
  const double t2 = rx_temperature(h0, t1, ys);
  rx_rhs(p, t2, ys, k3);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) ys[k] = y[k] + dt * k3[k];
  const double t3 = rx_temperature(h0, t2, ys);
  rx_rhs(p, t3, ys, k4);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k)
    y[k] = y[k] + dt / 6.0 * (k1[k] + 2.0 * k2[k] + 2.0 * k3[k] + k4[k]);
  t = rx_temperature(h0, t3, y);
}

// Trajectories of n reactors from (t0, p, y0 [7][n]).  Pass 1 finds the
// ignition step (T > T0 + dt_rise) within nmax steps; pass 2 restarts and
// records nsamp states at equal step spacing over min(2 x ignition, nmax)
// steps (nmax without ignition), each with the change of Y over the next
// step: states [nsamp][n][10] = (T, p, Y_1..7, dt), dy [nsamp][n][7].
static __global__ __launch_bounds__(64) void reactor_kernel(
    int64_t n, const double* __restrict__ t0, const double* __restrict__ p0,
    const double* __restrict__ y0, double dt, int64_t nmax, double t_rise, int nsamp,
    double* __restrict__ t_ign, double* __restrict__ states, double* __restrict__ dy) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p = p0[i];
  double y[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = y0[k * n + i];
  double t = t0[i];
  double h0, cp0;
  mgdg::mix_h_cp(t, y, h0, cp0);
  int64_t ign = -1;
  for (int64_t s = 0; s < nmax; ++s) {
    rx_step(p, h0, dt, t, y);
    if (t > t0[i] + t_rise) {
      ign = s + 1;
      break;
    }
  }
  t_ign[i] = ign > 0 ? double(ign) * dt : -1.0;
  if (nsamp <= 0) return;
  const int64_t nrec = ign > 0 ? (2 * ign < nmax ? 2 * ign : nmax) : nmax;
  int64_t stride = nrec / nsamp;
  if (stride < 1) stride = 1;
  t = t0[i];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = y0[k * n + i];
  int j = 0;
  for (int64_t s = 0; j < nsamp; ++s) {
    const bool rec = (s % stride) == 0;
    double* st = states + (int64_t(j) * n + i) * 10;
    if (rec) {
      st[0] = t;
      st[1] = p;
#pragma unroll
      for (int k = 0; k < MIX_NSP; ++k) st[2 + k] = y[k];
      st[9] = dt;
    }
    double yo[MIX_NSP];
#pragma unroll
    for (int k = 0; k < MIX_NSP; ++k) yo[k] = y[k];
    rx_step(p, h0, dt, t, y);
    if (rec) {
      double* d = dy + (int64_t(j) * n + i) * 7;
#pragma unroll
      for (int k = 0; k < MIX_NSP; ++k) d[k] = y[k] - yo[k];
      ++j;
    }
  }
}

// Change of Y over dt (nsub RK4 steps of dt / nsub at constant pressure)
// from states [n][10] = (T, p, Y, ...): dy [n][7].
static __global__ __launch_bounds__(64) void substep_kernel(int64_t n,
                                                     const double* __restrict__ states,
                                                     double dt, int nsub,
                                                     double* __restrict__ dy) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* st = states + i * 10;
  double t = st[0];
  const double p = st[1];
  double y[MIX_NSP], yo[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = yo[k] = st[2 + k];
  double h0, cp0;
  mgdg::mix_h_cp(t, y, h0, cp0);
  const double h = dt / nsub;
  for (int s = 0; s < nsub; ++s) rx_step(p, h0, h, t, y);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) dy[i * 7 + k] = y[k] - yo[k];
}

// }}}

// {{{ features and normalisation

struct Norm {
  double in_mean[kNIn], in_scale[kNIn];
  double out_mean[kNOut], out_scale[kNOut];
  double lo[kNIn], hi[kNIn];
  double check_tol;
  int sample_period, sample_offset;
};

// raw features (log T, log p, Y_1..7, dt) -> normalised FP32
__device__ __forceinline__ void features(double t, double p, const double* y, double dt,
                                         const Norm& nm, float* x, double* raw) {
  raw[0] = log(t);
  raw[1] = log(p);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) raw[2 + k] = y[k];
  raw[9] = dt;
#pragma unroll
  for (int f = 0; f < kNIn; ++f) x[f] = float((raw[f] - nm.in_mean[f]) / nm.in_scale[f]);
}

// Training inputs: states [n][10] -> x [n][32] BF16; dy [n][7] -> target
// [n][32] FP32 (normalised; both zero-padded).
static __global__ __launch_bounds__(256) void prepare_kernel(int64_t n,
                                                      const double* __restrict__ states,
                                                      const double* __restrict__ dy,
                                                      Norm nm, __bf16* __restrict__ x,
                                                      float* __restrict__ tgt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* st = states + i * 10;
  float xf[kNIn];
  double raw[kNIn];
  features(st[0], st[1], st + 2, st[9], nm, xf, raw);
#pragma unroll
  for (int f = 0; f < kIn; ++f) x[i * kIn + f] = __bf16(f < kNIn ? xf[f] : 0.0f);
  if (dy) {
#pragma unroll
    for (int k = 0; k < kOut; ++k)
      tgt[i * kOut + k] =
          k < kNOut ? float((dy[i * 7 + k] - nm.out_mean[k]) / nm.out_scale[k]) : 0.0f;
  }
}

// }}}

// {{{ training GEMM: C[M][N] = A[M][K] . Bt[N][K]^T on the matrix cores
//
// Both operands are K-contiguous, so every MFMA operand fragment is one
// 16-byte load per lane straight from global memory (L2-resident at these
// sizes): lane l of v_mfma_f32_16x16x32_bf16 holds A[l%16][8(l/16) .. +7]
// and Bt[l%16][8(l/16) .. +7] and receives D[4(l/16) .. +3][l%16].
// Workgroup tile 64 x 64 (four waves, 32 x 32 each: 2 x 2 fragments);
// M % 32 == 0, N % 16 == 0, K % 32 == 0 are checked on the host.  blockIdx.z
// splits K (partial products to C + z M N, summed by the Adam kernel).

constexpr int kEpiStore = 0;   // C (+ bias) -> FP32 [M][N]
constexpr int kEpiGelu = 1;    // z = C + bias: z FP32 [M][N]; gelu(z) BF16 [M][N] and [N][M]
constexpr int kEpiBwd = 2;     // dz = C * gelu'(zpre): BF16 [M][N] and [N][M]

struct GemmArgs {
  const __bf16* a;        // A [M][K], row stride lda
  const __bf16* bt;       // Bt [N][K], row stride ldb
  int64_t lda, ldb;
  int64_t m, n, k;
  int64_t ksplit;
  int64_t ldc, cslice;    // kEpiStore: C row stride, stride between K-split slices
  const float* bias;      // kEpiStore (optional) / kEpiGelu
  float* c;               // kEpiStore
  float* zout;            // kEpiGelu: pre-activation
  const float* zpre;      // kEpiBwd
  __bf16* out;            // kEpiGelu / kEpiBwd: [M][N]
  __bf16* out_t;          // kEpiGelu / kEpiBwd: [N][M] (optional)
};

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t m0 = int64_t(blockIdx.x) * 64 + (w >> 1) * 32;
  const int64_t n0 = int64_t(blockIdx.y) * 64 + (w & 1) * 32;
  if (m0 >= g.m) return;   // M % 32 == 0: a wave's 32 rows are all valid
  const int64_t kper = g.k / g.ksplit;
  const int64_t k_lo = int64_t(blockIdx.z) * kper;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const bool ncol1 = n0 + 16 < g.n;   // second column fragment exists
  if (n0 < g.n) {
    for (int64_t k0 = k_lo; k0 < k_lo + kper; k0 += 32) {
      bf16x8 fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        fa[i] = *reinterpret_cast<const bf16x8*>(g.a + (m0 + 16 * i + r) * g.lda + k0 + kq);
      fb[0] = *reinterpret_cast<const bf16x8*>(g.bt + (n0 + r) * g.ldb + k0 + kq);
      if (ncol1)
        fb[1] = *reinterpret_cast<const bf16x8*>(g.bt + (n0 + 16 + r) * g.ldb + k0 + kq);
      else
        fb[1] = fb[0];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0,
                                                              0, 0);
    }
  }
  if (n0 >= g.n) return;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t col = n0 + 16 * j + r;
      if (j == 1 && !ncol1) continue;
      const int64_t row0 = m0 + 16 * i + (lane >> 4) * 4;
      if constexpr (EPI == kEpiStore) {
        const float b = g.bias ? g.bias[col] : 0.0f;
        float* c = g.c + int64_t(blockIdx.z) * g.cslice;
#pragma unroll
        for (int e = 0; e < 4; ++e) c[(row0 + e) * g.ldc + col] = acc[i][j][e] + b;
      } else if constexpr (EPI == kEpiGelu) {
        const float b = g.bias[col];
        bf16x4 t;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float z = acc[i][j][e] + b;
          g.zout[(row0 + e) * g.n + col] = z;
          const __bf16 h = __bf16(gelu(z));
          g.out[(row0 + e) * g.n + col] = h;
          t[e] = h;
        }
        if (g.out_t) *reinterpret_cast<bf16x4*>(g.out_t + col * g.m + row0) = t;
      } else {
        bf16x4 t;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float z = g.zpre[(row0 + e) * g.n + col];
          const __bf16 d = __bf16(acc[i][j][e] * gelu_grad(z));
          g.out[(row0 + e) * g.n + col] = d;
          t[e] = d;
        }
        if (g.out_t) *reinterpret_cast<bf16x4*>(g.out_t + col * g.m + row0) = t;
      }
    }
  }
}

// Loss of a batch: out, tgt [B][32] (only the first 7 columns count).
// L = sum (out - tgt)^2 / (7 B); dL/dout as BF16 [B][32] and [32][B];
// per-block loss sums into loss_part[blockIdx.x].
static __global__ __launch_bounds__(256) void loss_kernel(int64_t b, const float* __restrict__ out,
                                                   const float* __restrict__ tgt,
                                                   __bf16* __restrict__ dout,
                                                   __bf16* __restrict__ dout_t,
                                                   float* __restrict__ loss_part) {
  __shared__ float red[256];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < b * kOut) {
    const int64_t row = i / kOut;
    const int col = int(i - row * kOut);
    float d = 0.0f;
    if (col < kNOut) {
      const float e = out[i] - tgt[i];
      l = e * e;
      d = 2.0f * e / float(kNOut * b);
    }
    dout[i] = __bf16(d);
    dout_t[int64_t(col) * b + row] = __bf16(d);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] / float(kNOut * b);
}

// Bias gradient db[n] = sum_m dz[m][n] from dz^T [N][M]: one workgroup per n,
// written to slice 0 of the gradient partials (other slices zeroed).
static __global__ __launch_bounds__(256) void bias_grad_kernel(int64_t m,
                                                        const __bf16* __restrict__ dzt,
                                                        float* __restrict__ gpart,
                                                        int64_t off, int ksplit) {
  __shared__ float red[256];
  const int64_t col = blockIdx.x;
  float s = 0.0f;
  for (int64_t r = threadIdx.x; r < m; r += blockDim.x) s += float(dzt[col * m + r]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int t = 128; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    gpart[off + col] = red[0];
    for (int z = 1; z < ksplit; ++z) gpart[int64_t(z) * kNParam + off + col] = 0.0f;
  }
}

// Where a master parameter index lives: layer (0..3), weight or bias, the
// layer's K_in / N_out, the element's (k, n) and its BF16 offset.
struct ParamSlot {
  bool weight;
  int kin, nout;
  int64_t k, n, bf;
};

__device__ __forceinline__ ParamSlot param_slot(int64_t i) {
  ParamSlot s{};
  int64_t woff, boff;
  if (i < kOffW2) {
    woff = kOffW1; boff = kOffB1; s.kin = kIn; s.nout = kH; s.bf = kBfW1;
  } else if (i < kOffW3) {
    woff = kOffW2; boff = kOffB2; s.kin = kH; s.nout = kH; s.bf = kBfW2;
  } else if (i < kOffW4) {
    woff = kOffW3; boff = kOffB3; s.kin = kH; s.nout = kH; s.bf = kBfW3;
  } else {
    woff = kOffW4; boff = kOffB4; s.kin = kH; s.nout = kOut; s.bf = kBfW4;
  }
  s.weight = i < boff;
  const int64_t loc = i - woff;
  s.k = loc / s.nout;
  s.n = loc - s.k * s.nout;
  return s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; the gradient is the
// sum of the ksplit partial slices.  Weights are re-emitted as BF16 in both
// layouts.  step counts from 1.
static __global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ prm,
                                                   float* __restrict__ mom,
                                                   float* __restrict__ vel,
                                                   const float* __restrict__ gpart,
                                                   int ksplit, float lr, float beta1,
                                                   float beta2, float eps, float bc1,
                                                   float bc2, __bf16* __restrict__ wbf,
                                                   float* __restrict__ grad_out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= kNParam) return;
  float g = 0.0f;
  for (int z = 0; z < ksplit; ++z) g += gpart[int64_t(z) * kNParam + i];
  if (grad_out) grad_out[i] = g;
  float p = prm[i];
  if (lr != 0.0f) {
    const float m = beta1 * mom[i] + (1.0f - beta1) * g;
    const float v = beta2 * vel[i] + (1.0f - beta2) * g * g;
    mom[i] = m;
    vel[i] = v;
    p -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    prm[i] = p;
  }
  const ParamSlot sl = param_slot(i);
  if (!sl.weight) return;
  wbf[sl.bf + sl.n * sl.kin + sl.k] = __bf16(p);                // Wt [N][K]
  wbf[kBfHalf + sl.bf + sl.k * sl.nout + sl.n] = __bf16(p);     // W  [K][N]
}

// }}}

// {{{ fused, gated inference into a right-hand side
//
// A workgroup (four waves) owns R = floor(128 / np) whole elements (rows
// = nodes, at most 128).  Activations live in LDS as BF16 [128][264] (row
// stride padded by 8 so that the 16 rows of an operand fragment start on
// different banks); each hidden layer is computed with every wave owning
// 64 output columns (8 row fragments x 4 column fragments, K in steps of
// 32; the weight fragments are read once per workgroup from L2), then
// GELU'd and written back over its input after a barrier.  The output layer
// (32 columns) splits the row fragments over the waves.  Gate per element:
// any node outside the training bounds, or a sampled node (global index
// = sample_offset mod sample_period) whose surrogate source misses the exact
// one by more than check_tol output scales, sends the whole element to the
// exact rates (mix_species_source).

constexpr int kRows = 128;
constexpr int kLdsStride = kH + 8;
constexpr int kOutStride = kLdsStride / 2;   // output rows over activation rows (floats)
static_assert(kOut <= kOutStride, "an output row fits an activation row");

// One hidden layer of the fused inference, in place on the activations
// act [128][kLdsStride] (BF16, LDS), formed transposed: D = W X^T with the
// weights W [out][k] as the A operand and the activations as the B operand,
// so that each lane ends with 4 consecutive output features of one row, which
// go to LDS as one 8-byte store after the bias, GELU and a packed BF16
// conversion.  Wave w owns output features [64 w, 64 w + 64) of every row:
// 4 x 8 accumulator fragments; the weight fragments of the next K step are
// fetched while the current one's 32 matrix-core products run (two register
// sets).
template <int KDIM>
__device__ __forceinline__ void hidden_layer(__bf16* act, const __bf16* __restrict__ wt,
                                             const float* __restrict__ bias, int w,
                                             int lane) {
  constexpr int NKS = KDIM / 32;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int c0 = w * 64;
  floatx4 acc[4][8];   // [out tile j][row tile i]: D[out][row]
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[j][i] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto load_w = [&](bf16x8* fw, int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      fw[j] = *reinterpret_cast<const bf16x8*>(wt + int64_t(c0 + 16 * j + r16) * KDIM + k0 +
                                               kq);
  };
  auto mma = [&](const bf16x8* fw, int k0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bf16x8 fx =
          *reinterpret_cast<const bf16x8*>(act + (16 * i + r16) * kLdsStride + k0 + kq);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw[j], fx, acc[j][i], 0, 0, 0);
    }
  };
  bf16x8 fw0[4], fw1[4];
  load_w(fw0, 0);
  if constexpr (NKS == 1) {
    mma(fw0, 0);
  } else {
#pragma unroll 1
    for (int ks = 0; ks < NKS; ks += 2) {
      load_w(fw1, (ks + 1) * 32);
      mma(fw0, ks * 32);
      if (ks + 2 < NKS) load_w(fw0, (ks + 2) * 32);
      mma(fw1, (ks + 1) * 32);
    }
  }
  __syncthreads();   // every wave has read this layer's input
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int o = c0 + 16 * j + 4 * (lane >> 4);   // this lane's 4 output features
    float b[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) b[e] = bias[o + e];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      floatx4 v;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = gelu(acc[j][i][e] + b[e]);
      *reinterpret_cast<bf16x4*>(act + (16 * i + r16) * kLdsStride + o) =
          __builtin_convertvector(v, bf16x4);
    }
  }
  __syncthreads();
}

// AS 1: global-qualified pointers, for argument blocks read from memory
// (see mgdg::KArgsT)
#define MG_P(T) typename mgdg::ArgPtr<AS, T>::type
template <int AS>
struct InferArgsT {
  int64_t n;          // nodes
  int np;             // nodes per element
  int dim;
  double dt;
  MG_P(const double) q;    // [F][n]
  MG_P(const double) temp; // [n]
  MG_P(double) out;        // add: rhs [F][n] (species rows += source); else src [7][n]
  int add;
  MG_P(const float) prm;
  MG_P(const __bf16) wbf;
  MG_P(int) fallback;      // optional counter of elements sent to the exact rates
  Norm nm;
};
#undef MG_P
using InferArgs = InferArgsT<0>;
static_assert(sizeof(InferArgsT<0>) == sizeof(InferArgsT<1>), "argument block layouts");

template <typename A>
__device__ __forceinline__ void infer_body(const A& a, int64_t blk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __bf16* act = reinterpret_cast<__bf16*>(smem_raw);                  // [128][264]
  // the output layer's results overlay the activations: row r at the start
  // of activation row r (each wave reads and writes only its own 32 rows
  // there), so that two workgroups fit a CU's LDS
  float* outp = reinterpret_cast<float*>(smem_raw);                    // [128][132]
  int* eflag = reinterpret_cast<int*>(act + kRows * kLdsStride);       // [128]
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int epb = kRows / a.np;
  const int rows = epb * a.np;
  const int64_t node0 = blk * rows;
  const int64_t n = a.n;
  const int sp0 = 2 + a.dim;

  // input features; out-of-bounds nodes flag their element
  if (threadIdx.x < kRows) eflag[threadIdx.x] = 0;
  __syncthreads();
  if (threadIdx.x < kRows) {
    const int rr = threadIdx.x;
    const int64_t g = node0 + rr;
    float xf[kNIn];
    bool ok = true;
    if (rr < rows && g < n) {
      const double rho = a.q[g];
      double y[MIX_NSP];
#pragma unroll
      for (int k = 0; k < MIX_NSP; ++k) y[k] = a.q[(sp0 + k) * n + g] / rho;
      const double t = a.temp[g];
      const double p = rho * mgdg::mix_gas_const(y) * t;
      double raw[kNIn];
      features(t, p, y, a.dt, a.nm, xf, raw);
#pragma unroll
      for (int f = 0; f < kNIn; ++f) ok = ok && raw[f] >= a.nm.lo[f] && raw[f] <= a.nm.hi[f];
    } else {
#pragma unroll
      for (int f = 0; f < kNIn; ++f) xf[f] = 0.0f;
    }
#pragma unroll
    for (int f = 0; f < kIn; ++f) act[rr * kLdsStride + f] = __bf16(f < kNIn ? xf[f] : 0.0f);
    if (!ok) atomicOr(&eflag[rr / a.np], 1);
  }
  __syncthreads();

  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  // hidden layers
  // generic views of the weights (the compiler keeps the global address
  // space the argument block states)
  const __bf16* wbf = (const __bf16*)a.wbf;
  const float* prm = (const float*)a.prm;
  hidden_layer<kIn>(act, wbf + kBfW1, prm + kOffB1, w, lane);
  hidden_layer<kH>(act, wbf + kBfW2, prm + kOffB2, w, lane);
  hidden_layer<kH>(act, wbf + kBfW3, prm + kOffB3, w, lane);
  // output layer (32 columns: 2 fragments): wave w owns row fragments
  // 2w, 2w + 1
  {
    const __bf16* wt = wbf + kBfW4;
    const float* bias = prm + kOffB4;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < kH; k0 += 32) {
      bf16x8 fb[2];
#pragma unroll
      for (int j = 0; j < 2; ++j)
        fb[j] = *reinterpret_cast<const bf16x8*>(wt + int64_t(16 * j + r16) * kH + k0 + kq);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const bf16x8 fa = *reinterpret_cast<const bf16x8*>(
            act + (16 * (2 * w + i) + r16) * kLdsStride + k0 + kq);
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb[j], acc[i][j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float b = bias[16 * j + r16];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row0 = 16 * (2 * w + i) + (lane >> 4) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          outp[(row0 + e) * kOutStride + 16 * j + r16] = acc[i][j][e] + b;
      }
    }
  }
  __syncthreads();

  // surrogate sources; sampled exact-rate check
  double src[MIX_NSP];
  const int rr = threadIdx.x;
  const int64_t g = node0 + rr;
  const bool mine = rr < kRows && rr < rows && g < n;
  double rho = 1.0, y[MIX_NSP];
  if (mine) {
    rho = a.q[g];
#pragma unroll
    for (int k = 0; k < MIX_NSP; ++k) y[k] = a.q[(sp0 + k) * n + g] / rho;
    const double irdt = rho / a.dt;
#pragma unroll
    for (int k = 0; k < MIX_NSP; ++k)
      src[k] = (double(outp[rr * kOutStride + k]) * a.nm.out_scale[k] + a.nm.out_mean[k]) *
               irdt;
    if (a.nm.sample_period > 0 && g % a.nm.sample_period == a.nm.sample_offset) {
      double ex[MIX_NSP];
      mgdg::mix_species_source(rho, a.temp[g], y, ex);
      bool bad = false;
#pragma unroll
      for (int k = 0; k < MIX_NSP; ++k)
        bad = bad || !(fabs(src[k] - ex[k]) <= a.nm.check_tol * a.nm.out_scale[k] * irdt);
      if (bad) atomicOr(&eflag[rr / a.np], 1);
    }
  }
  __syncthreads();
  if (mine) {
    if (eflag[rr / a.np]) mgdg::mix_species_source(rho, a.temp[g], y, src);
#pragma unroll
    for (int k = 0; k < MIX_NSP; ++k) {
      if (a.add) {
        a.out[(sp0 + k) * n + g] += src[k];
      } else {
        a.out[k * n + g] = src[k];
      }
    }
  }
  if (a.fallback && threadIdx.x == 0) {
    int c = 0;
    for (int e = 0; e < epb; ++e)
      if (node0 + int64_t(e) * a.np < n) c += eflag[e];
    if (c) atomicAdd((int*)a.fallback, c);
  }
}

static __global__ __launch_bounds__(256, 2) void infer_kernel(InferArgs a) {
  infer_body(a, blockIdx.x);
}

// LDS bytes of infer_body / infer_kernel
constexpr size_t kInferLds = sizeof(__bf16) * kRows * kLdsStride + sizeof(int) * kRows;

// }}}

}  // namespace mgchem

// host-side conversion and checks of mg_chem_norm
namespace {

mgchem::Norm to_norm(const mg_chem_norm* s) {
  mgchem::Norm n{};
  for (int i = 0; i < mgchem::kNIn; ++i) {
    n.in_mean[i] = s->in_mean[i];
    n.in_scale[i] = s->in_scale[i];
    n.lo[i] = s->lo[i];
    n.hi[i] = s->hi[i];
  }
  for (int k = 0; k < mgchem::kNOut; ++k) {
    n.out_mean[k] = s->out_mean[k];
    n.out_scale[k] = s->out_scale[k];
  }
  n.check_tol = s->check_tol;
  n.sample_period = s->sample_period;
  n.sample_offset = s->sample_offset;
  return n;
}

mg_status check_norm(const mg_chem_norm* s) {
  if (!s) return fail(MG_ERR_ARG, "normalisation NULL");
  for (int i = 0; i < mgchem::kNIn; ++i)
    if (!(s->in_scale[i] > 0.0)) return fail(MG_ERR_ARG, "in_scale must be positive");
  for (int k = 0; k < mgchem::kNOut; ++k)
    if (!(s->out_scale[k] > 0.0)) return fail(MG_ERR_ARG, "out_scale must be positive");
  if (s->sample_period < 0 || (s->sample_period > 0 && (s->sample_offset < 0 ||
                                                         s->sample_offset >= s->sample_period)))
    return fail(MG_ERR_ARG, "sample_offset must lie in [0, sample_period)");
  return MG_OK;
}

}  // namespace
