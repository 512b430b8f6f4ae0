// MI355X (gfx950) kernels of the DG right-hand side and their C ABI.
//
// Element-batched kernels: a workgroup owns EPB consecutive elements and runs
// three phases over them with LDS as the exchange buffer:
//   1. volume nodes   (thread per node): load the state (and gradients),
//      evaluate the pointwise physics and the contravariant fluxes
//      G_a = sum_i (dr_a/dx_i) F_i  ->  LDS
//   2. face nodes     (thread per face node): coalesced loads of the own
//      trace and indexed loads of the neighbour trace (interior_trace_pairs),
//      numerical flux times the face/volume Jacobian ratio  ->  LDS
//   3. volume nodes   (thread per node): the element-local contractions
//      sum_a Dw_a G_a - LIFT f*   (weak divergence, face mass and inverse
//      mass fused, mirgecom/operators.py:66-95), written once to HBM.
// The reference matrices (Dw_a = M^-1 D_a^T M, LIFT = M^-1 E) live in LDS,
// stored transposed so that consecutive nodes read consecutive banks.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>

#include <cstdio>
#include <cstring>
#include <string>

#include "mgdg.h"
#include "mgdg_physics.h"

namespace mgdg {

constexpr int kBlock = 256;
constexpr int kLdsBudget = 64 * 1024;

template <int DIM, int P>
struct Elem {
  static constexpr int NP = binom(P + DIM, DIM);
  static constexpr int NFP = binom(P + DIM - 1, DIM - 1);
  static constexpr int NF = DIM + 1;
  static constexpr int NFN = NF * NFP;
  static constexpr int MAT = DIM * NP * NP + NFN * NP;  // doubles of Dw + LIFT
};

// Argument block of the element-batched kernels.  AS 0: generic pointers
// (a kernel argument: the compiler treats its pointers as global memory);
// AS 1: global-qualified pointers, for blocks read from memory (ensemble
// descriptors), whose generic pointers would otherwise be accessed with
// flat instructions.  Both have the same layout.
template <int AS, typename T>
struct ArgPtr {
  using type = T*;
};
template <typename T>
struct ArgPtr<1, T> {
  using type = __attribute__((address_space(1))) T*;
};

#define MG_P(T) typename ArgPtr<AS, T>::type
template <int AS>
struct KArgsT {
  MG_P(const double) q;
  MG_P(const double) temp;
  MG_P(const double) grad;
  MG_P(const double) qb;
  MG_P(const double) tb;
  MG_P(double) out;
  MG_P(const double) rx;
  MG_P(const double) normals;
  MG_P(const double) fscale;
  MG_P(const int32_t) vmap_p;
  MG_P(const double) dw;
  MG_P(const double) lift;
  MG_P(const int32_t) fmask;
  MG_P(const double) ops;   // packed operators (mg_mesh_prepare)
  MG_P(const int32_t) border; // block visiting-order tables, or NULL
  MG_P(const double) fnet;  // nodal net flux [F*DIM][N] (kFlux)
  MG_P(const double) fnb;   // exterior net normal flux [F][nbnd] (kFlux)
  MG_P(const float) fnet32; // nodal net flux, FP32 (mixed precision)
  MG_P(float) fnet32_out;   // FP32 nodal net flux output (mixed precision)
  MG_P(double) grad_out;    // optional gradient output of the fused flux kernel
  MG_P(double) fnb_out;     // exterior net normal flux output of the flux kernel
  MG_P(float) grad32_out;   // FP32 gradient output (mixed-precision mixture path)
  MG_P(const float) grad32; // FP32 gradients read by the node-parallel net flux
  // Mengaldo boundary data per boundary node (NULL wflag: none)
  MG_P(const double) wq;     // CV_bc [F][nbnd]
  MG_P(const double) wt;     // T_bc [nbnd]
  MG_P(const double) wts;    // temperature of state_bc [nbnd] (mixtures)
  MG_P(const int32_t) wflag; // MG_BC_* bits [nbnd]
  MG_P(const int32_t) wvol;  // interior volume node of each boundary node [nbnd]
  MG_P(const double) wn;     // unit normal [DIM][nbnd]
  MG_P(double) wwork;        // boundary gradients / viscous fluxes [(F+1)*DIM][nbnd]
  int64_t K;
  int64_t nbnd;
  int64_t nnodes;      // K * NP
  GasDev gas;
  int flux_kind;
  int inviscid_on;
};
#undef MG_P
using KArgs = KArgsT<0>;
static_assert(sizeof(KArgsT<0>) == sizeof(KArgsT<1>), "argument block layouts");


template <int DIM, int P, typename A>
__device__ __forceinline__ void load_matrices(double* s_dw, double* s_lift,
                                              int* s_fmask, const A& a) {
  using E = Elem<DIM, P>;
  // s_dw[(r*NP + j)*NP + i] = Dw_r[i][j];  s_lift[m*NP + i] = LIFT[i][m]
  for (int t = threadIdx.x; t < DIM * E::NP * E::NP; t += blockDim.x) {
    const int r = t / (E::NP * E::NP);
    const int rem = t - r * E::NP * E::NP;
    const int i = rem / E::NP;
    const int j = rem - i * E::NP;
    s_dw[(r * E::NP + j) * E::NP + i] = a.dw[t];
  }
  for (int t = threadIdx.x; t < E::NP * E::NFN; t += blockDim.x) {
    const int i = t / E::NFN;
    const int m = t - i * E::NFN;
    s_lift[m * E::NP + i] = a.lift[t];
  }
  for (int t = threadIdx.x; t < E::NFN; t += blockDim.x) s_fmask[t] = a.fmask[t];
}

// PT: a pointer to const double, generic or global-qualified
template <int DIM, int NSP, typename PT>
__device__ __forceinline__ void load_q(Node<DIM, NSP>& s, PT q, int64_t stride, int64_t g) {
#pragma unroll
  for (int f = 0; f < Node<DIM, NSP>::NFLD; ++f) s.q[f] = q[f * stride + g];
}

template <int DIM, int NSP, typename PT>
__device__ __forceinline__ void load_grad(double (*gq)[DIM], double* gt, PT grad,
                                          int64_t stride, int64_t g) {
  constexpr int NFLD = 2 + DIM + NSP;
#pragma unroll
  for (int f = 0; f < NFLD; ++f)
#pragma unroll
    for (int d = 0; d < DIM; ++d) gq[f][d] = grad[(f * DIM + d) * stride + g];
#pragma unroll
  for (int d = 0; d < DIM; ++d) gt[d] = grad[(NFLD * DIM + d) * stride + g];
}

// {{{ packed element operators
//
// The contraction phase assigns one (element, field) column per lane and
// splits the element's output rows over the kWaves waves of a workgroup.
// Every lane of a wave then needs the same operator entry at the same time,
// so the operators are read with wave-uniform addresses (scalar loads into
// SGPRs) from a packed copy in which each wave's rows are contiguous:
//   dvol [w][j][r][rr]     Dw_r[i][j]       (divergence and gradient volume)
//   lift [w][m][rr]        LIFT[i][m]
// with i = w*RPW + rr (rows >= NP are zero).  The workspace holds the FP64
// copy followed by an FP32 copy for the mixed-precision kernels.

constexpr int kWaves = kBlock / 64;

template <int DIM, int P>
struct Ops {
  using E = Elem<DIM, P>;
  static constexpr int RPW = (E::NP + kWaves - 1) / kWaves;
  static constexpr int DVOL = kWaves * E::NP * DIM * RPW;
  static constexpr int LIFT = kWaves * E::NFN * RPW;
  static constexpr int TOTAL = DVOL + LIFT;
  // MFMA A-fragments (FP32, v_mfma_f32_16x16x4f32) of the same operators for
  // the mixed-precision kernels: 16-row tiles of Dw_r and of each face's
  // LIFT block, K in steps of 4, 64 floats per fragment in lane order
  // (lane l holds A[l % 16][l / 16]); rows >= NP and k past the end are 0.
  static constexpr int MT = (E::NP + 15) / 16;
  static constexpr int KSV = (E::NP + 3) / 4;
  static constexpr int KSF = (E::NFP + 3) / 4;
  static constexpr int MFV = DIM * MT * KSV * 64;
  static constexpr int MFL = E::NF * MT * KSF * 64;
  static constexpr int MF = MFV + MFL;
  // bytes: TOTAL doubles + TOTAL floats + MF floats
  static constexpr int BYTES = TOTAL * 12 + MF * 4;
};

template <int DIM, int P>
__global__ __launch_bounds__(kBlock) void pack_ops_kernel(const double* dw,
                                                          const double* lift,
                                                          double* ws) {
  using E = Elem<DIM, P>;
  using O = Ops<DIM, P>;
  constexpr int NP = E::NP, NFN = E::NFN, RPW = O::RPW;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < O::DVOL) {
    const int rr = t % RPW;
    const int r = (t / RPW) % DIM;
    const int j = (t / (RPW * DIM)) % NP;
    const int w = t / (RPW * DIM * NP);
    const int i = w * RPW + rr;
    v = i < NP ? dw[(r * NP + i) * NP + j] : 0.0;
  } else if (t < O::TOTAL) {
    const int u = t - O::DVOL;
    const int rr = u % RPW;
    const int m = (u / RPW) % NFN;
    const int w = u / (RPW * NFN);
    const int i = w * RPW + rr;
    v = i < NP ? lift[i * NFN + m] : 0.0;
  }
  if (t < O::TOTAL) {
    ws[t] = v;
    reinterpret_cast<float*>(ws + O::TOTAL)[t] = float(v);
  } else if (t < O::TOTAL + O::MF) {
    constexpr int NFP = E::NFP, MT = O::MT, KSV = O::KSV, KSF = O::KSF;
    int u = t - O::TOTAL;
    const int lane = u % 64;
    const int i0 = lane % 16, k0 = lane / 16;
    double a = 0.0;
    if (u < O::MFV) {
      const int frag = u / 64;
      const int st = frag % KSV, mt = (frag / KSV) % MT, r = frag / (KSV * MT);
      const int i = mt * 16 + i0, j = st * 4 + k0;
      if (i < NP && j < NP) a = dw[(r * NP + i) * NP + j];
    } else {
      const int frag = (u - O::MFV) / 64;
      const int st = frag % KSF, mt = (frag / KSF) % MT, f = frag / (KSF * MT);
      const int i = mt * 16 + i0, m = st * 4 + k0;
      if (i < NP && m < NFP) a = lift[i * NFN + f * NFP + m];
    }
    (reinterpret_cast<float*>(ws + O::TOTAL) + O::TOTAL)[u] = float(a);
  }
}

// MFMA A-fragments of the packed operators (see Ops).
template <int DIM, int P>
__device__ __forceinline__ const float* mfma_ops(const double* ws) {
  return reinterpret_cast<const float*>(ws + Ops<DIM, P>::TOTAL) + Ops<DIM, P>::TOTAL;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename R, int DIM, int P>
__device__ __forceinline__ const R* packed_ops(const double* ws) {
  if constexpr (sizeof(R) == 8) {
    return reinterpret_cast<const R*>(ws);
  } else {
    return reinterpret_cast<const R*>(ws + Ops<DIM, P>::TOTAL);
  }
}

// Coalesced write-out of a block's staged results: stage[plane][le*NP + i]
// -> out[plane*N + (e0 + le)*NP + i].
template <int NP, typename S, typename T>
__device__ __forceinline__ void write_staged(const S* stage, int nplanes,
                                             int epb, int pstride, int64_t e0,
                                             int64_t K, T* __restrict__ out) {
  const int64_t N = K * NP;
  const int64_t nvalid = (K - e0 < epb ? K - e0 : epb) * NP;
  const int per = epb * NP;
  for (int t = threadIdx.x; t < nplanes * per; t += blockDim.x) {
    const int plane = t / per;
    const int idx = t - plane * per;
    if (idx < nvalid) out[plane * N + e0 * NP + idx] = T(stage[plane * pstride + idx]);
  }
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (blockIdx.x % 8 picks the XCD), each with its own L2.  Consecutive
// element blocks are mesh neighbours (elements are ordered cube by cube), so
// the logical block index gives every XCD one contiguous range of the mesh:
// the neighbour gathers of a block then mostly hit the L2 that loaded the
// neighbouring blocks.  A bijection on [0, gridDim.x).
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t q = nb / kXcds, r = nb % kXcds;
  const int64_t x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

// Block visiting order.  A kernel taking EPB elements per workgroup has
// ceil(K / EPB) element blocks; with a traversal order of the elements
// (mg_mesh::elem_rank) the workgroup that xcd_block() numbers b takes the
// block whose first element is the b-th in that order.  Since xcd_block()
// hands each XCD a contiguous range of b, each XCD then works through a
// compact region of the mesh, block after neighbouring block, and the
// gathers of face neighbours find their lines in its L2 instead of
// fetching them again.  Layout (int32): K entries of scratch used while
// building, then one table per EPB of kOrderEpb.
constexpr int kOrderEpb[] = {4, 5, 6, 7, 8, 9, 10, 12, 16};
constexpr int kOrderTables = sizeof(kOrderEpb) / sizeof(int);

constexpr int order_slot(int epb) {
  for (int i = 0; i < kOrderTables; ++i)
    if (kOrderEpb[i] == epb) return i;
  return -1;
}

__host__ __device__ inline int64_t order_table_offset(int64_t K, int slot) {
  int64_t off = K;
  for (int i = 0; i < slot; ++i) off += (K + kOrderEpb[i] - 1) / kOrderEpb[i];
  return off;
}

template <int EPB, typename A>
__device__ __forceinline__ int64_t visit_block(const A& a, int64_t blk) {
  constexpr int slot = order_slot(EPB);
  if constexpr (slot < 0) {
    return blk;
  } else {
    if (!a.border) return blk;
    return a.border[order_table_offset(a.K, slot) + blk];
  }
}

// Building a table: mark[r] = the block whose first element has rank r
// (-1 elsewhere), then the marked entries in rank order.
static __global__ __launch_bounds__(kBlock) void order_fill_kernel(int32_t* mark, int64_t K) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < K) mark[i] = -1;
}

static __global__ __launch_bounds__(kBlock) void order_mark_kernel(const int32_t* rank, int32_t* mark,
                                                            int64_t K, int epb) {
  const int64_t nb = (K + epb - 1) / epb;
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int32_t r = rank[b * epb];
  if (r >= 0 && r < K) mark[r] = int32_t(b);
}

// One workgroup of 1024: a running compaction of mark into out; a table
// that does not come out with exactly ceil(K / epb) entries (a rank array
// that is not a permutation) is replaced by the natural order.
static __global__ __launch_bounds__(1024) void order_compact_kernel(const int32_t* mark, int32_t* out,
                                                             int64_t K, int epb) {
  __shared__ int s_wave[16];
  __shared__ int64_t s_base;
  const int64_t nb = (K + epb - 1) / epb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int64_t i0 = 0; i0 < K; i0 += 1024) {
    const int64_t i = i0 + threadIdx.x;
    const int32_t v = i < K ? mark[i] : -1;
    const bool keep = v >= 0;
    const uint64_t bal = __ballot(keep);
    const int before = __popcll(bal & ((uint64_t(1) << lane) - 1));
    if (lane == 0) s_wave[w] = __popcll(bal);
    __syncthreads();
    int64_t pos = s_base;
    for (int j = 0; j < w; ++j) pos += s_wave[j];
    pos += before;
    if (keep && pos < nb) out[pos] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = s_base;
      for (int j = 0; j < 16; ++j) t += s_wave[j];
      s_base = t;
    }
    __syncthreads();
  }
  if (s_base != nb)
    for (int64_t b = threadIdx.x; b < nb; b += 1024) out[b] = int32_t(b);
}

// LDS row strides are padded to an odd number of elements: rows written by
// consecutive lanes then fall into different banks (an even stride of 60-64
// doubles maps whole waves onto a few banks).
constexpr int odd_stride(int n) { return n | 1; }

// Reference pressure of the mixed-precision path: the pressure at the first
// node of the mesh.  Subtracting a constant p_ref from the momentum flux
// diagonal does not change the DG divergence (the weak divergence of a
// constant flux cancels its lift exactly), but it removes the large
// low-Mach background from the values that are rounded to FP32.  Only the
// pressure is formed (ideal gas: (gamma - 1) rho e; mixture: R T sum_k
// rho Y_k / W_k), so that every thread can afford it.
template <int DIM, int NSP, int MIX, typename A>
__device__ __forceinline__ double reference_pressure(const A& a) {
  const int64_t N = a.nnodes;
  if constexpr (MIX == 0) {
    const double rho = a.q[0];
    double mm = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) mm += a.q[(2 + d) * N] * a.q[(2 + d) * N];
    return (a.gas.gamma - 1.0) * (a.q[N] - 0.5 * mm / rho);
  } else {
    static_assert(NSP == MIX_NSP, "mixture path is the 7-species mechanism");
    double rs = 0.0;
#pragma unroll
    for (int k = 0; k < MIX_NSP; ++k) rs += a.q[(2 + DIM + k) * N] * c_mech.inv_wt[k];
    return c_mech.gas_constant * rs * a.temp[0];
  }
}

// Mengaldo boundary nodes (boundary.py:391-793): gradient values CV_bc and
// T_bc, and the viscous flux F_V(state_bc, grad_cv_bc, grad_t_bc) . n with
// the boundary conditions on the gradients applied to the interior ones.
constexpr int kBcMengaldo = 1, kBcSlip = 2, kBcImpermeable = 4, kBcAdiabatic = 8;

template <typename A>
__device__ __forceinline__ int wall_flags(const A& a, int64_t b) {
  return a.wflag ? a.wflag[b] : 0;
}

// Effective exterior viscous normal flux 2 F_V(bc).n - F_V(q-, grad-).n: the
// face flux averages both sides, which then gives F_V(bc).n.  gq/gt hold the
// interior gradients on entry and the boundary-condition gradients on exit.
template <int DIM, int NSP, int MIX, typename A>
__device__ void wall_viscous_flux_n(const A& a, int64_t b, int64_t gm, int fl,
                                    const double* n, double (*gq)[DIM], double* gt,
                                    double* fv) {
  constexpr int NFLD = 2 + DIM + NSP;
  Node<DIM, NSP> s;
  load_q<DIM, NSP>(s, a.q, a.nnodes, gm);
  node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.temp[gm] : 0.0);
  double fm[NFLD];
  viscous_flux_n<DIM, NSP, MIX>(s, gq, gt, n, fm);
  if (fl & kBcSlip) {
    // grad v_bc = G - (G n) n^T + n (n^T G n) n^T (boundary.py:159-173 in
    // frame-free form); grad(rho v)_bc = rho_bc grad v_bc + v_bc grad rho-
    const double rho_bc = a.wq[b];
    double gv[DIM][DIM], gn[DIM], ngn = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      gn[i] = 0.0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        gv[i][j] = (gq[2 + i][j] - s.v[i] * gq[0][j]) * s.inv_rho;
        gn[i] += gv[i][j] * n[j];
      }
      ngn += n[i] * gn[i];
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      const double vbc = a.wq[(2 + i) * a.nbnd + b] / rho_bc;
#pragma unroll
      for (int j = 0; j < DIM; ++j)
        gq[2 + i][j] = rho_bc * (gv[i][j] - gn[i] * n[j] + n[i] * ngn * n[j]) +
                       vbc * gq[0][j];
    }
  }
  if (fl & kBcImpermeable) {
    // grad Y_bc = grad Y - (grad Y . n) n (boundary.py:196-213)
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      double gy[DIM], gyn = 0.0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        gy[j] = (gq[2 + DIM + k][j] - s.y[k] * gq[0][j]) * s.inv_rho;
        gyn += gy[j] * n[j];
      }
#pragma unroll
      for (int j = 0; j < DIM; ++j)
        gq[2 + DIM + k][j] = s.q[0] * (gy[j] - gyn * n[j]) + s.y[k] * gq[0][j];
    }
  }
  if (fl & kBcAdiabatic) {
    double tn = 0.0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) tn += gt[j] * n[j];
#pragma unroll
    for (int j = 0; j < DIM; ++j) gt[j] -= tn * n[j];
  }
  load_q<DIM, NSP>(s, a.wq, a.nbnd, b);
  node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.wts[b] : 0.0);
  viscous_flux_n<DIM, NSP, MIX>(s, gq, gt, n, fv);
#pragma unroll
  for (int k = 0; k < NFLD; ++k) fv[k] = 2.0 * fv[k] - fm[k];
}

// }}}

// {{{ divergence (Euler / Navier-Stokes RHS)
//
// Net flux F = F_I - F_V per node, by mode:
//   kEuler  F_I(q) evaluated here
//   kGrad   F_I(q) - F_V(q, grad q, grad T) from given gradients
//   kFlux   read from the nodal-flux array written by grad_kernel<FLUX>
// Face flux from the net normal fluxes Fn = F.n of both sides:
//   Rusanov / central  (Fn- + Fn+ + lam (q- - q+)) / 2
//   HLL                HLL(F_I- . n, F_I+ . n) - (F_V- . n + F_V+ . n) / 2
// R is the type of the LDS buffers and of the contraction (double, or float
// for the mixed-precision path, which only exists for kFlux).  With float,
// each element's fluxes are stored relative to the element's first node
// (the weak divergence of that constant cancels its lift), so that FP32
// holds small variations; face fluxes are formed in FP64 and shifted alike.

constexpr int kEuler = 0, kGrad = 1, kFlux = 2;

template <int DIM, int P, int NSP, int MIX, int MODE, typename R>
struct RhsCfg {
  using E = Elem<DIM, P>;
  static constexpr int NFLD = 2 + DIM + NSP;
  static constexpr int NODEF = NFLD * DIM;        // net flux
  static constexpr int NODEP = odd_stride(NODEF);  // LDS node stride
  // LDS of a workgroup of epb elements (the few-field FP64 kernels alias the
  // face buffer, below, and never take two column groups)
  static constexpr size_t lds_bytes(int epb) {
    return sizeof(R) * (size_t(epb) * E::NP * NODEP + size_t(E::NFN) * odd_stride(epb * NFLD)) +
           sizeof(double) * (size_t(epb) * DIM * DIM + (sizeof(R) == 4 ? epb * NODEF : 0) +
                             size_t(epb) * E::NP) +
           sizeof(int) * E::NFN;
  }
  // (element, field) columns of the contraction: one per lane of a wave, or,
  // for the mixed-precision 12-field mixture kernels, two per lane: twice
  // the elements per workgroup, so the node and face phases keep twice the
  // lanes busy, where the LDS still holds the five workgroups per CU that
  // their registers allow (p <= 2; the FP64 buffers would halve them)
  static constexpr int CG =
      (NFLD > 6 && sizeof(R) == 4 && 5 * lds_bytes(128 / NFLD) <= 160 * 1024) ? 2 : 1;
  static constexpr int EPB = 64 * CG / NFLD;
  static constexpr int COLS = EPB * NFLD;
  static constexpr int FROW = odd_stride(COLS);    // LDS face-row stride
  static constexpr int SP = odd_stride(EPB * E::NP);  // staging plane stride
  static constexpr int SF = EPB * E::NP * NODEP;
  static constexpr int SFACE = E::NFN * FROW;
  static constexpr int SGEO = EPB * DIM * DIM;
  static constexpr int SREF = sizeof(R) == 4 ? EPB * NODEF : 0;
  static constexpr int SWS = EPB * E::NP;          // wavespeed |v| + c per node
  static constexpr int STAGE = NFLD * SP;
  static_assert(STAGE <= SF, "staging aliases the flux buffer");
  // FP64 with few fields: the face fluxes are held in registers until the
  // volume contraction has read the nodal fluxes, then written over them,
  // which cuts the LDS per workgroup by the face buffer (the occupancy of
  // these kernels is set by LDS otherwise).
  static constexpr bool ALIAS = sizeof(R) == 8 && NFLD <= 6;
  static constexpr int SFF = ALIAS ? (SF > SFACE ? SF : SFACE) : SF + SFACE;
  static constexpr size_t LDS = sizeof(R) * SFF + sizeof(double) * (SGEO + SREF + SWS) +
                                sizeof(int) * E::NFN;
  // waves per SIMD the register allocation aims at: 5 (96 VGPRs) for the
  // mixed-precision many-field kernels, which come out a few registers
  // above that otherwise; the compiler's choice elsewhere (forcing 5 on the
  // FP64 five-field kernels spills and runs slower)
  static constexpr int WAVES = (sizeof(R) == 4 && NFLD > 6 && MODE == kFlux) ? 5 : 1;
  static_assert(EPB >= 1, "too many fields");
  static_assert(COLS <= 64 * CG, "columns exceed the lanes' column groups");
  static_assert(!(ALIAS && CG > 1), "aliased face buffer with two column groups");
  static_assert(LDS <= 160 * 1024, "LDS per workgroup exceeds 160 KiB");
};

// HLL selects the HLL face flux (compiled separately: its two full face
// states cost registers that the Rusanov / central path does not need).
template <int DIM, int P, int NSP, int MIX, int MODE, typename R, bool HLL, typename A>
__device__ __forceinline__ void rhs_body(const A& a, const double* __restrict__ ws,
                                         double* __restrict__ out, int64_t blk) {
  using E = Elem<DIM, P>;
  using O = Ops<DIM, P>;
  using C = RhsCfg<DIM, P, NSP, MIX, MODE, R>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF, NFN = E::NFN;
  constexpr int NFLD = C::NFLD, EPB = C::EPB, COLS = C::COLS;
  constexpr int RPW = O::RPW, NODEF = C::NODEF, NODEP = C::NODEP;
  constexpr int FROW = C::FROW, SP = C::SP;
  constexpr bool VISC = MODE == kGrad;
  constexpr bool MIXED = sizeof(R) == 4;
  static_assert(!MIXED || MODE != kGrad, "mixed precision: Euler and the fused NS path");
  extern __shared__ double smem[];
  double* s_geo = smem;                                   // [EPB][DIM][DIM]
  double* s_ref = s_geo + C::SGEO;                        // [EPB][NODEF] (mixed)
  double* s_ws = s_ref + C::SREF;                         // [EPB][NP]
  R* s_f = reinterpret_cast<R*>(s_ws + C::SWS);           // [EPB][NP][NFLD][DIM]
  R* s_face = C::ALIAS ? s_f : s_f + C::SF;              // [NFN][FROW]
  int* s_fmask = reinterpret_cast<int*>(s_f + C::SFF);
  constexpr bool ALIAS = C::ALIAS;
  constexpr int NIT = (EPB * NFN + kBlock - 1) / kBlock;
  R hold[ALIAS ? NIT : 1][ALIAS ? NFLD : 1];  // face fluxes until s_f is free
  int hold_at[ALIAS ? NIT : 1];
  const R* ops = packed_ops<R, DIM, P>(ws);

  const int64_t N = a.K * NP;
  const int64_t e0 = visit_block<EPB>(a, blk) * EPB;
  const int64_t nel = a.K - e0 < EPB ? a.K - e0 : EPB;
  const int64_t g_lo = e0 * NP, g_hi = (e0 + nel) * NP;
  const bool inv = MODE == kEuler || a.inviscid_on;
  // FP32 fluxes exclude p_ref I (the stored net fluxes of the NS path, and
  // the Euler fluxes formed here)
  const double p_ref = MIXED ? reference_pressure<DIM, NSP, MIX>(a) : 0.0;
  for (int t = threadIdx.x; t < NFN; t += blockDim.x) s_fmask[t] = a.fmask[t];
  for (int t = threadIdx.x; t < EPB * DIM * DIM; t += blockDim.x)
    s_geo[t] = t < nel * DIM * DIM ? a.rx[e0 * DIM * DIM + t] : 0.0;

  // phase 1: net flux at the volume nodes (coalesced; fixed trip count)
#pragma unroll
  for (int it = 0; it < (EPB * NP + kBlock - 1) / kBlock; ++it) {
    const int t = int(threadIdx.x) + it * kBlock;
    if (t >= EPB * NP) continue;
    const int le = t / NP;
    const int i = t - le * NP;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int64_t g = e * NP + i;
    R* fdst = s_f + t * NODEP;
    if constexpr (MODE == kFlux) {
      Node<DIM, NSP> s;
      load_q<DIM, NSP>(s, a.q, N, g);
      node_state<DIM, NSP, MIX, false>(s, a.gas, MIX ? a.temp[g] : 0.0);
      s_ws[t] = s.speed + s.c;
      if constexpr (MIXED) {
        // relative to the element's first node: every thread of the element
        // loads that node's values (cache hits after the first), node 0's
        // thread keeps them in s_ref (exact in FP64) for the face phase,
        // which reads them after the barrier below
        const int64_t g0 = e * NP;
        float r[NODEF];
#pragma unroll
        for (int k = 0; k < NODEF; ++k) r[k] = a.fnet32[k * N + g0];
#pragma unroll
        for (int k = 0; k < NODEF; ++k) fdst[k] = a.fnet32[k * N + g] - r[k];
        if (i == 0) {
#pragma unroll
          for (int k = 0; k < NODEF; ++k) s_ref[le * NODEF + k] = double(r[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < NODEF; ++k) fdst[k] = a.fnet[k * N + g];
      }
    } else {
      Node<DIM, NSP> s;
      load_q<DIM, NSP>(s, a.q, N, g);
      node_state<DIM, NSP, MIX, VISC>(s, a.gas, MIX ? a.temp[g] : 0.0);
      s_ws[t] = s.speed + s.c;
      double F[NFLD][DIM];
      if (inv) {
        inviscid_flux<DIM, NSP>(s, F);
      } else {
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
#pragma unroll
          for (int d = 0; d < DIM; ++d) F[f][d] = 0.0;
      }
      if constexpr (VISC) {
        double gq[NFLD][DIM], gt[DIM], FV[NFLD][DIM];
        load_grad<DIM, NSP>(gq, gt, a.grad, N, g);
        viscous_flux<DIM, NSP, MIX>(s, gq, gt, FV);
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
#pragma unroll
          for (int d = 0; d < DIM; ++d) F[f][d] -= FV[f][d];
      }
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          fdst[f * DIM + d] = R(F[f][d] - ((MIXED && f == 2 + d) ? p_ref : 0.0));
    }
  }
  __syncthreads();
  // REL: s_f holds each element's fluxes relative to its first node (the
  // stored FP32 net fluxes); the Euler kernel keeps its FP32 fluxes absolute
  // and subtracts the first node's in the contraction (in FP32, where the
  // values are already rounded).  Either way s_ref holds the reference.
  constexpr bool REL = MIXED && MODE == kFlux;
  if constexpr (MIXED && MODE == kEuler) {
    for (int t = threadIdx.x; t < EPB * NODEF; t += blockDim.x) {
      const int le = t / NODEF;
      const int k = t - le * NODEF;
      s_ref[t] = le < nel ? double(s_f[(le * NP) * NODEP + k]) : 0.0;
    }
    __syncthreads();
  }

  // phase 2: numerical face fluxes (FP64).  The own side and neighbours
  // inside this workgroup come from LDS; other neighbours are gathered here.
  constexpr bool hll = HLL;
  // With many fields (the 12-field mixture) the Rusanov / central face flux
  // is formed field by field, without per-field arrays that would hold ~70
  // doubles per thread and cut occupancy to one wave per SIMD; with five
  // fields the arrays are cheap and let all loads issue up front.
  constexpr bool STREAM = !HLL && NFLD > 6;
  static_assert(!(STREAM && ALIAS), "streamed face fluxes are written to LDS directly");
  // fixed trip count, unrolled: the gathers of a thread's face nodes can
  // issue together
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int t = int(threadIdx.x) + it * kBlock;
    if constexpr (ALIAS) hold_at[it] = -1;
    if (t >= EPB * NFN) continue;
    const int le = t / NFN;
    const int fn = t - le * NFN;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int f = fn / NFP;
    const int im = s_fmask[fn];
    const int32_t nb = a.vmap_p[e * NFN + fn];
    double n[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + f) * DIM + d];
    const double fs = a.fscale[e * NF + f];
    const int lm = le * NP + im;
    if constexpr (!STREAM) {
      double fnm[NFLD], fnp[NFLD], qm[NFLD], qp[NFLD], shift[NFLD];
  #pragma unroll
      for (int k = 0; k < NFLD; ++k) {
        double c = 0.0, sh = 0.0;
  #pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double r0 = MIXED ? s_ref[le * NODEF + k * DIM + d] : 0.0;
          c += (double(s_f[lm * NODEP + k * DIM + d]) + (REL ? r0 : 0.0)) * n[d];
          sh += r0 * n[d];
        }
        fnm[k] = c;
        shift[k] = sh;
      }
      const int64_t gm = e * NP + im;
      const double tm = MIX ? a.temp[gm] : 0.0;
      const double wsm = s_ws[lm];
  #pragma unroll
      for (int k = 0; k < NFLD; ++k) qm[k] = a.q[k * N + gm];
      double tp, wsp;
      if (nb >= g_lo && nb < g_hi) {
        const int lp = int(nb - g_lo);
        const int ple = lp / NP;
        tp = MIX ? a.temp[nb] : 0.0;
        wsp = s_ws[lp];
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) qp[k] = a.q[k * N + nb];
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) {
          double c = 0.0;
  #pragma unroll
          for (int d = 0; d < DIM; ++d)
            c += (double(s_f[lp * NODEP + k * DIM + d]) +
                  (REL ? s_ref[ple * NODEF + k * DIM + d] : 0.0)) * n[d];
          fnp[k] = c;
        }
      } else {
        const bool bnd = nb < 0;
        const int64_t b = bnd ? int64_t(-nb - 1) : 0;
        Node<DIM, NSP> sp;
        if (!bnd) {
          load_q<DIM, NSP>(sp, a.q, N, nb);
          tp = MIX ? a.temp[nb] : 0.0;
        } else {
          load_q<DIM, NSP>(sp, a.qb, a.nbnd, b);
          tp = MIX ? a.tb[b] : 0.0;
        }
        node_state<DIM, NSP, MIX, VISC>(sp, a.gas, tp);
        wsp = sp.speed + sp.c;
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) qp[k] = sp.q[k];
        if constexpr (MODE == kFlux) {
          if (!bnd) {
  #pragma unroll
            for (int k = 0; k < NFLD; ++k) {
              double c = 0.0;
  #pragma unroll
              for (int d = 0; d < DIM; ++d)
                c += (MIXED ? double(a.fnet32[(k * DIM + d) * N + nb])
                            : a.fnet[(k * DIM + d) * N + nb]) * n[d];
              fnp[k] = c;
            }
          } else {
  #pragma unroll
            for (int k = 0; k < NFLD; ++k) fnp[k] = a.fnb[k * a.nbnd + b];
          }
        } else {
          if (inv) {
            inviscid_flux_n<DIM, NSP>(sp, n, fnp);
            if constexpr (MIXED) {
  #pragma unroll
              for (int d = 0; d < DIM; ++d) fnp[2 + d] -= p_ref * n[d];
            }
          } else {
  #pragma unroll
            for (int k = 0; k < NFLD; ++k) fnp[k] = 0.0;
          }
          if constexpr (VISC) {
            double gq[NFLD][DIM], gt[DIM], fv[NFLD];
            load_grad<DIM, NSP>(gq, gt, a.grad, N, bnd ? e * NP + im : int64_t(nb));
            if (bnd && (wall_flags(a, b) & kBcMengaldo)) {
              // precomputed by wall_flux_kernel
  #pragma unroll
              for (int k = 0; k < NFLD; ++k) fv[k] = a.wwork[k * a.nbnd + b];
            } else {
              viscous_flux_n<DIM, NSP, MIX>(sp, gq, gt, n, fv);
            }
  #pragma unroll
            for (int k = 0; k < NFLD; ++k) fnp[k] -= fv[k];
          }
        }
      }
      double fl[NFLD];
      if constexpr (hll) {
        Node<DIM, NSP> sm, sp;
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) {
          sm.q[k] = qm[k];
          sp.q[k] = qp[k];
        }
        node_state<DIM, NSP, MIX, false>(sm, a.gas, tm);
        node_state<DIM, NSP, MIX, false>(sp, a.gas, tp);
        double fim[NFLD], fip[NFLD];
        inviscid_flux_n<DIM, NSP>(sm, n, fim);
        inviscid_flux_n<DIM, NSP>(sp, n, fip);
        inviscid_numflux<DIM, NSP>(sm, sp, n, 1, fl);
        // stored net fluxes exclude p_ref I: F_V.n = F_I.n - (Fn + p_ref n)
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) {
          const double pn = (k >= 2 && k < 2 + DIM) ? p_ref * n[k - 2] : 0.0;
          fl[k] -= ((fim[k] - fnm[k] - pn) + (fip[k] - fnp[k] - pn)) / 2 + pn;
        }
      } else {
        const double lam = (inv && a.flux_kind == 0) ? fmax(wsm, wsp) : 0.0;
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) fl[k] = (fnm[k] + fnp[k] + lam * (qm[k] - qp[k])) / 2;
      }
      if constexpr (ALIAS) {
        hold_at[it] = fn * FROW + le * NFLD;
  #pragma unroll
        for (int k = 0; k < NFLD; ++k) hold[it][k] = R(-(fs * (fl[k] - shift[k])));
      } else {
  #pragma unroll
        for (int k = 0; k < NFLD; ++k)
          s_face[fn * FROW + le * NFLD + k] = R(-(fs * (fl[k] - shift[k])));
      }
    } else {
      // Rusanov / central, field by field
      const int64_t gm = e * NP + im;
      const double wsm = s_ws[lm];
      const bool local = nb >= g_lo && nb < g_hi;
      const bool bnd = nb < 0;
      const int64_t b = bnd ? int64_t(-nb - 1) : 0;
      const auto qsrc = bnd ? a.qb : a.q;
      const int64_t qstr = bnd ? a.nbnd : N;
      const int64_t qix = bnd ? b : int64_t(nb);
      const int lp = local ? int(nb - g_lo) : 0;
      const int ple = lp / NP;
      double wsp;
      Node<DIM, NSP> sp;
      double fv[VISC ? NFLD : 1];
      if (local) {
        wsp = s_ws[lp];
      } else {
        load_q<DIM, NSP>(sp, qsrc, qstr, qix);
        node_state<DIM, NSP, MIX, VISC>(sp, a.gas,
                                        MIX ? (bnd ? a.tb[b] : a.temp[nb]) : 0.0);
        wsp = sp.speed + sp.c;
        if constexpr (VISC) {
          double gq[NFLD][DIM], gt[DIM];
          load_grad<DIM, NSP>(gq, gt, a.grad, N, bnd ? e * NP + im : int64_t(nb));
          if (bnd && (wall_flags(a, b) & kBcMengaldo)) {
            // precomputed by wall_flux_kernel
#pragma unroll
            for (int k = 0; k < NFLD; ++k) fv[k] = a.wwork[k * a.nbnd + b];
          } else {
            viscous_flux_n<DIM, NSP, MIX>(sp, gq, gt, n, fv);
          }
        }
      }
      double vnp = 0.0, mnp = 0.0;
      if (!local) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          vnp += sp.v[d] * n[d];
          mnp += sp.q[2 + d] * n[d];
        }
      }
      const double lam = (inv && a.flux_kind == 0) ? fmax(wsm, wsp) : 0.0;
#pragma unroll
      for (int k = 0; k < NFLD; ++k) {
        double fm = 0.0, sh = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double r0 = MIXED ? s_ref[le * NODEF + k * DIM + d] : 0.0;
          fm += (double(s_f[lm * NODEP + k * DIM + d]) + (REL ? r0 : 0.0)) * n[d];
          sh += r0 * n[d];
        }
        const double qmk = a.q[k * N + gm];
        const double qpk = local ? a.q[k * N + nb] : sp.q[k];
        double fp = 0.0;
        if (local) {
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            fp += (double(s_f[lp * NODEP + k * DIM + d]) +
                   (REL ? s_ref[ple * NODEF + k * DIM + d] : 0.0)) * n[d];
        } else if constexpr (MODE == kFlux) {
          if (bnd) {
            fp = a.fnb[k * a.nbnd + b];
          } else {
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              fp += (MIXED ? double(a.fnet32[(k * DIM + d) * N + nb])
                           : a.fnet[(k * DIM + d) * N + nb]) * n[d];
          }
        } else {
          if (inv) {
            // component k of inviscid_flux_n (mgdg_physics.h)
            if (k == 0) fp = mnp;
            else if (k == 1) fp = vnp * (sp.q[1] + sp.p);
            else if (k < 2 + DIM) fp = sp.q[0] * sp.v[k - 2] * vnp + (sp.p - p_ref) * n[k - 2];
            else fp = vnp * sp.q[k];
          }
          if constexpr (VISC) fp -= fv[k];
        }
        const double flk = (fm + fp + lam * (qmk - qpk)) / 2;
        s_face[fn * FROW + le * NFLD + k] = R(-(fs * (flk - sh)));
      }
    }
  }
  if constexpr (!ALIAS) __syncthreads();

  // phase 3: rhs[i] = sum_j sum_r Dw_r[i][j] G_r[j] + sum_m LIFT[i][m] B[m]
  // per (element, field) column, G_r = sum_d rx[r][d] F_d; rows split over
  // the waves, operator entries wave-uniform
  // (each lane holds CG columns, c and c + 64)
  constexpr int CG = C::CG;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = threadIdx.x & 63;
  int cc[CG], cle[CG], cf[CG];
#pragma unroll
  for (int g = 0; g < CG; ++g) {
    cc[g] = c + 64 * g < COLS ? c + 64 * g : 0;
    cle[g] = cc[g] / NFLD;
    cf[g] = cc[g] - cle[g] * NFLD;
  }
  R rxl[CG][DIM][DIM];
#pragma unroll
  for (int g = 0; g < CG; ++g)
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int d = 0; d < DIM; ++d) rxl[g][r][d] = R(s_geo[cle[g] * DIM * DIM + r * DIM + d]);
  const R* DV = ops + w * (NP * DIM * RPW);
  const R* LF = ops + O::DVOL + w * (NFN * RPW);
  R acc[CG][RPW];
#pragma unroll
  for (int g = 0; g < CG; ++g)
#pragma unroll
    for (int rr = 0; rr < RPW; ++rr) acc[g][rr] = R(0);
  R f0[CG][DIM];   // the columns' references (absolute FP32 Euler fluxes)
#pragma unroll
  for (int g = 0; g < CG; ++g)
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      f0[g][d] = (MIXED && !REL) ? s_f[(cle[g] * NP) * NODEP + cf[g] * DIM + d] : R(0);
#pragma unroll 2
  for (int j = 0; j < NP; ++j) {
#pragma unroll
    for (int g = 0; g < CG; ++g) {
      const R* fj = s_f + (cle[g] * NP + j) * NODEP + cf[g] * DIM;
      R fd[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) fd[d] = fj[d] - f0[g][d];
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        R gr = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) gr += rxl[g][r][d] * fd[d];
#pragma unroll
        for (int rr = 0; rr < RPW; ++rr) acc[g][rr] += DV[(j * DIM + r) * RPW + rr] * gr;
      }
    }
  }
  if constexpr (ALIAS) {
    // every nodal flux read: the face fluxes go over them
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NIT; ++it)
      if (hold_at[it] >= 0)
#pragma unroll
        for (int k = 0; k < NFLD; ++k) s_face[hold_at[it] + k] = hold[it][k];
    __syncthreads();
  }
#pragma unroll 4
  for (int m = 0; m < NFN; ++m) {
#pragma unroll
    for (int g = 0; g < CG; ++g) {
      const R bv = s_face[m * FROW + cc[g]];
#pragma unroll
      for (int rr = 0; rr < RPW; ++rr) acc[g][rr] += LF[m * RPW + rr] * bv;
    }
  }
  __syncthreads();
  // phase 4: stage in LDS (over s_f), write coalesced
  R* stage = s_f;                            // [NFLD][SP]
#pragma unroll
  for (int g = 0; g < CG; ++g) {
    if (c + 64 * g < COLS) {
#pragma unroll
      for (int rr = 0; rr < RPW; ++rr) {
        const int i = w * RPW + rr;
        if (i < NP) stage[cf[g] * SP + cle[g] * NP + i] = acc[g][rr];
      }
    }
  }
  __syncthreads();
  write_staged<NP>(stage, NFLD, EPB, SP, e0, a.K, out);
}

template <int DIM, int P, int NSP, int MIX, int MODE, typename R, bool HLL>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(HLL ? 1 : RhsCfg<DIM, P, NSP, MIX, MODE, R>::WAVES)))
void rhs_kernel(KArgs a,
                                                     const double* __restrict__ ws,
                                                     double* __restrict__ out) {
  rhs_body<DIM, P, NSP, MIX, MODE, R, HLL>(a, ws, out, xcd_block());
}

// }}}

// {{{ gradient (and the fused nodal net flux)
//
// grad_kernel computes the DG gradients of the conserved fields and of T
// (central flux inside, exterior value on boundary faces).  With FLUX set it
// then evaluates, at every node, the net flux F_I - F_V from the state and
// the just-computed gradients (viscous stress, heat and species fluxes are
// pointwise) and writes that instead of (or besides) the gradients; for
// boundary face nodes it also writes (F_I - F_V)(q_ext, grad q-) . n, the
// exterior side of the central viscous boundary flux.
// With R = float (mixed precision) the traces are held relative to each
// element's first node (the DG gradient of a constant is zero, and the
// central face values are shifted by the same constant), the contraction
// runs in FP32, the physics in FP64 and the nodal fluxes are stored in FP32
// with p_ref removed from the momentum-flux diagonal.

template <int DIM, int P, int NSP, int MIX, typename R, bool FLUX = false>
struct GradCfg {
  using E = Elem<DIM, P>;
  static constexpr int NS = 3 + DIM + NSP;  // conserved fields + T
  static constexpr int KG = E::NP + E::NFN;
  static constexpr int GEO = DIM * DIM + E::NF * DIM;  // rx and normals
  // LDS of a workgroup of cg * 64 columns
  static constexpr size_t lds_bytes(int cg) {
    const int epb = 64 * cg / NS;
    const size_t sb = size_t(KG) * odd_stride(64 * cg);
    const size_t stage = size_t(NS) * DIM * odd_stride(epb * E::NP);
    return sizeof(R) * (sb > stage ? sb : stage) +
           sizeof(double) * (size_t(epb) * GEO + (sizeof(R) == 4 ? epb * NS : 0)) +
           sizeof(int) * E::NFN;
  }
  // elements resident per CU: workgroups by LDS, at most those the
  // registers allow (the mixed-precision gradient pass of the 13-scalar
  // mixture with one 16-row tile, p <= 2 in 3-D: 60 VGPRs, eight waves per
  // SIMD, with one column group; 80, six waves, with two.  With more tiles
  // the second group's accumulators cost a half of the waves or more.)
  static constexpr int resident(int cg) {
    const int by_lds = int((160 * 1024) / lds_bytes(cg));
    const int by_regs = cg == 1 ? 8 : 6;
    return (64 * cg / NS) * (by_lds < by_regs ? by_lds : by_regs);
  }
  // (element, scalar) columns: 64, or 128 (two column groups per wave) for
  // the mixed-precision mixture gradients where that keeps more elements
  // resident: twice the elements per workgroup, so the node and face phases
  // keep twice the lanes busy
  static constexpr int CG = (NS > 7 && !FLUX && sizeof(R) == 4 && E::NP <= 16 &&
                             resident(2) > resident(1)) ? 2 : 1;
  static constexpr int EPB = 64 * CG / NS;
  static constexpr int COLS = 64 * CG;
  static constexpr int ROW = odd_stride(COLS);         // LDS trace-row stride
  static constexpr int SP = odd_stride(EPB * E::NP);   // gradient plane stride
  static constexpr int STAGE = NS * DIM * SP;
  static constexpr int SB = KG * ROW > STAGE ? KG * ROW : STAGE;
  static constexpr int SREF = sizeof(R) == 4 ? EPB * NS : 0;
  static constexpr size_t LDS = sizeof(R) * SB + sizeof(double) * (EPB * GEO + SREF) +
                                sizeof(int) * E::NFN;
  static_assert(EPB >= 1, "too many fields");
  static_assert(CG == 1 || sizeof(R) == 4, "two column groups: matrix-core contraction only");
  static_assert(LDS == lds_bytes(CG), "LDS layout");
  static_assert(LDS <= 160 * 1024, "LDS per workgroup exceeds 160 KiB");
};

template <int DIM, int NSP, int MIX>
__device__ __forceinline__ double node_temperature(const double* qv,
                                                   const GasDev& g, double tmix) {
  if constexpr (MIX != 0) {
    return tmix;
  } else {
    double mm = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) mm += qv[2 + d] * qv[2 + d];
    const double irho = 1.0 / qv[0];
    const double e_int = qv[1] - 0.5 * mm * irho;
    return ((g.gamma - 1.0) / g.gas_const) * e_int * irho;
  }
}

template <int DIM, int P, int NSP, int MIX, bool FLUX, typename R, typename A>
__device__ __forceinline__ void grad_body(const A& a, const double* __restrict__ ws,
                                          double* __restrict__ out, int64_t blk) {
  using E = Elem<DIM, P>;
  using O = Ops<DIM, P>;
  using C = GradCfg<DIM, P, NSP, MIX, R, FLUX>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF, NFN = E::NFN;
  constexpr int NFLD = 2 + DIM + NSP, NS = C::NS, EPB = C::EPB, COLS = C::COLS;
  constexpr int ROW = C::ROW, SP = C::SP;
  constexpr int RPW = O::RPW, GEO = C::GEO;
  constexpr bool MIXED = sizeof(R) == 4;
  extern __shared__ double smem[];
  double* s_geo = smem;                          // [EPB][GEO]
  double* s_ref = s_geo + EPB * GEO;             // [EPB][NS] (mixed)
  R* s_b = reinterpret_cast<R*>(s_ref + C::SREF);  // [NP + NFN][ROW]
  R* stage = s_b;                                // [NS][DIM][EPB*NP], after phase 3
  int* s_fmask = reinterpret_cast<int*>(s_b + C::SB);
  const R* ops = packed_ops<R, DIM, P>(ws);

  const int64_t N = a.K * NP;
  const int64_t e0 = visit_block<EPB>(a, blk) * EPB;
  const int64_t nel = a.K - e0 < EPB ? a.K - e0 : EPB;
  const int64_t g_lo = e0 * NP, g_hi = (e0 + nel) * NP;
  for (int t = threadIdx.x; t < NFN; t += blockDim.x) s_fmask[t] = a.fmask[t];
  for (int t = threadIdx.x; t < EPB * GEO; t += blockDim.x) {
    const int le = t / GEO;
    const int u = t - le * GEO;
    const int64_t e = e0 + le;
    double v = 0.0;
    if (e < a.K)
      v = u < DIM * DIM ? a.rx[e * DIM * DIM + u] : a.normals[e * NF * DIM + u - DIM * DIM];
    s_geo[t] = v;
  }
  // phase 1: conserved fields and T at the volume nodes.  With one node per
  // thread the node's state stays in registers for the flux epilogue.
  // Mixed precision stores them relative to the element's first node (q and
  // T there): every thread of the element forms that reference itself (its
  // loads hit the cache after the first), node 0's thread keeps it in s_ref
  // for the face phase, which reads it after the barrier below.
  constexpr bool KEEPQ = FLUX && EPB * NP <= kBlock;
  double qkeep[KEEPQ ? NFLD : 1];
  for (int t = threadIdx.x; t < EPB * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int64_t g = e * NP + i;
    double qv[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) qv[f] = a.q[f * N + g];
    if constexpr (KEEPQ) {
#pragma unroll
      for (int f = 0; f < NFLD; ++f) qkeep[f] = qv[f];
    }
    const double tn = node_temperature<DIM, NSP, MIX>(qv, a.gas, MIX ? a.temp[g] : 0.0);
    double ref[NFLD + 1];
    if constexpr (MIXED) {
      const int64_t g0 = e * NP;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) ref[f] = a.q[f * N + g0];
      ref[NFLD] = node_temperature<DIM, NSP, MIX>(ref, a.gas, MIX ? a.temp[g0] : 0.0);
      if (i == 0) {
#pragma unroll
        for (int f = 0; f <= NFLD; ++f) s_ref[le * NS + f] = ref[f];
      }
    } else {
#pragma unroll
      for (int f = 0; f <= NFLD; ++f) ref[f] = 0.0;
    }
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_b[i * ROW + le * NS + f] = R(qv[f] - ref[f]);
    s_b[i * ROW + le * NS + NFLD] = R(tn - ref[NFLD]);
  }
  __syncthreads();

  // phase 2: central (interior) / exterior-value (boundary) traces x fscale;
  // neighbours inside this workgroup are read from LDS (fixed trip count,
  // unrolled, so that a thread's gathers can issue together)
#pragma unroll
  for (int it = 0; it < (EPB * NFN + kBlock - 1) / kBlock; ++it) {
    const int t = int(threadIdx.x) + it * kBlock;
    if (t >= EPB * NFN) continue;
    const int le = t / NFN;
    const int fn = t - le * NFN;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int f = fn / NFP;
    const int im = s_fmask[fn];
    const int32_t nb = a.vmap_p[e * NFN + fn];
    const double fs = a.fscale[e * NF + f];
    R* dst = s_b + (NP + fn) * ROW + le * NS;
    const R* own = s_b + im * ROW + le * NS;
    const double* ref = s_ref + le * NS;
    if (nb >= g_lo && nb < g_hi) {
      const int lp = int(nb - g_lo);
      const int ple = lp / NP;
      const R* nbv = s_b + (lp - ple * NP) * ROW + ple * NS;
      const double* pref = s_ref + ple * NS;
#pragma unroll
      for (int k = 0; k <= NFLD; ++k) {
        const double dref = MIXED ? pref[k] - ref[k] : 0.0;
        dst[k] = R(fs * ((double(nbv[k]) + dref + double(own[k])) / 2));
      }
    } else {
      double qp[NFLD];
      double tp;
      const bool bnd = nb < 0;
      if (!bnd) {
#pragma unroll
        for (int k = 0; k < NFLD; ++k) qp[k] = a.q[k * N + nb];
        tp = node_temperature<DIM, NSP, MIX>(qp, a.gas, MIX ? a.temp[nb] : 0.0);
      } else {
        const int64_t b = -int64_t(nb) - 1;
        if (wall_flags(a, b) & kBcMengaldo) {
          // CV_bc n and T_bc n (boundary.py:713-793)
#pragma unroll
          for (int k = 0; k < NFLD; ++k) qp[k] = a.wq[k * a.nbnd + b];
          tp = a.wt[b];
        } else {
#pragma unroll
          for (int k = 0; k < NFLD; ++k) qp[k] = a.qb[k * a.nbnd + b];
          tp = node_temperature<DIM, NSP, MIX>(qp, a.gas, MIX ? a.tb[b] : 0.0);
        }
      }
#pragma unroll
      for (int k = 0; k <= NFLD; ++k) {
        const double up = (k < NFLD ? qp[k] : tp) - (MIXED ? ref[k] : 0.0);
        dst[k] = R(bnd ? fs * up : fs * ((up + double(own[k])) / 2));
      }
    }
  }
  __syncthreads();

  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = threadIdx.x & 63;
  if constexpr (MIXED) {
    // phase 3 on the matrix cores (v_mfma_f32_16x16x4f32): wave w owns the
    // 16 columns [16w, 16w + 16) of each group of 64 (CG groups); per
    // 16-row tile of each output W_r = Dw_r u and L_f = LIFT_f (fs u*) the
    // K dimension (nodes) runs in steps of 4.  Lane l supplies
    // B[k = l / 16][col = l % 16] and holds D[4 (l / 16) + reg][l % 16] of
    // each tile; the groups share the operator fragments.
    using O = Ops<DIM, P>;
    constexpr int MT = O::MT, KSV = O::KSV, KSF = O::KSF, CG = C::CG;
    const float* mf = mfma_ops<DIM, P>(ws);
    const int col0 = w * 16 + (c & 15);
    const int kq = c >> 4;
    floatx4 aw[CG][DIM][MT], al[CG][NF][MT];
#pragma unroll
    for (int g = 0; g < CG; ++g)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int r = 0; r < DIM; ++r) aw[g][r][mt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int f = 0; f < NF; ++f) al[g][f][mt] = floatx4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
    for (int st = 0; st < KSV; ++st) {
      const int j = st * 4 + kq;
      float bv[CG];
#pragma unroll
      for (int g = 0; g < CG; ++g)  // A is 0 past NP
        bv[g] = float(s_b[(j < NP ? j : NP - 1) * ROW + 64 * g + col0]);
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const float av = mf[((r * MT + mt) * KSV + st) * 64 + c];
#pragma unroll
          for (int g = 0; g < CG; ++g)
            aw[g][r][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[g], aw[g][r][mt], 0, 0, 0);
        }
    }
    // faces innermost: NF * MT independent accumulations per step
#pragma unroll
    for (int st = 0; st < KSF; ++st) {
      const int m = st * 4 + kq;
      const int mm = m < NFP ? m : NFP - 1;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        float bv[CG];
#pragma unroll
        for (int g = 0; g < CG; ++g) bv[g] = float(s_b[(NP + f * NFP + mm) * ROW + 64 * g + col0]);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const float av = mf[O::MFV + ((f * MT + mt) * KSF + st) * 64 + c];
#pragma unroll
          for (int g = 0; g < CG; ++g)
            al[g][f][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[g], al[g][f][mt], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    // phase 4: stage the gradients [NS][DIM][EPB*NP] (over the trace buffer)
#pragma unroll
    for (int g = 0; g < CG; ++g) {
      const int col = 64 * g + col0;
      if (col < EPB * NS) {
        const int le = col / NS;
        const int sc = col - le * NS;
        const double* geo = s_geo + le * GEO;
        // one direction at a time: only its 7 geometric factors are live
#pragma unroll 1
        for (int d = 0; d < DIM; ++d) {
          float rxd[DIM], nfd[NF];
#pragma unroll
          for (int r = 0; r < DIM; ++r) rxd[r] = float(geo[r * DIM + d]);
#pragma unroll
          for (int f = 0; f < NF; ++f) nfd[f] = float(geo[DIM * DIM + f * DIM + d]);
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
              const int i = mt * 16 + 4 * kq + reg;
              if (i >= NP) continue;
              float vsum = 0.f, fsum = 0.f;
#pragma unroll
              for (int r = 0; r < DIM; ++r) vsum += rxd[r] * aw[g][r][mt][reg];
#pragma unroll
              for (int f = 0; f < NF; ++f) fsum += nfd[f] * al[g][f][mt][reg];
              stage[(sc * DIM + d) * SP + le * NP + i] = R(fsum - vsum);
            }
        }
      }
    }
  } else {
    // phase 3: per (element, scalar) column: W_r = Dw_r u, L_f = LIFT_f (fs u*),
    // grad_d = sum_f n_f[d] L_f - sum_r rx[r][d] W_r
    const R* GV = ops + w * (NP * DIM * RPW);
    const R* GL = ops + O::DVOL + w * (NFN * RPW);
    R accw[RPW][DIM], accl[RPW][NF];
  #pragma unroll
    for (int rr = 0; rr < RPW; ++rr) {
  #pragma unroll
      for (int r = 0; r < DIM; ++r) accw[rr][r] = R(0);
  #pragma unroll
      for (int f = 0; f < NF; ++f) accl[rr][f] = R(0);
    }
  #pragma unroll 2
    for (int j = 0; j < NP; ++j) {
      const R bv = s_b[j * ROW + c];
  #pragma unroll
      for (int r = 0; r < DIM; ++r)
  #pragma unroll
        for (int rr = 0; rr < RPW; ++rr) accw[rr][r] += GV[(j * DIM + r) * RPW + rr] * bv;
    }
  #pragma unroll
    for (int f = 0; f < NF; ++f) {
  #pragma unroll 2
      for (int m = 0; m < NFP; ++m) {
        const R bv = s_b[(NP + f * NFP + m) * ROW + c];
  #pragma unroll
        for (int rr = 0; rr < RPW; ++rr) accl[rr][f] += GL[(f * NFP + m) * RPW + rr] * bv;
      }
    }
    __syncthreads();
    // phase 4: stage the gradients [NS][DIM][EPB*NP] (over the trace buffer)
    if (c < EPB * NS) {
      const int le = c / NS;
      const int sc = c - le * NS;
      const double* geo = s_geo + le * GEO;
  #pragma unroll
      for (int d = 0; d < DIM; ++d) {
  #pragma unroll
        for (int rr = 0; rr < RPW; ++rr) {
          const int i = w * RPW + rr;
          if (i >= NP) continue;
          R vsum = R(0), fsum = R(0);
  #pragma unroll
          for (int r = 0; r < DIM; ++r) vsum += R(geo[r * DIM + d]) * accw[rr][r];
  #pragma unroll
          for (int f = 0; f < NF; ++f) fsum += R(geo[DIM * DIM + f * DIM + d]) * accl[rr][f];
          stage[(sc * DIM + d) * SP + le * NP + i] = fsum - vsum;
        }
      }
    }
  }
  __syncthreads();
  if constexpr (!FLUX) {
    if constexpr (MIXED)
      write_staged<NP>(stage, NS * DIM, EPB, SP, e0, a.K, (float*)a.grad32_out);
    else
      write_staged<NP>(stage, NS * DIM, EPB, SP, e0, a.K, out);
  } else {
    if (a.grad_out) write_staged<NP>(stage, NS * DIM, EPB, SP, e0, a.K, a.grad_out);
    const double p_ref = MIXED ? reference_pressure<DIM, NSP, MIX>(a) : 0.0;
    // phase 5: net flux F_I - F_V at every node, written coalesced (fixed
    // trip count)
#pragma unroll
    for (int it = 0; it < (EPB * NP + kBlock - 1) / kBlock; ++it) {
      const int t = int(threadIdx.x) + it * kBlock;
      if (t >= EPB * NP) continue;
      const int le = t / NP;
      const int i = t - le * NP;
      const int64_t e = e0 + le;
      if (e >= a.K) continue;
      const int64_t g = e * NP + i;
      Node<DIM, NSP> s;
      if constexpr (KEEPQ) {
#pragma unroll
        for (int f = 0; f < NFLD; ++f) s.q[f] = qkeep[f];
      } else {
        load_q<DIM, NSP>(s, a.q, N, g);
      }
      node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.temp[g] : 0.0);
      double gq[NFLD][DIM], gt[DIM], F[NFLD][DIM];
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
#pragma unroll
        for (int d = 0; d < DIM; ++d) gq[f][d] = double(stage[(f * DIM + d) * SP + t]);
#pragma unroll
      for (int d = 0; d < DIM; ++d) gt[d] = double(stage[(NFLD * DIM + d) * SP + t]);
      viscous_flux<DIM, NSP, MIX>(s, gq, gt, F);
      double FI[NFLD][DIM];
      if (a.inviscid_on) {
        inviscid_flux<DIM, NSP>(s, FI);
      } else {
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
#pragma unroll
          for (int d = 0; d < DIM; ++d) FI[f][d] = 0.0;
      }
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double v = FI[f][d] - F[f][d] - ((f == 2 + d) ? p_ref : 0.0);
          if constexpr (MIXED) {
            a.fnet32_out[(f * DIM + d) * N + g] = float(v);
          } else {
            out[(f * DIM + d) * N + g] = v;
          }
        }
    }
    // phase 6: exterior net normal flux at boundary face nodes: central
    // viscous flux with the interior gradients (boundary.py:944-970), or the
    // Mengaldo viscous flux F_V(bc) . n (boundary.py:630-710)
    if (a.nbnd > 0) {
      for (int t = threadIdx.x; t < EPB * NFN; t += blockDim.x) {
        const int le = t / NFN;
        const int fn = t - le * NFN;
        const int64_t e = e0 + le;
        if (e >= a.K) continue;
        const int32_t nb = a.vmap_p[e * NFN + fn];
        if (nb >= 0) continue;
        const int64_t b = -int64_t(nb) - 1;
        const int f = fn / NFP;
        const int im = s_fmask[fn];
        if (wall_flags(a, b) & kBcMengaldo) {
          // the interior gradients go to wall_flux_kernel, which forms this
          // node's exterior flux
          const int ti = le * NP + im;
#pragma unroll
          for (int k = 0; k < NS * DIM; ++k)
            a.wwork[k * a.nbnd + b] = double(stage[k * SP + ti]);
          continue;
        }
        double n[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) n[d] = s_geo[le * GEO + DIM * DIM + f * DIM + d];
        Node<DIM, NSP> s;
        load_q<DIM, NSP>(s, a.qb, a.nbnd, b);
        node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.tb[b] : 0.0);
        double gq[NFLD][DIM], gt[DIM], fv[NFLD], fi[NFLD];
        const int ti = le * NP + im;
#pragma unroll
        for (int k = 0; k < NFLD; ++k)
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            gq[k][d] = double(stage[(k * DIM + d) * SP + ti]);
#pragma unroll
        for (int d = 0; d < DIM; ++d) gt[d] = double(stage[(NFLD * DIM + d) * SP + ti]);
        if (a.inviscid_on) {
          inviscid_flux_n<DIM, NSP>(s, n, fi);
        } else {
#pragma unroll
          for (int k = 0; k < NFLD; ++k) fi[k] = 0.0;
        }
        viscous_flux_n<DIM, NSP, MIX>(s, gq, gt, n, fv);
#pragma unroll
        for (int k = 0; k < NFLD; ++k)
          a.fnb_out[k * a.nbnd + b] =
              fi[k] - fv[k] - ((k >= 2 && k < 2 + DIM) ? p_ref * n[k - 2] : 0.0);
      }
    }
  }
}

template <int DIM, int P, int NSP, int MIX, bool FLUX, typename R>
// The mixed-precision few-field flux kernel aims at 5 waves per SIMD (96
// VGPRs; its LDS allows 8 workgroups per CU): 0.114 -> 0.106 ms on the TGV
// p = 3 workload.  6 waves, and 5 for the mixed divergence kernel, were
// slower (spills); the FP64 flux kernel is held to 4 workgroups by LDS.
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu((FLUX && NSP == 0 && sizeof(R) == 4) ? 5 : 1)))
void grad_kernel(KArgs a,
                                                      const double* __restrict__ ws,
                                                      double* __restrict__ out) {
  grad_body<DIM, P, NSP, MIX, FLUX, R>(a, ws, out, xcd_block());
}

// }}}

// {{{ Euler RHS on the quadrature discretization (overintegration)
//
// A workgroup owns epb elements (host-chosen from the LDS budget).  LDS:
// s_q [F][epb][NP] the state (then the result), s_g [F*epb][gs] the inviscid
// flux in reference directions (rx . F) at the volume points (r-major rows
// of DIM*nq), s_f [F*epb][fs] minus the numerical flux times the face
// Jacobian ratio at the face points; the row strides gs, fs are 4 mod 32
// doubles so that the matrix-core operand reads of 16 rows spread over the
// LDS banks.  Lanes own quadrature points in phases 2-3; phase 4 is one
// (NP x K) x (K x F*epb) product on the matrix cores.

struct QuadArgs {
  int nq, nfq, epb, niter;
  int gs, fs;             // LDS row strides of s_g and s_f (doubles)
  const double* temp;     // [N] nodal temperature (mixtures: Newton seeds)
  const double* tbq;      // [nbq] exterior temperatures at boundary face points
  // operators transposed so that consecutive lanes (consecutive points in
  // phases 2-3, consecutive nodes in phase 4) read consecutive addresses
  const double* interp;   // [NP][nq]
  const double* wdiff;    // [DIM][nq][NP]
  const double* finterp;  // [NF][NFP][nfq]
  const double* lift;     // [NF*nfq][NP]
  const double* qbq;      // [F][nbq] exterior states at boundary face points
  int64_t nbq;
};

// Temperature of a mixture state at a quadrature point: the Newton
// inversion of its internal energy from an interpolated seed (the
// reference's project_fluid_state -> make_fluid_state, gas_model.py:494-588).
template <int DIM, int NSP>
__device__ __forceinline__ double point_temperature(const double* q, double seed,
                                                    int niter) {
  const double irho = 1.0 / q[0];
  double mm = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) mm += q[2 + d] * q[2 + d];
  const double e = (q[1] - 0.5 * mm * irho) * irho;
  double y[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = q[2 + DIM + k] * irho;
  return mix_temperature(e, seed, y, niter);
}

// acc0 + acc1 += A[i][0:kn) B[0:kn)[j] for one 16x16 tile on the matrix
// cores: a = &A[i][0] with k stride NP (global, read through the caches),
// b = &B[0][j] with k contiguous (LDS); kk = lane / 16 mod 4.  Reads are
// clamped into [0, kn) and B is zero beyond kn, so the k padding adds
// nothing.  Two register sets of 4 * kQuadSet k each alternate, so that the
// loads of one are in flight while the other feeds the matrix cores.
constexpr int kQuadSet = 4;

template <int NP>
__device__ __forceinline__ void quad_mfma_load(const double* a, const double* b, int kn,
                                               int k0, int kk, double (&av)[kQuadSet],
                                               double (&bv)[kQuadSet]) {
#pragma unroll
  for (int u = 0; u < kQuadSet; ++u) {
    const int k = k0 + 4 * u + kk;
    const bool ok = k < kn;
    av[u] = a[(ok ? k : 0) * NP];
    const double y = b[ok ? k : 0];
    bv[u] = ok ? y : 0.0;
  }
}

__device__ __forceinline__ void quad_mfma_step(const double (&av)[kQuadSet],
                                               const double (&bv)[kQuadSet],
                                               doublex4& acc0, doublex4& acc1) {
#pragma unroll
  for (int u = 0; u < kQuadSet; u += 2) {
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u + 1], bv[u + 1], acc1, 0, 0, 0);
  }
}

template <int NP>
__device__ __forceinline__ void quad_mfma_panel(const double* a, const double* b, int kn,
                                                int kk, doublex4& acc0, doublex4& acc1) {
  constexpr int KS = 4 * kQuadSet;
  double ax[kQuadSet], bx[kQuadSet], ay[kQuadSet], by[kQuadSet];
  quad_mfma_load<NP>(a, b, kn, 0, kk, ax, bx);
  for (int k0 = 0; k0 < kn; k0 += 2 * KS) {
    quad_mfma_load<NP>(a, b, kn, k0 + KS, kk, ay, by);
    quad_mfma_step(ax, bx, acc0, acc1);
    quad_mfma_load<NP>(a, b, kn, k0 + 2 * KS, kk, ax, bx);
    quad_mfma_step(ay, by, acc0, acc1);
  }
}

template <int DIM, int P, int NSP, int MIX>
__global__ __launch_bounds__(kBlock) void euler_quad_kernel(KArgs a, QuadArgs qa) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP;
  extern __shared__ double lds[];
  const int nq = qa.nq, nfq = qa.nfq, epb = qa.epb, nfq_all = NF * qa.nfq;
  const int gs = qa.gs, fs = qa.fs;
  double* s_q = lds;
  double* s_g = s_q + NFLD * epb * NP;
  double* s_f = s_g + NFLD * epb * gs;
  double* s_p = s_f + NFLD * epb * fs;        // [F][epb][NF*NFP] exterior
                                              // face-node values
  double* s_t = s_p + NFLD * epb * NF * NFP;  // [epb][NP] temperature (MIX)
  double* s_tp = s_t + epb * NP;              // [epb][NF*NFP] exterior (MIX)
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  // 1: the elements' state (each field's rows are contiguous in HBM) and
  // the neighbours' values at this element's face nodes (interior faces)
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int f = t / (ne * NP);
    const int r = t - f * ne * NP;
    s_q[f * epb * NP + r] = a.q[f * N + e0 * NP + r];
  }
  for (int t = threadIdx.x; t < ne * NF * NFP; t += blockDim.x) {
    const int le = t / (NF * NFP);
    const int fn = t - le * NF * NFP;
    const int32_t nb = a.vmap_p[(e0 + le) * NF * NFP + fn];
    if (nb < 0) continue;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_p[(f * epb + le) * NF * NFP + fn] = a.q[f * N + nb];
    if constexpr (MIX != 0) s_tp[le * NF * NFP + fn] = qa.temp[nb];
  }
  if constexpr (MIX != 0) {
    for (int t = threadIdx.x; t < ne * NP; t += blockDim.x)
      s_t[t] = qa.temp[e0 * NP + t];
  }
  __syncthreads();
  // 2: inviscid flux at the volume points, in reference directions
  for (int t = threadIdx.x; t < ne * nq; t += blockDim.x) {
    const int le = t / nq;
    const int m = t - le * nq;
    const double* w = qa.interp + m;
    Node<DIM, NSP> s;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s.q[f] = 0.0;
    double tseed = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double wi = w[i * nq];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s.q[f] += wi * s_q[(f * epb + le) * NP + i];
      if constexpr (MIX != 0) tseed += wi * s_t[le * NP + i];
    }
    double tq = 0.0;
    if constexpr (MIX != 0) tq = point_temperature<DIM, NSP>(s.q, tseed, qa.niter);
    node_state<DIM, NSP, MIX, false>(s, a.gas, tq);
    double F[NFLD][DIM];
    inviscid_flux<DIM, NSP>(s, F);
    const double* rx = a.rx + (e0 + le) * DIM * DIM;
#pragma unroll
    for (int f = 0; f < NFLD; ++f)
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double g = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) g += rx[r * DIM + d] * F[f][d];
        s_g[(f * epb + le) * gs + r * nq + m] = g;
      }
  }
  // 3: numerical flux at the face points (both traces interpolated from the
  // face nodes in this element's face-node order)
  for (int t = threadIdx.x; t < ne * nfq_all; t += blockDim.x) {
    const int le = t / nfq_all;
    const int rem = t - le * nfq_all;
    const int fc = rem / nfq;
    const int m = rem - fc * nfq;
    const int64_t e = e0 + le;
    const double* w = qa.finterp + fc * NFP * nfq + m;
    const int32_t* vp = a.vmap_p + (e * NF + fc) * NFP;
    const int* fm = a.fmask + fc * NFP;
    Node<DIM, NSP> sm, sp;
    double seed_m = 0.0, seed_p = 0.0, tm = 0.0, tp = 0.0;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) sm.q[f] = 0.0;
#pragma unroll
    for (int j = 0; j < NFP; ++j) {
      const double wj = w[j * nfq];
      const int ij = fm[j];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sm.q[f] += wj * s_q[(f * epb + le) * NP + ij];
      if constexpr (MIX != 0) seed_m += wj * s_t[le * NP + ij];
    }
    const int32_t v0 = vp[0];
    if (v0 >= 0) {
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sp.q[f] = 0.0;
#pragma unroll
      for (int j = 0; j < NFP; ++j) {
        const double wj = w[j * nfq];
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
          sp.q[f] += wj * s_p[(f * epb + le) * NF * NFP + fc * NFP + j];
        if constexpr (MIX != 0) seed_p += wj * s_tp[le * NF * NFP + fc * NFP + j];
      }
      if constexpr (MIX != 0) tp = point_temperature<DIM, NSP>(sp.q, seed_p, qa.niter);
    } else {
      const int64_t bq = (-int64_t(v0) - 1) / NFP * nfq + m;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sp.q[f] = qa.qbq[f * qa.nbq + bq];
      if constexpr (MIX != 0) tp = qa.tbq[bq];
    }
    if constexpr (MIX != 0) tm = point_temperature<DIM, NSP>(sm.q, seed_m, qa.niter);
    node_state<DIM, NSP, MIX, false>(sm, a.gas, tm);
    node_state<DIM, NSP, MIX, false>(sp, a.gas, tp);
    double n[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + fc) * DIM + d];
    double fstar[NFLD];
    inviscid_numflux<DIM, NSP>(sm, sp, n, a.flux_kind, fstar);
    const double sc = a.fscale[e * NF + fc];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_f[(f * epb + le) * fs + rem] = -sc * fstar[f];
  }
  __syncthreads();
  // 4: dQ/dt = [Wq_0 .. Wq_DIM-1 Lq] [G; -fscale F*]: rows are the NP nodes,
  // columns j = f * epb + le the (field, element) pairs, the contraction
  // runs over the DIM*nq volume points, then the NF*nfq face points.  Each
  // wave forms 16x16 output tiles with v_mfma_f64_16x16x4f64 (A[i][k] in
  // lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16 (i % 4),
  // register i / 4).  Every read stays in bounds: points past the rule enter
  // with B = 0; rows past NP (read from row 0) and columns of absent
  // elements are formed but not stored.
  {
    constexpr int RT = (NP + 15) / 16;
    const int ncol = NFLD * epb;
    const int ntile = RT * ((ncol + 15) / 16);
    const int li = threadIdx.x & 15, kk = (threadIdx.x >> 4) & 3;
    for (int tile = threadIdx.x >> 6; tile < ntile; tile += blockDim.x >> 6) {
      const int rt = tile % RT, ct = tile / RT;
      const int row = rt * 16 + li;
      const int j = ct * 16 + li;
      const int lej = j % epb;
      const bool rok = row < NP, cok = j < ncol && lej < ne;
      doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
      const int jc = j < ncol ? j : 0;
      const int r0 = rok ? row : 0;
      // volume points: k = r * nq + m; face points: k = face * nfq + m
      quad_mfma_panel<NP>(qa.wdiff + r0, s_g + jc * gs, DIM * nq, kk, acc0, acc1);
      quad_mfma_panel<NP>(qa.lift + r0, s_f + jc * fs, nfq_all, kk, acc0, acc1);
      if (cok) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int i = rt * 16 + 4 * g + kk;
          if (i < NP) s_q[j * NP + i] = acc0[g] + acc1[g];
        }
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int f = t / (ne * NP);
    const int r = t - f * ne * NP;
    a.out[f * N + e0 * NP + r] = s_q[f * epb * NP + r];
  }
}

// }}}

// {{{ overintegrated Navier-Stokes (ns_operator with a quadrature tag)
//
// ns_grad_quad_kernel: grad_cv_operator / grad_t_operator on the points.
// The conserved state and T at the volume points and at both sides' face
// points (T from the EOS there; mixtures: Newton from the interpolated
// nodal T), the central face value (boundaries: the exterior state), then,
// per (field, element) column, the DIM reference-direction volume
// contractions and the NF face contractions on the matrix cores, combined
// per lane with -rx and n into the nodal gradient [NS][DIM][N] (NS = F + 1,
// T last).  LDS: s_q [F][epb][NP], s_p [F][epb][NF*NFP] the neighbours'
// face-node values, s_u [NS*epb][gs] the point values (gs >= nq), s_s
// [NS*epb][fs] fscale times the face values, s_o [NS*DIM][epb][NP] the
// result, s_t / s_tp the nodal temperatures (mixtures).
//
// ns_div_quad_kernel: the layout of euler_quad_kernel with F_I - F_V at the
// volume points (the nodal gradients interpolated there) and F_I* - F_V* at
// the face points: the central viscous flux of both sides' states and
// interpolated gradient traces (boundary points: the exterior state with
// the interior gradients, which for a dummy boundary -- exterior state =
// interior state -- is the one-sided flux).  Extra LDS: s_gr [NS*DIM][epb]
// [NP] the element's gradients, s_gp [NS*DIM][epb][NF*NFP] the neighbours'
// at this element's face nodes.

template <int DIM>
__device__ __forceinline__ double ideal_temperature(const double* q, const GasDev& g) {
  const double irho = 1.0 / q[0];
  double mm = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) mm += q[2 + d] * q[2 + d];
  return ((g.gamma - 1.0) / g.gas_const) * (q[1] - 0.5 * mm * irho) * irho;
}

template <int DIM, int NSP, int MIX>
__device__ __forceinline__ double quad_point_temperature(const double* q, double seed,
                                                         const GasDev& g, int niter) {
  if constexpr (MIX != 0)
    return point_temperature<DIM, NSP>(q, seed, niter);
  else
    return ideal_temperature<DIM>(q, g);
}

// Phase 1 of both kernels: the elements' state and the neighbours' values
// at this element's face nodes (interior faces), and, for mixtures, the
// nodal temperatures of both.
template <int DIM, int P, int NSP, int MIX>
__device__ __forceinline__ void quad_load_states(const KArgs& a, const QuadArgs& qa,
                                                 int64_t e0, int ne, int epb, double* s_q,
                                                 double* s_p, double* s_t, double* s_tp) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP;
  const int64_t N = a.nnodes;
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int f = t / (ne * NP);
    const int r = t - f * ne * NP;
    s_q[f * epb * NP + r] = a.q[f * N + e0 * NP + r];
  }
  for (int t = threadIdx.x; t < ne * NF * NFP; t += blockDim.x) {
    const int le = t / (NF * NFP);
    const int fn = t - le * NF * NFP;
    const int32_t nb = a.vmap_p[(e0 + le) * NF * NFP + fn];
    if (nb < 0) continue;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_p[(f * epb + le) * NF * NFP + fn] = a.q[f * N + nb];
    if constexpr (MIX != 0) s_tp[le * NF * NFP + fn] = qa.temp[nb];
  }
  if constexpr (MIX != 0) {
    for (int t = threadIdx.x; t < ne * NP; t += blockDim.x) s_t[t] = qa.temp[e0 * NP + t];
  }
}

template <int DIM, int P, int NSP, int MIX>
__global__ __launch_bounds__(kBlock) void ns_grad_quad_kernel(KArgs a, QuadArgs qa) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP, NS = NFLD + 1;
  extern __shared__ double lds[];
  const int nq = qa.nq, nfq = qa.nfq, epb = qa.epb, nfq_all = NF * qa.nfq;
  const int gs = qa.gs, fs = qa.fs;
  double* s_q = lds;
  double* s_p = s_q + NFLD * epb * NP;
  double* s_u = s_p + NFLD * epb * NF * NFP;
  double* s_s = s_u + NS * epb * gs;
  double* s_o = s_s + NS * epb * fs;
  double* s_t = s_o + NS * DIM * epb * NP;
  double* s_tp = s_t + epb * NP;
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  quad_load_states<DIM, P, NSP, MIX>(a, qa, e0, ne, epb, s_q, s_p, s_t, s_tp);
  __syncthreads();
  // 2: state and temperature at the volume points
  for (int t = threadIdx.x; t < ne * nq; t += blockDim.x) {
    const int le = t / nq;
    const int m = t - le * nq;
    const double* w = qa.interp + m;
    double u[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) u[f] = 0.0;
    double tseed = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double wi = w[i * nq];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) u[f] += wi * s_q[(f * epb + le) * NP + i];
      if constexpr (MIX != 0) tseed += wi * s_t[le * NP + i];
    }
    const double tq = quad_point_temperature<DIM, NSP, MIX>(u, tseed, a.gas, qa.niter);
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_u[(f * epb + le) * gs + m] = u[f];
    s_u[(NFLD * epb + le) * gs + m] = tq;
  }
  // 3: the gradient flux value at the face points (central inside, the
  // exterior value on boundaries), times the face Jacobian ratio
  for (int t = threadIdx.x; t < ne * nfq_all; t += blockDim.x) {
    const int le = t / nfq_all;
    const int rem = t - le * nfq_all;
    const int fc = rem / nfq;
    const int m = rem - fc * nfq;
    const int64_t e = e0 + le;
    const double* w = qa.finterp + fc * NFP * nfq + m;
    const int32_t v0 = a.vmap_p[(e * NF + fc) * NFP];
    const int* fm = a.fmask + fc * NFP;
    double um[NFLD], up[NFLD];
    double seed_m = 0.0, seed_p = 0.0, tp;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) um[f] = 0.0;
#pragma unroll
    for (int j = 0; j < NFP; ++j) {
      const double wj = w[j * nfq];
      const int ij = fm[j];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) um[f] += wj * s_q[(f * epb + le) * NP + ij];
      if constexpr (MIX != 0) seed_m += wj * s_t[le * NP + ij];
    }
    const double tm = quad_point_temperature<DIM, NSP, MIX>(um, seed_m, a.gas, qa.niter);
    const double sc = a.fscale[e * NF + fc];
    if (v0 >= 0) {
#pragma unroll
      for (int f = 0; f < NFLD; ++f) up[f] = 0.0;
#pragma unroll
      for (int j = 0; j < NFP; ++j) {
        const double wj = w[j * nfq];
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
          up[f] += wj * s_p[(f * epb + le) * NF * NFP + fc * NFP + j];
        if constexpr (MIX != 0) seed_p += wj * s_tp[le * NF * NFP + fc * NFP + j];
      }
      tp = quad_point_temperature<DIM, NSP, MIX>(up, seed_p, a.gas, qa.niter);
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
        s_s[(f * epb + le) * fs + rem] = sc * ((um[f] + up[f]) / 2);
      s_s[(NFLD * epb + le) * fs + rem] = sc * ((tm + tp) / 2);
    } else {
      const int64_t bq = (-int64_t(v0) - 1) / NFP * nfq + m;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) up[f] = qa.qbq[f * qa.nbq + bq];
      if constexpr (MIX != 0)
        tp = qa.tbq[bq];
      else
        tp = ideal_temperature<DIM>(up, a.gas);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s_s[(f * epb + le) * fs + rem] = sc * up[f];
      s_s[(NFLD * epb + le) * fs + rem] = sc * tp;
    }
  }
  __syncthreads();
  // 4: per column j = f * epb + le: D_r = Wq_r u (r < DIM) and L_c = Lq_c
  // (fscale u*) (c < NF) on the matrix cores (operand layout as in
  // euler_quad_kernel: each lane's accumulators hold its own column j), then
  // grad_d = sum_r -rx[r][d] D_r + sum_c n_c[d] L_c with the element's
  // factors read per lane.
  {
    constexpr int RT = (NP + 15) / 16;
    const int ncol = NS * epb;
    const int ntile = RT * ((ncol + 15) / 16);
    const int li = threadIdx.x & 15, kk = (threadIdx.x >> 4) & 3;
    for (int tile = threadIdx.x >> 6; tile < ntile; tile += blockDim.x >> 6) {
      const int rt = tile % RT, ct = tile / RT;
      const int row = rt * 16 + li;
      const int j = ct * 16 + li;
      const int lej = j % epb;
      const bool cok = j < ncol && lej < ne;
      const int jc = j < ncol ? j : 0;
      const int r0 = row < NP ? row : 0;
      const int64_t ej = e0 + (cok ? lej : 0);
      doublex4 g[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) g[d] = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
        quad_mfma_panel<NP>(qa.wdiff + r * nq * NP + r0, s_u + jc * gs, nq, kk, acc0, acc1);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double c = -a.rx[ej * DIM * DIM + r * DIM + d];
          g[d] += c * (acc0 + acc1);
        }
      }
#pragma unroll
      for (int fc = 0; fc < NF; ++fc) {
        doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
        quad_mfma_panel<NP>(qa.lift + fc * nfq * NP + r0, s_s + jc * fs + fc * nfq, nfq, kk,
                            acc0, acc1);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double c = a.normals[(ej * NF + fc) * DIM + d];
          g[d] += c * (acc0 + acc1);
        }
      }
      if (cok) {
        const int f = j / epb;
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const int i = rt * 16 + 4 * q4 + kk;
          if (i < NP) {
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              s_o[((f * DIM + d) * epb + lej) * NP + i] = g[d][q4];
          }
        }
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NS * DIM * ne * NP; t += blockDim.x) {
    const int k = t / (ne * NP);
    const int r = t - k * ne * NP;
    a.out[k * N + e0 * NP + r] = s_o[k * epb * NP + r];
  }
}

template <int DIM, int P, int NSP, int MIX>
__global__ __launch_bounds__(kBlock) void ns_div_quad_kernel(KArgs a, QuadArgs qa) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP, NS = NFLD + 1, NSD = NS * DIM;
  extern __shared__ double lds[];
  const int nq = qa.nq, nfq = qa.nfq, epb = qa.epb, nfq_all = NF * qa.nfq;
  const int gs = qa.gs, fs = qa.fs;
  double* s_q = lds;
  double* s_g = s_q + NFLD * epb * NP;
  double* s_f = s_g + NFLD * epb * gs;
  double* s_p = s_f + NFLD * epb * fs;
  double* s_gr = s_p + NFLD * epb * NF * NFP;
  double* s_gp = s_gr + NSD * epb * NP;
  double* s_t = s_gp + NSD * epb * NF * NFP;
  double* s_tp = s_t + epb * NP;
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  quad_load_states<DIM, P, NSP, MIX>(a, qa, e0, ne, epb, s_q, s_p, s_t, s_tp);
  for (int t = threadIdx.x; t < NSD * ne * NP; t += blockDim.x) {
    const int k = t / (ne * NP);
    const int r = t - k * ne * NP;
    s_gr[k * epb * NP + r] = a.grad[k * N + e0 * NP + r];
  }
  for (int t = threadIdx.x; t < ne * NF * NFP; t += blockDim.x) {
    const int le = t / (NF * NFP);
    const int fn = t - le * NF * NFP;
    const int32_t nb = a.vmap_p[(e0 + le) * NF * NFP + fn];
    if (nb < 0) continue;
#pragma unroll
    for (int k = 0; k < NSD; ++k) s_gp[(k * epb + le) * NF * NFP + fn] = a.grad[k * N + nb];
  }
  __syncthreads();
  // 2: F_I - F_V at the volume points, in reference directions (the node
  // loops stay rolled: unrolled, the interpolation's LDS reads are hoisted
  // and spill)
  for (int t = threadIdx.x; t < ne * nq; t += blockDim.x) {
    const int le = t / nq;
    const int m = t - le * nq;
    const double* w = qa.interp + m;
    Node<DIM, NSP> s;
    double gacc[NSD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s.q[f] = 0.0;
#pragma unroll
    for (int k = 0; k < NSD; ++k) gacc[k] = 0.0;
    double tseed = 0.0;
#pragma unroll 1
    for (int i = 0; i < NP; ++i) {
      const double wi = w[i * nq];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s.q[f] += wi * s_q[(f * epb + le) * NP + i];
#pragma unroll
      for (int k = 0; k < NSD; ++k) gacc[k] += wi * s_gr[(k * epb + le) * NP + i];
      if constexpr (MIX != 0) tseed += wi * s_t[le * NP + i];
    }
    double tq = 0.0;
    if constexpr (MIX != 0) tq = point_temperature<DIM, NSP>(s.q, tseed, qa.niter);
    node_state<DIM, NSP, MIX>(s, a.gas, tq);
    double F[NFLD][DIM];
    viscous_flux<DIM, NSP, MIX>(s, reinterpret_cast<const double(*)[DIM]>(gacc),
                                gacc + NFLD * DIM, F);
    if (a.inviscid_on) {
      double FI[NFLD][DIM];
      inviscid_flux<DIM, NSP>(s, FI);
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
#pragma unroll
        for (int d = 0; d < DIM; ++d) F[f][d] = FI[f][d] - F[f][d];
    } else {
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
#pragma unroll
        for (int d = 0; d < DIM; ++d) F[f][d] = -F[f][d];
    }
    const double* rx = a.rx + (e0 + le) * DIM * DIM;
#pragma unroll
    for (int f = 0; f < NFLD; ++f)
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double g = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) g += rx[r * DIM + d] * F[f][d];
        s_g[(f * epb + le) * gs + r * nq + m] = g;
      }
  }
  // 3: F_I* - F_V* at the face points: the own side's state and viscous
  // flux first, then the other side's, so that one gradient set is live
  for (int t = threadIdx.x; t < ne * nfq_all; t += blockDim.x) {
    const int le = t / nfq_all;
    const int rem = t - le * nfq_all;
    const int fc = rem / nfq;
    const int m = rem - fc * nfq;
    const int64_t e = e0 + le;
    const double* w = qa.finterp + fc * NFP * nfq + m;
    const int32_t v0 = a.vmap_p[(e * NF + fc) * NFP];
    const int* fm = a.fmask + fc * NFP;
    double n[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + fc) * DIM + d];
    Node<DIM, NSP> sm, sp;
    double fstar[NFLD], fv[NFLD];
    {
      double seed_m = 0.0, tm = 0.0;
      double gacc[NSD];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sm.q[f] = 0.0;
#pragma unroll
      for (int k = 0; k < NSD; ++k) gacc[k] = 0.0;
#pragma unroll 1
      for (int j = 0; j < NFP; ++j) {
        const double wj = w[j * nfq];
        const int ij = fm[j];
#pragma unroll
        for (int f = 0; f < NFLD; ++f) sm.q[f] += wj * s_q[(f * epb + le) * NP + ij];
#pragma unroll
        for (int k = 0; k < NSD; ++k) gacc[k] += wj * s_gr[(k * epb + le) * NP + ij];
        if constexpr (MIX != 0) seed_m += wj * s_t[le * NP + ij];
      }
      if constexpr (MIX != 0) tm = point_temperature<DIM, NSP>(sm.q, seed_m, qa.niter);
      node_state<DIM, NSP, MIX>(sm, a.gas, tm);
      viscous_flux_n<DIM, NSP, MIX>(sm, reinterpret_cast<const double(*)[DIM]>(gacc),
                                    gacc + NFLD * DIM, n, fv);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) fstar[f] = -0.5 * fv[f];
    }
    double tp = 0.0;
    if (v0 >= 0) {
      double seed_p = 0.0;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sp.q[f] = 0.0;
#pragma unroll 1
      for (int j = 0; j < NFP; ++j) {
        const double wj = w[j * nfq];
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
          sp.q[f] += wj * s_p[(f * epb + le) * NF * NFP + fc * NFP + j];
        if constexpr (MIX != 0) seed_p += wj * s_tp[le * NF * NFP + fc * NFP + j];
      }
      if constexpr (MIX != 0) tp = point_temperature<DIM, NSP>(sp.q, seed_p, qa.niter);
    } else {
      const int64_t bq = (-int64_t(v0) - 1) / NFP * nfq + m;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sp.q[f] = qa.qbq[f * qa.nbq + bq];
      if constexpr (MIX != 0) tp = qa.tbq[bq];
    }
    node_state<DIM, NSP, MIX>(sp, a.gas, tp);
    if (a.inviscid_on) {
      inviscid_numflux<DIM, NSP>(sm, sp, n, a.flux_kind, fv);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) fstar[f] += fv[f];
    }
    {
      // the other side's gradient trace: the neighbour's (interior faces)
      // or this element's own (boundary faces)
      double gacc[NSD];
#pragma unroll
      for (int k = 0; k < NSD; ++k) gacc[k] = 0.0;
      if (v0 >= 0) {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
          const double wj = w[j * nfq];
#pragma unroll
          for (int k = 0; k < NSD; ++k)
            gacc[k] += wj * s_gp[(k * epb + le) * NF * NFP + fc * NFP + j];
        }
      } else {
#pragma unroll 1
        for (int j = 0; j < NFP; ++j) {
          const double wj = w[j * nfq];
          const int ij = fm[j];
#pragma unroll
          for (int k = 0; k < NSD; ++k) gacc[k] += wj * s_gr[(k * epb + le) * NP + ij];
        }
      }
      viscous_flux_n<DIM, NSP, MIX>(sp, reinterpret_cast<const double(*)[DIM]>(gacc),
                                    gacc + NFLD * DIM, n, fv);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) fstar[f] -= 0.5 * fv[f];
    }
    const double sc = a.fscale[e * NF + fc];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_f[(f * epb + le) * fs + rem] = -sc * fstar[f];
  }
  __syncthreads();
  // 4: the contraction of euler_quad_kernel
  {
    constexpr int RT = (NP + 15) / 16;
    const int ncol = NFLD * epb;
    const int ntile = RT * ((ncol + 15) / 16);
    const int li = threadIdx.x & 15, kk = (threadIdx.x >> 4) & 3;
    for (int tile = threadIdx.x >> 6; tile < ntile; tile += blockDim.x >> 6) {
      const int rt = tile % RT, ct = tile / RT;
      const int row = rt * 16 + li;
      const int j = ct * 16 + li;
      const int lej = j % epb;
      const bool rok = row < NP, cok = j < ncol && lej < ne;
      doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
      const int jc = j < ncol ? j : 0;
      const int r0 = rok ? row : 0;
      quad_mfma_panel<NP>(qa.wdiff + r0, s_g + jc * gs, DIM * nq, kk, acc0, acc1);
      quad_mfma_panel<NP>(qa.lift + r0, s_f + jc * fs, nfq_all, kk, acc0, acc1);
      if (cok) {
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const int i = rt * 16 + 4 * q4 + kk;
          if (i < NP) s_q[j * NP + i] = acc0[q4] + acc1[q4];
        }
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int f = t / (ne * NP);
    const int r = t - f * ne * NP;
    a.out[f * N + e0 * NP + r] = s_q[f * epb * NP + r];
  }
}

// }}}

// {{{ entropy-stable Euler RHS (flux differencing, ideal gas)
//
// esdg_project_kernel: projected entropy variables v~ = P_q v(V_q u) per
// element (grudge's volume_quadrature_project), written to ev.
// esdg_kernel: the state at the nh hybridized points from v~ (own element)
// and the face fluxes from both sides' v~; each lane then owns one
// hybridized point h of one element and sums over the points j its row of
// (Q_h - Q_h^T) o F_S (volume rows: all nh points, face rows: the nq volume
// points -- the face-face block vanishes); the sums are mapped back with
// M^-1 V_h^T.  Geometric factors enter as rx and n * fscale (the Jacobian
// cancels against the mass matrix on affine elements).

struct EsdgArgs {
  int nq, nfq, nh, epb;
  const double* interp;   // [NP][nq]
  const double* proj;     // [nq][NP]  P_q transposed
  const double* vh;       // [NP][nh]  V_h transposed
  const double* skew;     // [DIM][nq(j)][nq(h)]: (Q_r - Q_r^T)[h][j]
  const double* ext;      // [NF*nfq][nq]  E
  const double* ext_col;  // [nq][NF*nfq]  E transposed
  const double* apply;    // [nh][NP]  (M^-1 V_h^T) transposed
  const double* finterp;  // [NF][NFP][nfq]
  const double* wf;       // [nfq]
  const double* qbq;      // [F][nbq]
  int64_t nbq;
  double* ev;             // [F][N] projected entropy variables
  int dissipation;
  // nq = 64, NF*nfq = 192, 256 threads: the paired phase 4 (each two-point
  // flux once, for both rows) with the operators along cyclic diagonals
  int paired;
  const double* skew_diag;  // [DIM][nq(s)][nq(h)]: (Q_r - Q_r^T)[h][(h+s) % nq]
  const double* ext_diag;   // [nq(s)][NF*nfq(m)]: E[m][(m+s) % nq]
  // mixtures: Newton steps per point, exterior temperatures at the
  // boundary points [nbq]; es_volume: the volume points' state is the
  // entropy-stable projected one (euler_operator) rather than the plain
  // projected one (the states ns_operator hands to its inviscid operator)
  int niter;
  const double* tbq;
  int es_volume;
};

template <int DIM, int P, int NSP>
__global__ __launch_bounds__(kBlock) void esdg_project_kernel(KArgs a, EsdgArgs ea) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP;
  constexpr int NFLD = 2 + DIM + NSP;
  extern __shared__ double lds[];
  const int nq = ea.nq, epb = ea.epb;
  double* s_e = lds;  // [F][epb][nq]
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  const double gamma = a.gas.gamma;
  for (int t = threadIdx.x; t < ne * nq; t += blockDim.x) {
    const int le = t / nq;
    const int m = t - le * nq;
    const int64_t base = (e0 + le) * NP;
    double u[NFLD], v[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) u[f] = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double w = ea.interp[i * nq + m];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) u[f] += w * a.q[f * N + base + i];
    }
    cons_to_entropy<DIM, NSP>(u, gamma, v);
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_e[(f * epb + le) * nq + m] = v[f];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ne * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    double v[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) v[f] = 0.0;
    for (int m = 0; m < nq; ++m) {
      const double w = ea.proj[m * NP + i];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) v[f] += w * s_e[(f * epb + le) * nq + m];
    }
#pragma unroll
    for (int f = 0; f < NFLD; ++f) ea.ev[f * N + e0 * NP + t] = v[f];
  }
}

// Phase 4 of esdg_kernel when nq = 64 and NF*nfq = 192 (nh = 256, four
// waves): each two-point flux is evaluated once.  (Q_h - Q_h^T) is skew and
// F_S symmetric, so the pair (h, j) adds S[h][j] F_S to row h and the
// negative to row j.  Wave 0 owns the 64 volume rows and meets the pairs
// (h, (h+s) mod 64), s = 1..31, in step s; the contributions to the partner
// rows travel in a register vector that moves one lane down per step, so
// that lane h always holds row (h+s) mod 64 (s = 32 pairs each lane with the
// lane that owns the same pair: both add their own row).  Wave 1 + b owns
// face rows 64b..64b+63 and meets the volume points (l+s) mod 64 in step s,
// s = 0..63 (the face-face block is zero); its moving vector collects that
// block's part of the volume rows and is back in place after 64 steps.
// Results: volume rows in the four nq-long pieces of s_acc (piece 0 from
// wave 0, piece 1 + b from wave 1 + b; phase 5 adds them), face rows plus
// the face flux in the face slots.  One element per workgroup: thread
// 64 + m has computed face point m's flux (fkeep) in phase 3.
template <int DIM, int NSP>
__device__ __forceinline__ void esdg_pairs_64(const KArgs& a, const EsdgArgs& ea,
                                              const EcPair<DIM, NSP>* s_prim,
                                              const double (&fkeep)[2 + DIM + NSP],
                                              double* s_acc, int64_t e, int nacc,
                                              double gamma) {
  constexpr int NF = DIM + 1;
  constexpr int NFLD = 2 + DIM + NSP;
  using Prim = EcPair<DIM, NSP>;
  constexpr int NQ = 64, NFT = 192;
  const int nfq = ea.nfq;
  const int l = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int down = (l + 1) & 63;
  {
    const int le = 0, epb = 1;
    const Prim* pj = s_prim;
    double acc[NFLD], mv[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) acc[f] = mv[f] = 0.0;
    if (wave == 0) {
      const Prim ph = pj[l];
      double rx[DIM * DIM];
#pragma unroll
      for (int k = 0; k < DIM * DIM; ++k) rx[k] = a.rx[e * DIM * DIM + k];
      for (int s = 1; s <= 32; ++s) {
        double sd[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sd[d] = 0.0;
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
          const double sr = ea.skew_diag[(r * NQ + s) * NQ + l];
#pragma unroll
          for (int d = 0; d < DIM; ++d) sd[d] += rx[r * DIM + d] * sr;
        }
        double fv[NFLD];
        ec_flux_dot<DIM, NSP, true>(ph, pj[(l + s) & 63], sd, gamma, fv);
        if (s < 32) {
#pragma unroll
          for (int f = 0; f < NFLD; ++f) {
            acc[f] += fv[f];
            mv[f] = __shfl(mv[f] - fv[f], down, 64);
          }
        } else {
#pragma unroll
          for (int f = 0; f < NFLD; ++f) acc[f] += fv[f];
        }
      }
      // lane l holds row (l + 32) mod 64
#pragma unroll
      for (int f = 0; f < NFLD; ++f) {
        acc[f] += __shfl(mv[f], (l + 32) & 63, 64);
        s_acc[(f * epb + le) * nacc + l] = acc[f];
      }
    } else {
      const int m = (wave - 1) * NQ + l;
      const Prim pm = pj[NQ + m];
      const int fc = m / nfq;
      const double bsc = a.fscale[e * NF + fc] * ea.wf[m - fc * nfq];
      double nrm[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) nrm[d] = bsc * a.normals[(e * NF + fc) * DIM + d];
      for (int s = 0; s < NQ; ++s) {
        const double b = ea.ext_diag[s * NFT + m];
        double sd[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sd[d] = b * nrm[d];
        double fv[NFLD];
        ec_flux_dot<DIM, NSP, true>(pm, pj[(l + s) & 63], sd, gamma, fv);
#pragma unroll
        for (int f = 0; f < NFLD; ++f) {
          acc[f] -= fv[f];
          mv[f] = __shfl(mv[f] + fv[f], down, 64);
        }
      }
#pragma unroll
      for (int f = 0; f < NFLD; ++f) {
        double* row = s_acc + (f * epb + le) * nacc;
        row[wave * NQ + l] = mv[f];
        row[4 * NQ + m] = acc[f] + fkeep[f];
      }
    }
  }
}

// Entropy-stable face flux times fscale * w at face point m of element
// e0 + le (phase 3 of esdg_kernel): the own trace state from s_qf, the
// neighbour's from its projected entropy variables at the shared face nodes
// (s_vp), on boundary faces the re-projected interior state and the
// boundary condition's exterior state.
template <int DIM, int P, int NSP>
__device__ __forceinline__ void esdg_face_flux(const KArgs& a, const EsdgArgs& ea,
                                               const double* s_qf, const double* s_vp,
                                               const double* s_q, int64_t e0, int le, int m,
                                               double gamma, double* out) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP;
  using Prim = EcPrim<DIM, NSP>;
  const int nfq = ea.nfq, epb = ea.epb, nft = NF * ea.nfq;
  const int fc = m / nfq;
  const int mm = m - fc * nfq;
  const int64_t e = e0 + le;
  const double* w = ea.finterp + fc * NFP * nfq + mm;
  const int32_t v0 = a.vmap_p[(e * NF + fc) * NFP];
  double qm[NFLD], qp[NFLD];
  if (v0 >= 0) {
#pragma unroll
    for (int f = 0; f < NFLD; ++f) qm[f] = s_qf[(f * epb + le) * nft + m];
    double vp[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) vp[f] = 0.0;
#pragma unroll
    for (int j = 0; j < NFP; ++j) {
      const double wj = w[j * nfq];
#pragma unroll
      for (int f = 0; f < NFLD; ++f)
        vp[f] += wj * s_vp[(f * epb + le) * NF * NFP + fc * NFP + j];
    }
    entropy_to_cons<DIM, NSP>(vp, gamma, qp);
  } else {
    // boundary: the interpolated interior state (through the entropy
    // variables and back, as the reference's projection does) and the
    // boundary condition's exterior state
    double ui[NFLD], vi[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) ui[f] = 0.0;
#pragma unroll
    for (int j = 0; j < NFP; ++j) {
      const double wj = w[j * nfq];
      const int ij = a.fmask[fc * NFP + j];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) ui[f] += wj * s_q[(f * epb + le) * NP + ij];
    }
    cons_to_entropy<DIM, NSP>(ui, gamma, vi);
    entropy_to_cons<DIM, NSP>(vi, gamma, qm);
    const int64_t bq = (-int64_t(v0) - 1) / NFP * nfq + mm;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) qp[f] = ea.qbq[f * ea.nbq + bq];
  }
  Prim pm, pp;
  ec_prim<DIM, NSP>(qm, gamma, pm);
  ec_prim<DIM, NSP>(qp, gamma, pp);
  double n[DIM], fs[NFLD];
#pragma unroll
  for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + fc) * DIM + d];
#pragma unroll
  for (int f = 0; f < NFLD; ++f) fs[f] = 0.0;
  ec_flux_dot<DIM, NSP>(pm, pp, n, gamma, fs);
  if (ea.dissipation) {
    const double alpha = fmax(pm.c, pp.c);
#pragma unroll
    for (int f = 0; f < NFLD; ++f) fs[f] -= 0.5 * alpha * (qp[f] - qm[f]);
  }
  const double sc = a.fscale[e * NF + fc] * ea.wf[mm];
#pragma unroll
  for (int f = 0; f < NFLD; ++f) out[f] = sc * fs[f];
}

// Phase 4 of the entropy-stable kernels (row by row): each work item sums
// one row piece of (Q_h - Q_h^T) o F_S with the two-point flux `flux`
// (called as flux(point_h, point_j, direction, acc)); face rows add the
// face flux s_fs.  Slots in s_acc as described in esdg_kernel.
template <int DIM, int NFLD, typename Pair, typename Flux>
__device__ __forceinline__ void esdg_rows(const KArgs& a, const EsdgArgs& ea,
                                          const Pair* s_prim, const double* s_fs,
                                          double* s_acc, int64_t e0, int ne, int nacc,
                                          int C, Flux flux) {
  constexpr int NF = DIM + 1;
  const int nq = ea.nq, nfq = ea.nfq, nh = ea.nh, epb = ea.epb, nft = NF * ea.nfq;
  const int nitems = nq * C + nft;
  for (int t = threadIdx.x; t < ne * nitems; t += blockDim.x) {
    const int le = t / nitems;
    const int it = t - le * nitems;
    int h, c, j0, j1;
    if (it < nq * C) {
      c = it / nq;
      h = it - c * nq;
      j0 = c * nq;
      j1 = j0 + nq < nh ? j0 + nq : nh;
    } else {
      c = 0;
      h = nq + (it - nq * C);
      j0 = 0;
      j1 = nq;
    }
    const int64_t e = e0 + le;
    const Pair ph = s_prim[le * nh + h];
    const Pair* pj = s_prim + le * nh;
    double acc[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) acc[f] = 0.0;
    if (h < nq) {
      double rx[DIM * DIM];
#pragma unroll
      for (int k = 0; k < DIM * DIM; ++k) rx[k] = a.rx[e * DIM * DIM + k];
      const int jv = j1 < nq ? j1 : nq;
      for (int j = j0; j < jv; ++j) {
        double sr[DIM], sd[DIM];
#pragma unroll
        for (int r = 0; r < DIM; ++r) sr[r] = ea.skew[(r * nq + j) * nq + h];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          double v = 0.0;
#pragma unroll
          for (int r = 0; r < DIM; ++r) v += rx[r * DIM + d] * sr[r];
          sd[d] = v;
        }
        flux(ph, pj[j], sd, acc);
      }
      for (int j = (j0 > nq ? j0 : nq); j < j1; ++j) {
        const int m = j - nq;
        const int fc = m / nfq;
        const double b = ea.ext[m * nq + h] * a.fscale[e * NF + fc] * ea.wf[m - fc * nfq];
        double sd[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sd[d] = b * a.normals[(e * NF + fc) * DIM + d];
        flux(ph, pj[j], sd, acc);
      }
    } else {
      const int m = h - nq;
      const int fc = m / nfq;
      const double bsc = -a.fscale[e * NF + fc] * ea.wf[m - fc * nfq];
      double nrm[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) nrm[d] = bsc * a.normals[(e * NF + fc) * DIM + d];
      for (int j = 0; j < nq; ++j) {
        const double ej = ea.ext_col[j * nft + m];
        double sd[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sd[d] = ej * nrm[d];
        flux(ph, pj[j], sd, acc);
      }
#pragma unroll
      for (int f = 0; f < NFLD; ++f) acc[f] += s_fs[(f * epb + le) * nft + m];
    }
    const int slot = h < nq ? c * nq + h : C * nq + (h - nq);
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_acc[(f * epb + le) * nacc + slot] = acc[f];
  }
}

// Phase 5 of the entropy-stable kernels: dQ/dt = -M^-1 V_h^T acc.
template <int DIM, int P, int NFLD>
__device__ __forceinline__ void esdg_apply(const KArgs& a, const EsdgArgs& ea,
                                           double* s_acc, int64_t e0, int ne,
                                           int nacc, int C) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NF = E::NF;
  const int nq = ea.nq, epb = ea.epb, nft = NF * ea.nfq;
  const int64_t N = a.nnodes;
  // the volume rows' pieces summed in place,
  // then one (field, element, node) output per thread
  if (C > 1) {
    for (int t = threadIdx.x; t < NFLD * ne * nq; t += blockDim.x) {
      const int fl = t / nq;
      const int h = t - fl * nq;
      const int f = fl / ne, le = fl - f * ne;
      double* sa = s_acc + (f * epb + le) * nacc + h;
      double v = sa[0];
      for (int cc = 1; cc < C; ++cc) v += sa[cc * nq];
      sa[0] = v;
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int fl = t / NP;
    const int i = t - fl * NP;
    const int f = fl / ne, le = fl - f * ne;
    const double* sa = s_acc + (f * epb + le) * nacc;
    const double* w = ea.apply + i;
    double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0;
    int h = 0;
    for (; h + 4 <= nq; h += 4) {
      o0 += w[h * NP] * sa[h];
      o1 += w[(h + 1) * NP] * sa[h + 1];
      o2 += w[(h + 2) * NP] * sa[h + 2];
      o3 += w[(h + 3) * NP] * sa[h + 3];
    }
    for (; h < nq; ++h) o0 += w[h * NP] * sa[h];
    const double* sf = sa + C * nq;
    const double* wf = w + nq * NP;
    int m = 0;
    for (; m + 4 <= nft; m += 4) {
      o0 += wf[m * NP] * sf[m];
      o1 += wf[(m + 1) * NP] * sf[m + 1];
      o2 += wf[(m + 2) * NP] * sf[m + 2];
      o3 += wf[(m + 3) * NP] * sf[m + 3];
    }
    for (; m < nft; ++m) o0 += wf[m * NP] * sf[m];
    a.out[f * N + (e0 + le) * NP + i] = -((o0 + o1) + (o2 + o3));
  }
}

template <int DIM, int P, int NSP>
__global__ __launch_bounds__(kBlock) void esdg_kernel(KArgs a, EsdgArgs ea) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP;
  using Pair = EcPair<DIM, NSP>;
  extern __shared__ double lds[];
  const int nq = ea.nq, nfq = ea.nfq, nh = ea.nh, epb = ea.epb, nft = NF * ea.nfq;
  Pair* s_prim = reinterpret_cast<Pair*>(lds);             // [epb][nh]
  double* s_v = reinterpret_cast<double*>(s_prim + epb * nh);  // [F][epb][NP]
  double* s_q = s_v + NFLD * epb * NP;                      // [F][epb][NP]
  double* s_vp = s_q + NFLD * epb * NP;                     // [F][epb][NF*NFP]
  // phase 4 splits volume rows into C pieces of nq entries
  const int C = (nh + nq - 1) / nq;
  const int nacc = C * nq + nft;                            // row sums per field
  // face fluxes [F][epb][nft] (not in the paired layout: kept in registers)
  double* s_fs = s_vp + NFLD * epb * NF * NFP;
  double* s_qf = ea.paired ? s_fs : s_fs + NFLD * epb * nft;  // [F][epb][nft], 2-3
  double* s_acc = s_qf;                                     // [F][epb][nacc], phases 4-5
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  const double gamma = a.gas.gamma;
  // 1: own v~ and q, the neighbours' v~ at this element's face nodes
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int f = t / (ne * NP);
    const int r = t - f * ne * NP;
    s_v[f * epb * NP + r] = ea.ev[f * N + e0 * NP + r];
    s_q[f * epb * NP + r] = a.q[f * N + e0 * NP + r];
  }
  for (int t = threadIdx.x; t < ne * NF * NFP; t += blockDim.x) {
    const int le = t / (NF * NFP);
    const int fn = t - le * NF * NFP;
    const int32_t nb = a.vmap_p[(e0 + le) * NF * NFP + fn];
    if (nb < 0) continue;
#pragma unroll
    for (int f = 0; f < NFLD; ++f)
      s_vp[(f * epb + le) * NF * NFP + fn] = ea.ev[f * N + nb];
  }
  __syncthreads();
  // 2: the state at the hybridized points
  for (int t = threadIdx.x; t < ne * nh; t += blockDim.x) {
    const int le = t / nh;
    const int h = t - le * nh;
    double v[NFLD], u[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) v[f] = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double w = ea.vh[i * nh + h];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) v[f] += w * s_v[(f * epb + le) * NP + i];
    }
    entropy_to_cons<DIM, NSP>(v, gamma, u);
    EcPrim<DIM, NSP> pr;
    ec_prim<DIM, NSP>(u, gamma, pr);
    s_prim[le * nh + h] = pr;
    if (h >= nq) {
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s_qf[(f * epb + le) * nft + (h - nq)] = u[f];
    }
  }
  __syncthreads();
  // 3: entropy-stable face flux at the face points, times fscale * w
  double fkeep[NFLD];
#pragma unroll
  for (int f = 0; f < NFLD; ++f) fkeep[f] = 0.0;
  if (ea.paired) {
    // one element per workgroup: thread 64 + m keeps face point m's flux
    // for its face row in phase 4
    const int m = int(threadIdx.x) - 64;
    if (m >= 0 && m < nft && ne == 1)
      esdg_face_flux<DIM, P, NSP>(a, ea, s_qf, s_vp, s_q, e0, 0, m, gamma, fkeep);
  } else {
    for (int t = threadIdx.x; t < ne * nft; t += blockDim.x) {
      const int le = t / nft;
      const int m = t - le * nft;
      double fs[NFLD];
      esdg_face_flux<DIM, P, NSP>(a, ea, s_qf, s_vp, s_q, e0, le, m, gamma, fs);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s_fs[(f * epb + le) * nft + m] = fs[f];
    }
  }
  __syncthreads();
  // 4: rows of (Q_h - Q_h^T) o F_S, plus the face flux on face rows.  A
  // volume row has nh entries, a face row nq (the face-face block is zero):
  // volume rows are cut into C = ceil(nh / nq) pieces of nq entries so that
  // every work item sums nq pairs; the pieces are added in phase 5.
  if (ea.paired) {
    esdg_pairs_64<DIM, NSP>(a, ea, s_prim, fkeep, s_acc, e0, nacc, gamma);
  } else {
    esdg_rows<DIM, NFLD>(a, ea, s_prim, s_fs, s_acc, e0, ne, nacc, C,
                         [gamma](const Pair& x, const Pair& y, const double* sd,
                                 double* acc) {
                           ec_flux_dot<DIM, NSP>(x, y, sd, gamma, acc);
                         });
  }
  __syncthreads();
  // 5: dQ/dt = -M^-1 V_h^T acc
  esdg_apply<DIM, P, NFLD>(a, ea, s_acc, e0, ne, nacc, C);
}

// }}}

// {{{ entropy-stable Euler RHS for the uiuc_7sp mixture (Renac's flux)
//
// The reference's mixture path (mirgecom/euler.py:108-310 with
// project_fluid_state(entropy_stable=True), gas_model.py:494-588, and
// make_entropy_projected_fluid_state, gas_model.py:1000-1023):
// esdg_project_mix_kernel: at the volume points the interpolated state, its
// entropy variables with the state's own p and gamma and back (the
// "entropy-stable" projected state: energy p/(gamma-1) + rho|u|^2/2), the
// state made again; its entropy variables L2-projected to ev.
// esdg_mix_kernel: at the hybridized points the conserved state from the
// projected variables and the interpolated nodal gamma(cv, T), T by Newton
// from the interpolated nodal T; the neighbours' face states from their
// projected variables with this element's interpolated gamma and their
// nodal T as the seed; boundary faces the entropy-stable projected interior
// state and the condition's exterior state; Renac's two-point flux; phases
// 4-5 as esdg_kernel (row by row).

// The "entropy-stable" projected state of conserved values u with Newton
// seed `seed` (project_fluid_state(..., entropy_stable=True)).
template <int DIM>
__device__ __forceinline__ void es_projected_state(const double* u, double seed,
                                                   const GasDev& g, int niter,
                                                   Node<DIM, MIX_NSP>& s) {
  constexpr int NFLD = 2 + DIM + MIX_NSP;
  Node<DIM, MIX_NSP> s0;
#pragma unroll
  for (int f = 0; f < NFLD; ++f) s0.q[f] = u[f];
  const double t0 = point_temperature<DIM, MIX_NSP>(s0.q, seed, niter);
  node_state<DIM, MIX_NSP, 1, false>(s0, g, t0);
  double v[NFLD];
  cons_to_entropy_p<DIM, MIX_NSP>(s0.q, s0.p, s0.gamma, v);
  entropy_to_cons<DIM, MIX_NSP>(v, s0.gamma, s.q);
  const double t1 = point_temperature<DIM, MIX_NSP>(s.q, seed, niter);
  node_state<DIM, MIX_NSP, 1, false>(s, g, t1);
}

template <int DIM, int P>
__global__ __launch_bounds__(kBlock) void esdg_project_mix_kernel(KArgs a, EsdgArgs ea) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP;
  constexpr int NFLD = 2 + DIM + MIX_NSP;
  extern __shared__ double lds[];
  const int nq = ea.nq, epb = ea.epb;
  double* s_e = lds;  // [F][epb][nq]
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  for (int t = threadIdx.x; t < ne * nq; t += blockDim.x) {
    const int le = t / nq;
    const int m = t - le * nq;
    const int64_t base = (e0 + le) * NP;
    double u[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) u[f] = 0.0;
    double seed = 0.0;
#pragma unroll 1
    for (int i = 0; i < NP; ++i) {
      const double w = ea.interp[i * nq + m];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) u[f] += w * a.q[f * N + base + i];
      seed += w * a.temp[base + i];
    }
    Node<DIM, MIX_NSP> s;
    if (ea.es_volume) {
      es_projected_state<DIM>(u, seed, a.gas, ea.niter, s);
    } else {
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s.q[f] = u[f];
      const double t0 = point_temperature<DIM, MIX_NSP>(s.q, seed, ea.niter);
      node_state<DIM, MIX_NSP, 1, false>(s, a.gas, t0);
    }
    double v[NFLD];
    cons_to_entropy_p<DIM, MIX_NSP>(s.q, s.p, s.gamma, v);
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_e[(f * epb + le) * nq + m] = v[f];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ne * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    double v[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) v[f] = 0.0;
    for (int m = 0; m < nq; ++m) {
      const double w = ea.proj[m * NP + i];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) v[f] += w * s_e[(f * epb + le) * nq + m];
    }
#pragma unroll
    for (int f = 0; f < NFLD; ++f) ea.ev[f * N + e0 * NP + t] = v[f];
  }
}

template <int DIM, int P>
__global__ __launch_bounds__(kBlock) void esdg_mix_kernel(KArgs a, EsdgArgs ea) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF;
  constexpr int NFLD = 2 + DIM + MIX_NSP;
  using Pair = RnPair<DIM>;
  extern __shared__ double lds[];
  const int nq = ea.nq, nfq = ea.nfq, nh = ea.nh, epb = ea.epb, nft = NF * ea.nfq;
  Pair* s_prim = reinterpret_cast<Pair*>(lds);                 // [epb][nh]
  double* s_v = reinterpret_cast<double*>(s_prim + epb * nh);  // [F][epb][NP]
  double* s_q = s_v + NFLD * epb * NP;                         // [F][epb][NP]
  double* s_vp = s_q + NFLD * epb * NP;                        // [F][epb][NF*NFP]
  double* s_tn = s_vp + NFLD * epb * NF * NFP;                 // [epb][NP] nodal T
  double* s_gb = s_tn + epb * NP;                              // [epb][NP] gamma(cv, T)
  double* s_tp = s_gb + epb * NP;                              // [epb][NF*NFP]
  double* s_gf = s_tp + epb * NF * NFP;                        // [epb][nft] gamma_h
  double* s_cf = s_gf + epb * nft;                             // [epb][nft] |u| + c
  const int C = (nh + nq - 1) / nq;
  const int nacc = C * nq + nft;
  double* s_fs = s_cf + epb * nft;                             // [F][epb][nft]
  double* s_qf = s_fs + NFLD * epb * nft;                      // [F][epb][nft], 2-3
  double* s_acc = s_qf;                                        // [F][epb][nacc], 4-5
  const int64_t e0 = int64_t(blockIdx.x) * epb;
  const int ne = a.K - e0 < epb ? int(a.K - e0) : epb;
  const int64_t N = a.nnodes;
  // 1: own v~, q, T and gamma(cv, T); the neighbours' v~ and T at this
  // element's face nodes
  for (int t = threadIdx.x; t < NFLD * ne * NP; t += blockDim.x) {
    const int f = t / (ne * NP);
    const int r = t - f * ne * NP;
    s_v[f * epb * NP + r] = ea.ev[f * N + e0 * NP + r];
    s_q[f * epb * NP + r] = a.q[f * N + e0 * NP + r];
  }
  for (int t = threadIdx.x; t < ne * NP; t += blockDim.x) {
    const int64_t g = e0 * NP + t;
    Node<DIM, MIX_NSP> s;
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s.q[f] = a.q[f * N + g];
    node_state<DIM, MIX_NSP, 1, false>(s, a.gas, a.temp[g]);
    s_tn[t] = a.temp[g];
    s_gb[t] = s.gamma;
  }
  for (int t = threadIdx.x; t < ne * NF * NFP; t += blockDim.x) {
    const int le = t / (NF * NFP);
    const int fn = t - le * NF * NFP;
    const int32_t nb = a.vmap_p[(e0 + le) * NF * NFP + fn];
    if (nb < 0) continue;
#pragma unroll
    for (int f = 0; f < NFLD; ++f)
      s_vp[(f * epb + le) * NF * NFP + fn] = ea.ev[f * N + nb];
    s_tp[le * NF * NFP + fn] = a.temp[nb];
  }
  __syncthreads();
  // 2: the state at the hybridized points
  for (int t = threadIdx.x; t < ne * nh; t += blockDim.x) {
    const int le = t / nh;
    const int h = t - le * nh;
    double v[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) v[f] = 0.0;
    double gh = 0.0, seed = 0.0;
#pragma unroll 1
    for (int i = 0; i < NP; ++i) {
      const double w = ea.vh[i * nh + h];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) v[f] += w * s_v[(f * epb + le) * NP + i];
      gh += w * s_gb[le * NP + i];
      seed += w * s_tn[le * NP + i];
    }
    Node<DIM, MIX_NSP> s;
    entropy_to_cons<DIM, MIX_NSP>(v, gh, s.q);
    const double th = point_temperature<DIM, MIX_NSP>(s.q, seed, ea.niter);
    node_state<DIM, MIX_NSP, 1, false>(s, a.gas, th);
    RnPrim<DIM> pr;
    rn_prim<DIM>(s, pr);
    s_prim[le * nh + h] = pr;
    if (h >= nq) {
      const int m = h - nq;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) s_qf[(f * epb + le) * nft + m] = s.q[f];
      s_gf[le * nft + m] = gh;
      s_cf[le * nft + m] = pr.c;
    }
  }
  __syncthreads();
  // 3: entropy-stable face flux at the face points, times fscale * w
  for (int t = threadIdx.x; t < ne * nft; t += blockDim.x) {
    const int le = t / nft;
    const int m = t - le * nft;
    const int fc = m / nfq;
    const int mm = m - fc * nfq;
    const int64_t e = e0 + le;
    const double* w = ea.finterp + fc * NFP * nfq + mm;
    const int32_t v0 = a.vmap_p[(e * NF + fc) * NFP];
    double qm[NFLD], qp[NFLD];
    RnPrim<DIM> pm, pp;
    Node<DIM, MIX_NSP> sp;
    if (v0 >= 0) {
#pragma unroll
      for (int f = 0; f < NFLD; ++f) qm[f] = s_qf[(f * epb + le) * nft + m];
      static_cast<Pair&>(pm) = s_prim[le * nh + nq + m];
      pm.c = s_cf[le * nft + m];
      double vp[NFLD];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) vp[f] = 0.0;
      double seed = 0.0;
#pragma unroll 1
      for (int j = 0; j < NFP; ++j) {
        const double wj = w[j * nfq];
#pragma unroll
        for (int f = 0; f < NFLD; ++f)
          vp[f] += wj * s_vp[(f * epb + le) * NF * NFP + fc * NFP + j];
        seed += wj * s_tp[le * NF * NFP + fc * NFP + j];
      }
      entropy_to_cons<DIM, MIX_NSP>(vp, s_gf[le * nft + m], sp.q);
      const double tp = point_temperature<DIM, MIX_NSP>(sp.q, seed, ea.niter);
      node_state<DIM, MIX_NSP, 1, false>(sp, a.gas, tp);
    } else {
      // boundary: the entropy-stable projected interior state and the
      // boundary condition's exterior state
      double ui[NFLD];
#pragma unroll
      for (int f = 0; f < NFLD; ++f) ui[f] = 0.0;
      double seed = 0.0;
#pragma unroll 1
      for (int j = 0; j < NFP; ++j) {
        const double wj = w[j * nfq];
        const int ij = a.fmask[fc * NFP + j];
#pragma unroll
        for (int f = 0; f < NFLD; ++f) ui[f] += wj * s_q[(f * epb + le) * NP + ij];
        seed += wj * s_tn[le * NP + ij];
      }
      Node<DIM, MIX_NSP> sm;
      es_projected_state<DIM>(ui, seed, a.gas, ea.niter, sm);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) qm[f] = sm.q[f];
      rn_prim<DIM>(sm, pm);
      const int64_t bq = (-int64_t(v0) - 1) / NFP * nfq + mm;
#pragma unroll
      for (int f = 0; f < NFLD; ++f) sp.q[f] = ea.qbq[f * ea.nbq + bq];
      node_state<DIM, MIX_NSP, 1, false>(sp, a.gas, ea.tbq[bq]);
    }
#pragma unroll
    for (int f = 0; f < NFLD; ++f) qp[f] = sp.q[f];
    rn_prim<DIM>(sp, pp);
    double n[DIM], fs[NFLD];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + fc) * DIM + d];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) fs[f] = 0.0;
    rn_flux_dot<DIM>(pm, pp, n, fs);
    if (ea.dissipation) {
      const double alpha = fmax(pm.c, pp.c);
#pragma unroll
      for (int f = 0; f < NFLD; ++f) fs[f] -= 0.5 * alpha * (qp[f] - qm[f]);
    }
    const double sc = a.fscale[e * NF + fc] * ea.wf[mm];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_fs[(f * epb + le) * nft + m] = sc * fs[f];
  }
  __syncthreads();
  // 4: rows of (Q_h - Q_h^T) o F_S, plus the face flux on face rows
  esdg_rows<DIM, NFLD>(a, ea, s_prim, s_fs, s_acc, e0, ne, nacc, C,
                       [](const Pair& x, const Pair& y, const double* sd, double* acc) {
                         rn_flux_dot<DIM>(x, y, sd, acc);
                       });
  __syncthreads();
  // 5: dQ/dt = -M^-1 V_h^T acc
  esdg_apply<DIM, P, NFLD>(a, ea, s_acc, e0, ne, nacc, C);
}

// }}}

// {{{ nodal net flux from stored gradients
//
// The pointwise phases (5 and 6) of grad_kernel<FLUX> as node-parallel
// passes over gradients that grad_kernel<FLUX=false> stored.  Used for the
// 12-field mixture: its flux kernel holds only 4 elements (40 nodes at p = 2)
// per workgroup, so the pointwise phase, which carries the NASA7 and
// transport work, kept 40 of 256 lanes busy.

//
// F32: the mixed-precision form, from FP32 gradients (a.grad32) to FP32 net
// fluxes (a.fnet32_out) that exclude p_ref I, as the mixed flux kernel
// writes them.
template <int DIM, int NSP, int MIX, bool F32, typename A>
__device__ __forceinline__ void node_flux_body(const A& a, int64_t g, double p_ref) {
  constexpr int NFLD = 2 + DIM + NSP;
  const int64_t N = a.nnodes;
  Node<DIM, NSP> s;
  load_q<DIM, NSP>(s, a.q, N, g);
  node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.temp[g] : 0.0);
  double gq[NFLD][DIM], gt[DIM], F[NFLD][DIM];
  if constexpr (F32)
    load_grad<DIM, NSP>(gq, gt, a.grad32, N, g);
  else
    load_grad<DIM, NSP>(gq, gt, a.grad, N, g);
  viscous_flux<DIM, NSP, MIX>(s, gq, gt, F);
  double FI[NFLD][DIM];
  if (a.inviscid_on) {
    inviscid_flux<DIM, NSP>(s, FI);
  } else {
#pragma unroll
    for (int f = 0; f < NFLD; ++f)
#pragma unroll
      for (int d = 0; d < DIM; ++d) FI[f][d] = 0.0;
  }
#pragma unroll
  for (int f = 0; f < NFLD; ++f)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      if constexpr (F32)
        a.fnet32_out[(f * DIM + d) * N + g] =
            float(FI[f][d] - F[f][d] - ((f == 2 + d) ? p_ref : 0.0));
      else
        a.out[(f * DIM + d) * N + g] = FI[f][d] - F[f][d];
    }
}

template <int DIM, int NSP, int MIX, bool F32 = false>
__global__ __launch_bounds__(kBlock, 2) void node_flux_kernel(KArgs a) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double p_ref = 0.0;
  if constexpr (F32) {
    __shared__ double s_pref;
    if (threadIdx.x == 0) s_pref = reference_pressure<DIM, NSP, MIX>(a);
    __syncthreads();
    p_ref = s_pref;
  }
  if (g < a.nnodes) node_flux_body<DIM, NSP, MIX, F32>(a, g, p_ref);
}

// Boundary face nodes (one thread per element face node; interior ones
// return at once): the exterior net normal flux with the interior gradients
// (boundary.py:944-970), or, on Mengaldo nodes, the interior gradients
// copied to wwork for wall_flux_kernel.
// F32: gradients from a.grad32 and exterior fluxes that exclude p_ref n (the
// mixed-precision form).
template <int DIM, int P, int NSP, int MIX, bool F32 = false>
__global__ __launch_bounds__(kBlock) void bnd_node_flux_kernel(KArgs a) {
  using E = Elem<DIM, P>;
  constexpr int NP = E::NP, NFP = E::NFP, NFN = E::NFN, NF = E::NF;
  constexpr int NFLD = 2 + DIM + NSP, NS = NFLD + 1;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double p_ref = 0.0;
  if constexpr (F32) {
    __shared__ double s_pref;
    if (threadIdx.x == 0) s_pref = reference_pressure<DIM, NSP, MIX>(a);
    __syncthreads();
    p_ref = s_pref;
  }
  if (t >= a.K * NFN) return;
  const int32_t nb = a.vmap_p[t];
  if (nb >= 0) return;
  const int64_t e = t / NFN;
  const int fn = int(t - e * NFN);
  const int64_t b = -int64_t(nb) - 1;
  const int f = fn / NFP;
  const int64_t N = a.nnodes;
  const int64_t gi = e * NP + a.fmask[fn];
  if (wall_flags(a, b) & kBcMengaldo) {
#pragma unroll
    for (int k = 0; k < NS * DIM; ++k)
      a.wwork[k * a.nbnd + b] = F32 ? double(a.grad32[k * N + gi]) : a.grad[k * N + gi];
    return;
  }
  double n[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + f) * DIM + d];
  Node<DIM, NSP> s;
  load_q<DIM, NSP>(s, a.qb, a.nbnd, b);
  node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.tb[b] : 0.0);
  double gq[NFLD][DIM], gt[DIM], fv[NFLD], fi[NFLD];
  if constexpr (F32)
    load_grad<DIM, NSP>(gq, gt, a.grad32, N, gi);
  else
    load_grad<DIM, NSP>(gq, gt, a.grad, N, gi);
  if (a.inviscid_on) {
    inviscid_flux_n<DIM, NSP>(s, n, fi);
  } else {
#pragma unroll
    for (int k = 0; k < NFLD; ++k) fi[k] = 0.0;
  }
  viscous_flux_n<DIM, NSP, MIX>(s, gq, gt, n, fv);
#pragma unroll
  for (int k = 0; k < NFLD; ++k)
    a.fnb_out[k * a.nbnd + b] =
        fi[k] - fv[k] - ((k >= 2 && k < 2 + DIM) ? p_ref * n[k - 2] : 0.0);
}

// }}}

// {{{ Mengaldo boundary nodes
//
// One thread per boundary node; nodes without MG_BC_MENGALDO return at once.
// FUSED: the gradients come from wwork (stored by grad_kernel<FLUX>) and the
// node's exterior net normal flux F_I(q+).n - [2 F_V(bc) - F_V(q-)].n (- p_ref n
// for the mixed path) is written to fnb_out.  Otherwise the gradients come
// from the full gradient array and the effective exterior viscous flux
// 2 F_V(bc).n - F_V(q-).n is written to wwork for rhs_kernel<kGrad>.
// SHIFT (FUSED only): the exterior fluxes exclude p_ref n (mixed precision).
template <int DIM, int NSP, int MIX, bool FUSED, bool SHIFT>
__global__ __launch_bounds__(kBlock) void wall_flux_kernel(KArgs a) {
  constexpr int NFLD = 2 + DIM + NSP;
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= a.nbnd) return;
  const int fl = a.wflag[b];
  if (!(fl & kBcMengaldo)) return;
  const int64_t gm = a.wvol[b];
  double n[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) n[d] = a.wn[d * a.nbnd + b];
  double gq[NFLD][DIM], gt[DIM], fv[NFLD];
  if constexpr (FUSED) {
#pragma unroll
    for (int k = 0; k < NFLD; ++k)
#pragma unroll
      for (int d = 0; d < DIM; ++d) gq[k][d] = a.wwork[(k * DIM + d) * a.nbnd + b];
#pragma unroll
    for (int d = 0; d < DIM; ++d) gt[d] = a.wwork[(NFLD * DIM + d) * a.nbnd + b];
  } else {
    load_grad<DIM, NSP>(gq, gt, a.grad, a.nnodes, gm);
  }
  wall_viscous_flux_n<DIM, NSP, MIX>(a, b, gm, fl, n, gq, gt, fv);
  if constexpr (FUSED) {
    double fi[NFLD];
    if (a.inviscid_on) {
      Node<DIM, NSP> s;
      load_q<DIM, NSP>(s, a.qb, a.nbnd, b);
      node_state<DIM, NSP, MIX, false>(s, a.gas, MIX ? a.tb[b] : 0.0);
      inviscid_flux_n<DIM, NSP>(s, n, fi);
    } else {
#pragma unroll
      for (int k = 0; k < NFLD; ++k) fi[k] = 0.0;
    }
    double p_ref = 0.0;
    if constexpr (SHIFT) p_ref = reference_pressure<DIM, NSP, MIX>(a);
#pragma unroll
    for (int k = 0; k < NFLD; ++k)
      a.fnb_out[k * a.nbnd + b] =
          fi[k] - fv[k] - ((k >= 2 && k < 2 + DIM) ? p_ref * n[k - 2] : 0.0);
  } else {
#pragma unroll
    for (int k = 0; k < NFLD; ++k) a.wwork[k * a.nbnd + b] = fv[k];
  }
}

// }}}

// {{{ pointwise mixture kernels

template <int DIM>
__global__ __launch_bounds__(kBlock) void mix_temperature_kernel(
    int64_t n, const double* __restrict__ q, const double* __restrict__ tseed,
    int niter, double* __restrict__ tout) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double rho = q[g];
  double mm = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) mm += q[(2 + d) * n + g] * q[(2 + d) * n + g];
  const double e = (q[n + g] - 0.5 * mm / rho) / rho;
  double y[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = q[(2 + DIM + k) * n + g] / rho;
  tout[g] = mix_temperature(e, tseed[g], y, niter);
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void mix_temperature_tol_kernel(
    int64_t n, const double* __restrict__ q, const double* __restrict__ tseed, double tol,
    int max_iter, double* __restrict__ tout, int32_t* __restrict__ iters) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double rho = q[g];
  double mm = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) mm += q[(2 + d) * n + g] * q[(2 + d) * n + g];
  const double e = (q[n + g] - 0.5 * mm / rho) / rho;
  double y[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = q[(2 + DIM + k) * n + g] / rho;
  int it;
  tout[g] = mix_temperature_tol(e, tseed[g], y, tol, max_iter, &it);
  if (iters) iters[g] = it;
}

// src [7][n] = W_k omega_k, or with ADD rhs species rows += W_k omega_k
// (rhs [F][n], the species rows start at 2 + DIM).
template <int DIM, bool ADD>
__global__ __launch_bounds__(kBlock) void species_source_kernel(
    int64_t n, const double* __restrict__ q, const double* __restrict__ temp,
    double* __restrict__ src) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double rho = q[g];
  double y[MIX_NSP], s[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = q[(2 + DIM + k) * n + g] / rho;
  mix_species_source(rho, temp[g], y, s);
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    if constexpr (ADD) {
      double* r = src + (2 + DIM + k) * n + g;
      *r = *r + s[k];
    } else {
      src[k * n + g] = s[k];
    }
  }
}

// Node-local stable timestep (inviscid.py:379-422, viscous.py:547-607):
//   dt = h_e / (|v| + c)                                   inviscid
//   dt = h_e / (|v| + c + (mu / rho + max_k d_k) / h_e)     viscous
// with h_e the characteristic length of the node's element.
template <int DIM, int NSP, int MIX>
__global__ __launch_bounds__(kBlock) void local_timestep_kernel(
    int64_t n, int np, const double* __restrict__ q, const double* __restrict__ temp,
    const double* __restrict__ hlen, GasDev gas, int viscous,
    double* __restrict__ dt) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  Node<DIM, NSP> s;
  load_q<DIM, NSP>(s, q, n, g);
  node_state<DIM, NSP, MIX, false>(s, gas, MIX ? temp[g] : 0.0);
  const double h = hlen[g / np];
  double rate = s.speed + s.c;
  if (viscous) {
    node_transport<DIM, NSP, MIX>(s, gas);
    double dmax = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) dmax = k == 0 ? s.d[0] : fmax(dmax, s.d[k]);
    rate += (s.mu * s.inv_rho + dmax) / h;
  }
  dt[g] = h / rate;
}

// Transport variables of every node, from the same node_state /
// node_transport code the Navier-Stokes kernels run.
template <int DIM, int NSP, int MIX>
__global__ __launch_bounds__(kBlock) void transport_vars_kernel(
    GasDev gas, int64_t n, const double* __restrict__ q,
    const double* __restrict__ temp, double* mu, double* mu_b, double* kappa,
    double* diff) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  Node<DIM, NSP> s;
#pragma unroll
  for (int f = 0; f < Node<DIM, NSP>::NFLD; ++f) s.q[f] = q[f * n + g];
  node_state<DIM, NSP, MIX, true>(s, gas, MIX ? temp[g] : 0.0);
  mu[g] = s.mu;
  mu_b[g] = s.mu_b;
  kappa[g] = s.kappa;
#pragma unroll
  for (int k = 0; k < NSP; ++k) diff[k * n + g] = s.d[k];
}

static __global__ __launch_bounds__(kBlock) void rk_axpy_kernel(
    int64_t n, const double* y, const double* k, double a, double* out,
    double* acc, double w) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double kv = k[g];
  out[g] = y[g] + a * kv;
  if (acc) acc[g] = acc[g] + w * kv;
}


// Bound-preserving limiter (mirgecom/limiter.py:39-125; Zhang & Shu 2011,
// eq. 2.9): u <- avg + theta (u - avg) per element, theta the largest factor
// in [0, 1] that brings the element's extrema inside [mmin, mmax].  The cell
// average is the mass-weighted mean: on an affine simplex the Jacobian is
// constant and cancels, so avg = sum_j w_j u_j / sum_j w_j with w the row
// sums of the reference mass matrix.  One thread per element; a block stages
// kLimElems consecutive elements of one field through LDS so that loads and
// stores stay contiguous (the element rows are np doubles long).
constexpr int kLimElems = 128;
constexpr int kLimFlagMax = 1;
constexpr int kLimFlagModifyAvg = 2;

template <int NP>
__global__ __launch_bounds__(kLimElems) void bound_limiter_kernel(
    int64_t nelements, const double* __restrict__ weights,
    const double* __restrict__ field, double* out, double mmin, double mmax,
    int flags) {
  constexpr int S = NP | 1;  // odd LDS row stride
  __shared__ double tile[kLimElems * S];
  const int64_t e0 = int64_t(blockIdx.x) * kLimElems;
  const int64_t ne = nelements - e0 < kLimElems ? nelements - e0 : kLimElems;
  const int64_t base = (int64_t(blockIdx.y) * nelements + e0) * NP;
  const int span = int(ne) * NP;
  const int t = threadIdx.x;
  for (int i = t; i < span; i += kLimElems) {
    const int e = i / NP;
    tile[e * S + (i - e * NP)] = field[base + i];
  }
  __syncthreads();
  if (t < ne) {
    double* u = tile + t * S;
    double num = 0.0, den = 0.0, umin = u[0], umax = u[0];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      num += weights[j] * u[j];
      den += weights[j];
      umin = fmin(umin, u[j]);
      umax = fmax(umax, u[j]);
    }
    double avg = num / fabs(den);
    const bool modify = (flags & kLimFlagModifyAvg) != 0;
    if (modify && !(avg > mmin)) avg = mmin;
    double theta = 1.0;
    if (umin < mmin) theta = fmin(theta, fabs((mmin - avg) / (umin - avg + 1e-13)));
    if (flags & kLimFlagMax) {
      if (modify && avg > mmax) avg = mmax;
      if (umax > mmax) theta = fmin(theta, fabs((mmax - avg) / (umax - avg + 1e-13)));
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) u[j] = theta * (u[j] - avg) + avg;
  }
  __syncthreads();
  for (int i = t; i < span; i += kLimElems) {
    const int e = i / NP;
    out[base + i] = tile[e * S + (i - e * NP)];
  }
}

// }}}

}  // namespace mgdg
