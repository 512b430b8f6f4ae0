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
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/mgdg.h"
#include "mgdg_physics.h"

namespace mgdg {

__constant__ MechDev c_mech;

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

struct KArgs {
  const double* q;
  const double* temp;
  const double* grad;
  const double* qb;
  const double* tb;
  double* out;
  const double* rx;
  const double* normals;
  const double* fscale;
  const int32_t* vmap_p;
  const double* dw;
  const double* lift;
  const int32_t* fmask;
  int64_t K;
  int64_t nbnd;
  GasDev gas;
  int flux_kind;
  int inviscid_on;
};

template <int DIM, int P>
__device__ __forceinline__ void load_matrices(double* s_dw, double* s_lift,
                                              int* s_fmask, const KArgs& a) {
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

template <int DIM, int NSP>
__device__ __forceinline__ void load_q(Node<DIM, NSP>& s, const double* q,
                                       int64_t stride, int64_t g) {
#pragma unroll
  for (int f = 0; f < Node<DIM, NSP>::NFLD; ++f) s.q[f] = q[f * stride + g];
}

template <int DIM, int NSP>
__device__ __forceinline__ void load_grad(double (*gq)[DIM], double* gt,
                                          const double* grad, int64_t stride,
                                          int64_t g) {
  constexpr int NFLD = 2 + DIM + NSP;
#pragma unroll
  for (int f = 0; f < NFLD; ++f)
#pragma unroll
    for (int d = 0; d < DIM; ++d) gq[f][d] = grad[(f * DIM + d) * stride + g];
#pragma unroll
  for (int d = 0; d < DIM; ++d) gt[d] = grad[(NFLD * DIM + d) * stride + g];
}

// {{{ divergence (Euler / Navier-Stokes RHS)

template <int DIM, int P, int NSP, bool MIX, bool VISC>
struct RhsCfg {
  using E = Elem<DIM, P>;
  static constexpr int NFLD = 2 + DIM + NSP;
  static constexpr int PER_ELEM = NFLD * DIM * E::NP + NFLD * E::NFN;
  static constexpr int EPB_LDS = (kLdsBudget / 8 - E::MAT - E::NFN) / PER_ELEM;
  static constexpr int EPB_THR = kBlock / E::NP;
  static constexpr int EPB = EPB_LDS < EPB_THR ? (EPB_LDS > 0 ? EPB_LDS : 1) : EPB_THR;
  static constexpr size_t LDS = sizeof(double) * (E::MAT + EPB * PER_ELEM) +
                                sizeof(int) * E::NFN;
  static_assert(LDS <= 160 * 1024, "LDS per workgroup exceeds 160 KiB");
};

template <int DIM, int P, int NSP, bool MIX, bool VISC>
__global__ __launch_bounds__(kBlock) void rhs_kernel(KArgs a) {
  using E = Elem<DIM, P>;
  using C = RhsCfg<DIM, P, NSP, MIX, VISC>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF, NFN = E::NFN;
  constexpr int NFLD = C::NFLD, EPB = C::EPB;
  extern __shared__ double smem[];
  double* s_dw = smem;
  double* s_lift = s_dw + DIM * NP * NP;
  double* s_vol = s_lift + NFN * NP;             // [EPB][NFLD][DIM][NP]
  double* s_face = s_vol + EPB * NFLD * DIM * NP; // [EPB][NFLD][NFN]
  int* s_fmask = reinterpret_cast<int*>(s_face + EPB * NFLD * NFN);

  const int64_t N = a.K * NP;
  const int64_t e0 = int64_t(blockIdx.x) * EPB;
  load_matrices<DIM, P>(s_dw, s_lift, s_fmask, a);

  // phase 1: volume fluxes
  for (int t = threadIdx.x; t < EPB * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int64_t g = e * NP + i;
    Node<DIM, NSP> s;
    load_q<DIM, NSP>(s, a.q, N, g);
    node_state<DIM, NSP, MIX>(s, a.gas, MIX ? a.temp[g] : 0.0);
    double F[NFLD][DIM];
    if (a.inviscid_on) {
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
    double rxe[DIM][DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int d = 0; d < DIM; ++d) rxe[r][d] = a.rx[(e * DIM + r) * DIM + d];
#pragma unroll
    for (int f = 0; f < NFLD; ++f)
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double gsum = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) gsum += rxe[r][d] * F[f][d];
        s_vol[((le * NFLD + f) * DIM + r) * NP + i] = gsum;
      }
  }
  __syncthreads();

  // phase 2: face fluxes
  for (int t = threadIdx.x; t < EPB * NFN; t += blockDim.x) {
    const int le = t / NFN;
    const int fn = t - le * NFN;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int f = fn / NFP;
    const int64_t gm = e * NP + s_fmask[fn];
    const int32_t nb = a.vmap_p[e * NFN + fn];
    double n[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n[d] = a.normals[(e * NF + f) * DIM + d];
    const double fs = a.fscale[e * NF + f];
    Node<DIM, NSP> sm, sp;
    load_q<DIM, NSP>(sm, a.q, N, gm);
    double tp = 0.0;
    const bool bnd = nb < 0;
    const int64_t b = bnd ? int64_t(-nb - 1) : 0;
    if (!bnd) {
      load_q<DIM, NSP>(sp, a.q, N, nb);
      if (MIX) tp = a.temp[nb];
    } else {
      load_q<DIM, NSP>(sp, a.qb, a.nbnd, b);
      if (MIX) tp = a.tb[b];
    }
    node_state<DIM, NSP, MIX>(sm, a.gas, MIX ? a.temp[gm] : 0.0);
    node_state<DIM, NSP, MIX>(sp, a.gas, tp);
    double fl[NFLD];
    if (a.inviscid_on) {
      inviscid_numflux<DIM, NSP>(sm, sp, n, a.flux_kind, fl);
    } else {
#pragma unroll
      for (int k = 0; k < NFLD; ++k) fl[k] = 0.0;
    }
    if constexpr (VISC) {
      double gq[NFLD][DIM], gt[DIM], fvm[NFLD][DIM], fvp[NFLD][DIM];
      load_grad<DIM, NSP>(gq, gt, a.grad, N, gm);
      viscous_flux<DIM, NSP, MIX>(sm, gq, gt, fvm);
      if (!bnd) load_grad<DIM, NSP>(gq, gt, a.grad, N, nb);
      viscous_flux<DIM, NSP, MIX>(sp, gq, gt, fvp);
#pragma unroll
      for (int k = 0; k < NFLD; ++k) {
        double c = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) c += (fvp[k][d] + fvm[k][d]) / 2 * n[d];
        fl[k] -= c;
      }
    }
#pragma unroll
    for (int k = 0; k < NFLD; ++k) s_face[(le * NFLD + k) * NFN + fn] = fs * fl[k];
  }
  __syncthreads();

  // phase 3: weak divergence + lift
  for (int t = threadIdx.x; t < EPB * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int64_t g = e * NP + i;
#pragma unroll 1
    for (int f = 0; f < NFLD; ++f) {
      const double* gv = s_vol + (le * NFLD + f) * DIM * NP;
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int j = 0; j < NP; ++j) acc += s_dw[(r * NP + j) * NP + i] * gv[r * NP + j];
      const double* fv = s_face + (le * NFLD + f) * NFN;
      double lf = 0.0;
#pragma unroll
      for (int m = 0; m < NFN; ++m) lf += s_lift[m * NP + i] * fv[m];
      a.out[f * N + g] = acc - lf;
    }
  }
}

// }}}

// {{{ gradient

template <int DIM, int P, int NSP, bool MIX>
struct GradCfg {
  using E = Elem<DIM, P>;
  static constexpr int NS = 3 + DIM + NSP;  // conserved fields + T
  static constexpr int PER_ELEM = NS * E::NP + NS * E::NFN;
  static constexpr int EPB_LDS = (kLdsBudget / 8 - E::MAT - E::NFN) / PER_ELEM;
  static constexpr int EPB_THR = kBlock / E::NP;
  static constexpr int EPB = EPB_LDS < EPB_THR ? (EPB_LDS > 0 ? EPB_LDS : 1) : EPB_THR;
  static constexpr size_t LDS = sizeof(double) * (E::MAT + EPB * PER_ELEM) +
                                sizeof(int) * E::NFN;
  static_assert(LDS <= 160 * 1024, "LDS per workgroup exceeds 160 KiB");
};

template <int DIM, int NSP, bool MIX>
__device__ __forceinline__ double node_temperature(const double* qv,
                                                   const GasDev& g, double tmix) {
  if constexpr (MIX) {
    return tmix;
  } else {
    double mm = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) mm += qv[2 + d] * qv[2 + d];
    const double e_int = qv[1] - 0.5 * mm / qv[0];
    return ((g.gamma - 1.0) / g.gas_const) * e_int / qv[0];
  }
}

template <int DIM, int P, int NSP, bool MIX>
__global__ __launch_bounds__(kBlock) void grad_kernel(KArgs a) {
  using E = Elem<DIM, P>;
  using C = GradCfg<DIM, P, NSP, MIX>;
  constexpr int NP = E::NP, NFP = E::NFP, NF = E::NF, NFN = E::NFN;
  constexpr int NFLD = 2 + DIM + NSP, NS = C::NS, EPB = C::EPB;
  extern __shared__ double smem[];
  double* s_dw = smem;
  double* s_lift = s_dw + DIM * NP * NP;
  double* s_u = s_lift + NFN * NP;           // [EPB][NS][NP]
  double* s_us = s_u + EPB * NS * NP;        // [EPB][NS][NFN]
  int* s_fmask = reinterpret_cast<int*>(s_us + EPB * NS * NFN);

  const int64_t N = a.K * NP;
  const int64_t e0 = int64_t(blockIdx.x) * EPB;
  load_matrices<DIM, P>(s_dw, s_lift, s_fmask, a);

  for (int t = threadIdx.x; t < EPB * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int64_t g = e * NP + i;
    double qv[NFLD];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) qv[f] = a.q[f * N + g];
#pragma unroll
    for (int f = 0; f < NFLD; ++f) s_u[(le * NS + f) * NP + i] = qv[f];
    s_u[(le * NS + NFLD) * NP + i] =
        node_temperature<DIM, NSP, MIX>(qv, a.gas, MIX ? a.temp[g] : 0.0);
  }
  __syncthreads();

  for (int t = threadIdx.x; t < EPB * NFN; t += blockDim.x) {
    const int le = t / NFN;
    const int fn = t - le * NFN;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int f = fn / NFP;
    const int im = s_fmask[fn];
    const int32_t nb = a.vmap_p[e * NFN + fn];
    const double fs = a.fscale[e * NF + f];
    if (nb >= 0) {
      double qp[NFLD];
#pragma unroll
      for (int k = 0; k < NFLD; ++k) qp[k] = a.q[k * N + nb];
      const double tp = node_temperature<DIM, NSP, MIX>(qp, a.gas, MIX ? a.temp[nb] : 0.0);
#pragma unroll
      for (int k = 0; k < NFLD; ++k)
        s_us[(le * NS + k) * NFN + fn] = fs * ((qp[k] + s_u[(le * NS + k) * NP + im]) / 2);
      s_us[(le * NS + NFLD) * NFN + fn] =
          fs * ((tp + s_u[(le * NS + NFLD) * NP + im]) / 2);
    } else {
      const int64_t b = -int64_t(nb) - 1;
      double qp[NFLD];
#pragma unroll
      for (int k = 0; k < NFLD; ++k) qp[k] = a.qb[k * a.nbnd + b];
      const double tp = node_temperature<DIM, NSP, MIX>(qp, a.gas, MIX ? a.tb[b] : 0.0);
#pragma unroll
      for (int k = 0; k < NFLD; ++k) s_us[(le * NS + k) * NFN + fn] = fs * qp[k];
      s_us[(le * NS + NFLD) * NFN + fn] = fs * tp;
    }
  }
  __syncthreads();

  for (int t = threadIdx.x; t < EPB * NP; t += blockDim.x) {
    const int le = t / NP;
    const int i = t - le * NP;
    const int64_t e = e0 + le;
    if (e >= a.K) continue;
    const int64_t g = e * NP + i;
    double rxe[DIM][DIM], nrm[NF][DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int d = 0; d < DIM; ++d) rxe[r][d] = a.rx[(e * DIM + r) * DIM + d];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int d = 0; d < DIM; ++d) nrm[f][d] = a.normals[(e * NF + f) * DIM + d];
#pragma unroll 1
    for (int s = 0; s < NS; ++s) {
      const double* u = s_u + (le * NS + s) * NP;
      const double* us = s_us + (le * NS + s) * NFN;
      double dr[DIM];
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j) acc += s_dw[(r * NP + j) * NP + i] * u[j];
        dr[r] = acc;
      }
      double lf[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < NFP; ++m) acc += s_lift[(f * NFP + m) * NP + i] * us[f * NFP + m];
        lf[f] = acc;
      }
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        double vsum = 0.0, fsum = 0.0;
#pragma unroll
        for (int r = 0; r < DIM; ++r) vsum += rxe[r][d] * dr[r];
#pragma unroll
        for (int f = 0; f < NF; ++f) fsum += nrm[f][d] * lf[f];
        a.out[(s * DIM + d) * N + g] = fsum - vsum;
      }
    }
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
__global__ __launch_bounds__(kBlock) void species_source_kernel(
    int64_t n, const double* __restrict__ q, const double* __restrict__ temp,
    double* __restrict__ src) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double rho = q[g];
  const double t = temp[g];
  double conc[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    const double c = c_mech.inv_wt[k] * rho * (q[(2 + DIM + k) * n + g] / rho);
    conc[k] = c < 0.0 ? 0.0 : c;
  }
  const double logt = log(t);
  double kf[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    kf[r] = exp(c_mech.log_a[r] + c_mech.b[r] * logt - c_mech.ea_r[r] / t);
  // reverse of R2 (CO + 0.5 O2 <=> CO2): 1/Kc = exp(dG/RT + 0.5 ln(p0/RT))
  double g_rt[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double cp_r, h_rt;
    nasa7_cp_h(k, t, cp_r, h_rt);
    g_rt[k] = h_rt - nasa7_s_r(k, t);
  }
  const double c0 = log(c_mech.one_atm / (c_mech.gas_constant * t));
  const double log_rkc = g_rt[2] - (g_rt[3] + 0.5 * g_rt[1]) + 0.5 * c0;
  const double r1 = kf[0] * pow(conc[0], 0.5) * pow(conc[1], 0.65);
  const double r2 = kf[1] * (conc[3] * pow(conc[1], 0.5) - exp(log_rkc) * conc[2]);
  const double r3 = kf[2] * pow(conc[5], 0.75) * pow(conc[1], 0.5);
  const double w[MIX_NSP] = {-r1, -r1 - 0.5 * r2 - 0.5 * r3, r2, 2 * r1 - r2, r3,
                             2 * r1 - r3, 0.0};
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) src[k * n + g] = c_mech.wt[k] * w[k];
}

__global__ __launch_bounds__(kBlock) void rk_axpy_kernel(
    int64_t n, const double* y, const double* k, double a, double* out,
    double* acc, double w) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double kv = k[g];
  out[g] = y[g] + a * kv;
  if (acc) acc[g] = acc[g] + w * kv;
}

// }}}

}  // namespace mgdg

// {{{ host side: dispatch and C ABI

namespace {

thread_local std::string g_err;

mg_status fail(mg_status code, const std::string& msg) {
  g_err = msg;
  return code;
}

mg_status check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(MG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return MG_OK;
}

mgdg::GasDev to_dev(const mg_gas* g) {
  mgdg::GasDev d{};
  d.eos_kind = g->eos_kind;
  d.transport_kind = g->transport_kind;
  d.gamma = g->gamma;
  d.gas_const = g->gas_const;
  d.mu = g->mu;
  d.mu_b = g->mu_b;
  d.kappa = g->kappa;
  d.pl_alpha = g->pl_alpha;
  d.pl_beta = g->pl_beta;
  d.pl_sigma = g->pl_sigma;
  d.pl_n = g->pl_n;
  d.has_lewis = g->has_lewis;
  for (int k = 0; k < 8; ++k) d.diff[k] = g->species_diffusivity[k];
  return d;
}

mg_status check_mesh(const mg_mesh* m) {
  using mgdg::binom;
  if (!m) return fail(MG_ERR_ARG, "mesh is NULL");
  if (m->dim < 2 || m->dim > 3) return fail(MG_ERR_UNSUPPORTED, "dim must be 2 or 3");
  if (m->order < 1 || m->order > 4) return fail(MG_ERR_UNSUPPORTED, "order must be 1..4");
  if (m->np != binom(m->order + m->dim, m->dim) ||
      m->nfp != binom(m->order + m->dim - 1, m->dim - 1) || m->nfaces != m->dim + 1)
    return fail(MG_ERR_ARG, "np/nfp/nfaces inconsistent with dim and order");
  if (m->nelements <= 0) return fail(MG_ERR_ARG, "nelements must be positive");
  if (m->nelements * m->np >= (int64_t(1) << 31))
    return fail(MG_ERR_ARG, "node count exceeds int32 connectivity");
  if (!m->rx || !m->normals || !m->fscale || !m->vmap_p || !m->weak_diff ||
      !m->lift || !m->face_mask)
    return fail(MG_ERR_ARG, "mesh has a NULL array");
  return MG_OK;
}

mgdg::KArgs make_args(const mg_mesh* m, const mg_gas* g) {
  mgdg::KArgs a{};
  a.rx = m->rx;
  a.normals = m->normals;
  a.fscale = m->fscale;
  a.vmap_p = m->vmap_p;
  a.dw = m->weak_diff;
  a.lift = m->lift;
  a.fmask = m->face_mask;
  a.K = m->nelements;
  a.nbnd = m->nbnd_nodes;
  a.gas = to_dev(g);
  return a;
}

template <int DIM, int P, int NSP, bool MIX, bool VISC>
mg_status launch_rhs(const mgdg::KArgs& a, hipStream_t st) {
  using C = mgdg::RhsCfg<DIM, P, NSP, MIX, VISC>;
  auto kern = mgdg::rhs_kernel<DIM, P, NSP, MIX, VISC>;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(C::LDS));
    if (e != hipSuccess)
      return fail(MG_ERR_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    attr = true;
  }
  const int64_t nblk = (a.K + C::EPB - 1) / C::EPB;
  hipLaunchKernelGGL(kern, dim3(unsigned(nblk)), dim3(mgdg::kBlock), C::LDS, st, a);
  return check_launch("rhs_kernel");
}

template <int DIM, int P, int NSP, bool MIX>
mg_status launch_grad(const mgdg::KArgs& a, hipStream_t st) {
  using C = mgdg::GradCfg<DIM, P, NSP, MIX>;
  auto kern = mgdg::grad_kernel<DIM, P, NSP, MIX>;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(C::LDS));
    if (e != hipSuccess)
      return fail(MG_ERR_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    attr = true;
  }
  const int64_t nblk = (a.K + C::EPB - 1) / C::EPB;
  hipLaunchKernelGGL(kern, dim3(unsigned(nblk)), dim3(mgdg::kBlock), C::LDS, st, a);
  return check_launch("grad_kernel");
}

template <int DIM>
mg_status dispatch_rhs_dim(int order, const mgdg::KArgs& a, int nsp, bool mix,
                           bool visc, hipStream_t st);

#define MG_RHS_CASES(DIM)                                                     \
  template <>                                                                 \
  mg_status dispatch_rhs_dim<DIM>(int order, const mgdg::KArgs& a, int nsp,   \
                                  bool mix, bool visc, hipStream_t st) {      \
    switch (order) {                                                          \
      case 1: return dispatch_rhs_p2<DIM, 1>(a, nsp, mix, visc, st);          \
      case 2: return dispatch_rhs_p2<DIM, 2>(a, nsp, mix, visc, st);          \
      case 3: return dispatch_rhs_p2<DIM, 3>(a, nsp, mix, visc, st);          \
      case 4: return dispatch_rhs_p2<DIM, 4>(a, nsp, mix, visc, st);          \
    }                                                                         \
    return fail(MG_ERR_UNSUPPORTED, "order");                                 \
  }

template <int DIM, int P>
mg_status dispatch_rhs_p2(const mgdg::KArgs& a, int nsp, bool mix, bool visc,
                          hipStream_t st) {
  if (visc) {
    if (nsp == 0 && !mix) return launch_rhs<DIM, P, 0, false, true>(a, st);
    if (nsp == 7 && !mix) return launch_rhs<DIM, P, 7, false, true>(a, st);
    if (nsp == 7 && mix) return launch_rhs<DIM, P, 7, true, true>(a, st);
  } else {
    if (nsp == 0 && !mix) return launch_rhs<DIM, P, 0, false, false>(a, st);
    if (nsp == 7 && !mix) return launch_rhs<DIM, P, 7, false, false>(a, st);
    if (nsp == 7 && mix) return launch_rhs<DIM, P, 7, true, false>(a, st);
  }
  return fail(MG_ERR_UNSUPPORTED,
              "unsupported species count / EOS combination (nspecies 0 or 7)");
}

MG_RHS_CASES(2)
MG_RHS_CASES(3)

template <int DIM, int P>
mg_status dispatch_grad_p(const mgdg::KArgs& a, int nsp, bool mix, hipStream_t st) {
  if (nsp == 0 && !mix) return launch_grad<DIM, P, 0, false>(a, st);
  if (nsp == 7 && !mix) return launch_grad<DIM, P, 7, false>(a, st);
  if (nsp == 7 && mix) return launch_grad<DIM, P, 7, true>(a, st);
  return fail(MG_ERR_UNSUPPORTED,
              "unsupported species count / EOS combination (nspecies 0 or 7)");
}

template <int DIM>
mg_status dispatch_grad_dim(int order, const mgdg::KArgs& a, int nsp, bool mix,
                            hipStream_t st) {
  switch (order) {
    case 1: return dispatch_grad_p<DIM, 1>(a, nsp, mix, st);
    case 2: return dispatch_grad_p<DIM, 2>(a, nsp, mix, st);
    case 3: return dispatch_grad_p<DIM, 3>(a, nsp, mix, st);
    case 4: return dispatch_grad_p<DIM, 4>(a, nsp, mix, st);
  }
  return fail(MG_ERR_UNSUPPORTED, "order");
}

mg_status check_gas(const mg_gas* g, bool need_transport) {
  if (!g) return fail(MG_ERR_ARG, "gas is NULL");
  if (g->eos_kind == 1 && g->nspecies != 7)
    return fail(MG_ERR_ARG, "the mixture EOS needs the 7 uiuc_7sp species");
  if (g->eos_kind != 0 && g->eos_kind != 1) return fail(MG_ERR_ARG, "eos_kind");
  if (g->nspecies < 0 || g->nspecies > 8) return fail(MG_ERR_ARG, "nspecies");
  if (need_transport && g->transport_kind == 0)
    return fail(MG_ERR_ARG, "Navier-Stokes operator expects a viscous gas model");
  return MG_OK;
}

bool g_mech_loaded = false;

mg_status ensure_mech() {
  if (g_mech_loaded) return MG_OK;
  return fail(MG_ERR_ARG, "mixture mechanism not loaded (mg_load_mechanism)");
}

}  // namespace

extern "C" {

const char* mg_last_error(void) { return g_err.c_str(); }

int32_t mg_abi_version(void) { return 1; }

/* Upload the uiuc_7sp mechanism constants (host array of 7*7*2 + 7 + 7 + 7
 * + 2 + 9 doubles, packed as in mgdg::MechDev). */
mg_status mg_load_mechanism(const double* packed, int64_t count) {
  if (count != int64_t(sizeof(mgdg::MechDev) / sizeof(double)))
    return fail(MG_ERR_ARG, "mechanism size mismatch");
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(mgdg::c_mech), packed,
                                   sizeof(mgdg::MechDev), 0, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(MG_ERR_HIP, hipGetErrorString(e));
  g_mech_loaded = true;
  return MG_OK;
}

mg_status mg_euler_rhs(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                       const double* temperature, const double* q_bnd,
                       const double* t_bnd, int32_t flux_kind, double* rhs,
                       void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  const bool mix = gas->eos_kind == 1;
  if (mix && (s = ensure_mech())) return s;
  if (!q || !rhs) return fail(MG_ERR_ARG, "q/rhs NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  if (flux_kind < 0 || flux_kind > 2) return fail(MG_ERR_ARG, "flux_kind");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.qb = q_bnd;
  a.tb = t_bnd;
  a.out = rhs;
  a.flux_kind = flux_kind;
  a.inviscid_on = 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (mesh->dim == 2)
    return dispatch_rhs_dim<2>(mesh->order, a, gas->nspecies, mix, false, st);
  return dispatch_rhs_dim<3>(mesh->order, a, gas->nspecies, mix, false, st);
}

mg_status mg_ns_gradient(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                         const double* temperature, const double* q_bnd,
                         const double* t_bnd, double* grad, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  const bool mix = gas->eos_kind == 1;
  if (!q || !grad) return fail(MG_ERR_ARG, "q/grad NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.qb = q_bnd;
  a.tb = t_bnd;
  a.out = grad;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (mesh->dim == 2) return dispatch_grad_dim<2>(mesh->order, a, gas->nspecies, mix, st);
  return dispatch_grad_dim<3>(mesh->order, a, gas->nspecies, mix, st);
}

mg_status mg_ns_rhs(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                    const double* temperature, const double* grad,
                    const double* q_bnd, const double* t_bnd, int32_t flux_kind,
                    int32_t inviscid_on, double* rhs, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const bool mix = gas->eos_kind == 1;
  if (mix && (s = ensure_mech())) return s;
  if (!q || !grad || !rhs) return fail(MG_ERR_ARG, "q/grad/rhs NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  if (flux_kind < 0 || flux_kind > 2) return fail(MG_ERR_ARG, "flux_kind");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.grad = grad;
  a.qb = q_bnd;
  a.tb = t_bnd;
  a.out = rhs;
  a.flux_kind = flux_kind;
  a.inviscid_on = inviscid_on ? 1 : 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (mesh->dim == 2)
    return dispatch_rhs_dim<2>(mesh->order, a, gas->nspecies, mix, true, st);
  return dispatch_rhs_dim<3>(mesh->order, a, gas->nspecies, mix, true, st);
}

mg_status mg_mixture_temperature(const mg_gas* gas, int32_t dim, int64_t n,
                                 const double* q, const double* t_seed,
                                 int32_t niter, double* t_out, void* stream) {
  mg_status s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 1) return fail(MG_ERR_ARG, "mixture EOS required");
  if ((s = ensure_mech())) return s;
  if (!q || !t_seed || !t_out || n < 0 || niter < 0)
    return fail(MG_ERR_ARG, "bad arguments");
  if (n == 0) return MG_OK;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dim == 2)
    hipLaunchKernelGGL(mgdg::mix_temperature_kernel<2>, dim3(nblk), dim3(mgdg::kBlock), 0,
                       st, n, q, t_seed, niter, t_out);
  else if (dim == 3)
    hipLaunchKernelGGL(mgdg::mix_temperature_kernel<3>, dim3(nblk), dim3(mgdg::kBlock), 0,
                       st, n, q, t_seed, niter, t_out);
  else
    return fail(MG_ERR_ARG, "dim");
  return check_launch("mix_temperature_kernel");
}

mg_status mg_species_source(const mg_gas* gas, int32_t dim, int64_t n,
                            const double* q, const double* temperature,
                            double* src, void* stream) {
  mg_status s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 1) return fail(MG_ERR_ARG, "mixture EOS required");
  if ((s = ensure_mech())) return s;
  if (!q || !temperature || !src || n < 0) return fail(MG_ERR_ARG, "bad arguments");
  if (n == 0) return MG_OK;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dim == 2)
    hipLaunchKernelGGL(mgdg::species_source_kernel<2>, dim3(nblk), dim3(mgdg::kBlock), 0,
                       st, n, q, temperature, src);
  else if (dim == 3)
    hipLaunchKernelGGL(mgdg::species_source_kernel<3>, dim3(nblk), dim3(mgdg::kBlock), 0,
                       st, n, q, temperature, src);
  else
    return fail(MG_ERR_ARG, "dim");
  return check_launch("species_source_kernel");
}

mg_status mg_rk_axpy(int64_t n, const double* y, const double* k, double a,
                     double* out, double* acc, double w, void* stream) {
  if (!y || !k || !out || n < 0) return fail(MG_ERR_ARG, "bad arguments");
  if (n == 0) return MG_OK;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipLaunchKernelGGL(mgdg::rk_axpy_kernel, dim3(nblk), dim3(mgdg::kBlock), 0,
                     static_cast<hipStream_t>(stream), n, y, k, a, out, acc, w);
  return check_launch("rk_axpy_kernel");
}

}  // extern "C"

// }}}
