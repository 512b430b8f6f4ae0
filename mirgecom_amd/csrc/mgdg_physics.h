// Pointwise gas physics evaluated inside the DG kernels (FP64).
//
// Formulas follow the reference:
//   ideal gas EOS           mirgecom/eos.py:302-537
//   uiuc_7sp mixture EOS    mirgecom/eos.py:540-888, thermochemistry.py:87-167
//   transport               mirgecom/transport.py:143-321
//   inviscid flux           mirgecom/inviscid.py:65-102
//   Rusanov / central / HLL mirgecom/inviscid.py:105-287, flux.py:45-177
//   viscous flux            mirgecom/viscous.py:72-293
// The operation order follows the reference expressions where it costs
// nothing, so that the FP64 results agree with the oracle to a few ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgdg {

constexpr int binom(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Gas model parameters passed by value to every kernel.
struct GasDev {
  int eos_kind;        // 0 ideal single gas, 1 uiuc_7sp mixture
  int transport_kind;  // 0 none, 1 constant (SimpleTransport), 2 power law
  double gamma, gas_const;
  double mu, mu_b, kappa;
  double pl_alpha, pl_beta, pl_sigma, pl_n;
  int has_lewis;
  double diff[8];      // species diffusivities, or Lewis numbers if has_lewis
  double ma_alpha, ma_factor, ma_epsilon, ma_singular;  // mixture-averaged
};

// {{{ uiuc_7sp thermochemistry (species order C2H4, O2, CO2, CO, H2O, H2, N2)

constexpr int MIX_NSP = 7;

struct MechDev {
  double nasa_lo[MIX_NSP][7];
  double nasa_hi[MIX_NSP][7];
  double t_mid[MIX_NSP];
  double wt[MIX_NSP];
  double inv_wt[MIX_NSP];
  double gas_constant;   // J/(kmol K)
  double one_atm;        // Pa
  // Arrhenius: k = exp(log_a + b log T - ea_r / T)
  double log_a[3], b[3], ea_r[3];
};

// one copy per translation unit (each .hip file is its own device
// module); mg_load_mechanism / mg_load_transport fill every copy
static __constant__ MechDev c_mech;

// NASA7 polynomials rearranged for Horner evaluation, derived on the host
// from the mechanism's coefficients a0..a6 when it is loaded ([0] low range,
// [1] high range):
//   cp/R = a0 + a1 T + a2 T^2 + a3 T^3 + a4 T^4
//   h/RT = a0 + a1/2 T + a2/3 T^2 + a3/4 T^3 + a4/5 T^4 + a5 / T
//   s/R  = a0 ln T + a1 T + a2/2 T^2 + a3/3 T^3 + a4/4 T^4 + a6
struct PolyDev {
  double cp[2][MIX_NSP][5];
  double h[2][MIX_NSP][6];
  double s[2][MIX_NSP][6];
};

// one copy per translation unit (each .hip file is its own device
// module); mg_load_mechanism / mg_load_transport fill every copy
static __constant__ PolyDev c_poly;

// Lowest and highest NASA7 switch temperature over the species: a wave whose
// lanes all lie above the highest (below the lowest) needs only the high
// (low) range, which the callers below evaluate alone in that case (the
// same values as the per-lane selection).
__device__ __forceinline__ void nasa7_tmid_range(double& lo, double& hi) {
  lo = c_mech.t_mid[0];
  hi = c_mech.t_mid[0];
#pragma unroll
  for (int k = 1; k < MIX_NSP; ++k) {
    lo = fmin(lo, c_mech.t_mid[k]);
    hi = fmax(hi, c_mech.t_mid[k]);
  }
}

// The temperature range is chosen per lane, the species index is
// wave-uniform: both ranges' polynomials are evaluated with scalar-operand
// coefficients and the result is selected per lane (indexing the table by
// the range would turn every coefficient into a per-lane vector load).
__device__ __forceinline__ double horner5(const double (&c)[5], double t) {
  return c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * c[4])));
}

__device__ __forceinline__ void nasa7_cp_h(int k, double t, double inv_t, double& cp_r,
                                           double& h_rt) {
  const bool high = t > c_mech.t_mid[k];
  const double cp_lo = horner5(c_poly.cp[0][k], t);
  const double cp_hi = horner5(c_poly.cp[1][k], t);
  const double* hl = c_poly.h[0][k];
  const double* hh = c_poly.h[1][k];
  const double h_lo = hl[0] + t * (hl[1] + t * (hl[2] + t * (hl[3] + t * hl[4]))) +
                      hl[5] * inv_t;
  const double h_hi = hh[0] + t * (hh[1] + t * (hh[2] + t * (hh[3] + t * hh[4]))) +
                      hh[5] * inv_t;
  cp_r = high ? cp_hi : cp_lo;
  h_rt = high ? h_hi : h_lo;
}

__device__ __forceinline__ double nasa7_s_r(int k, double t, double log_t) {
  const bool high = t > c_mech.t_mid[k];
  const double* sl = c_poly.s[0][k];
  const double* sh = c_poly.s[1][k];
  const double s_lo = sl[0] * log_t + t * (sl[1] + t * (sl[2] + t * (sl[3] + t * sl[4]))) + sl[5];
  const double s_hi = sh[0] * log_t + t * (sh[1] + t * (sh[2] + t * (sh[3] + t * sh[4]))) + sh[5];
  return high ? s_hi : s_lo;
}

// Mixture internal energy and cv (mass basis) at T.
__device__ __forceinline__ void mix_e_cv(double t, const double* y, double& e,
                                         double& cv) {
  double se = 0.0, sc = 0.0;
  const double inv_t = 1.0 / t;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double cp_r, h_rt;
    nasa7_cp_h(k, t, inv_t, cp_r, h_rt);
    se += c_mech.inv_wt[k] * (h_rt - 1.0) * y[k];
    sc += c_mech.inv_wt[k] * (cp_r - 1.0) * y[k];
  }
  e = c_mech.gas_constant * t * se;
  cv = c_mech.gas_constant * sc;
}

// e and cv of one temperature range (RANGE 0 low, 1 high) for all species.
template <int RANGE>
__device__ __forceinline__ void mix_e_cv_range(double t, const double* y, double& e,
                                               double& cv) {
  double se = 0.0, sc = 0.0;
  const double inv_t = 1.0 / t;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    const double* h = c_poly.h[RANGE][k];
    const double cp_r = horner5(c_poly.cp[RANGE][k], t);
    const double h_rt = h[0] + t * (h[1] + t * (h[2] + t * (h[3] + t * h[4]))) +
                        h[5] * inv_t;
    se += c_mech.inv_wt[k] * (h_rt - 1.0) * y[k];
    sc += c_mech.inv_wt[k] * (cp_r - 1.0) * y[k];
  }
  e = c_mech.gas_constant * t * se;
  cv = c_mech.gas_constant * sc;
}

// Fixed-iteration Newton solve for T from mass-specific internal energy
// (thermochemistry.py:126-167, update at :99-102).  For the standalone
// kernel: when the whole wave is on one side of every species' T_mid (the
// common case), only that range's polynomials are evaluated.
__device__ __forceinline__ double mix_temperature(double e_in, double t_seed,
                                                  const double* y, int niter) {
  double tmid_lo, tmid_hi;
  nasa7_tmid_range(tmid_lo, tmid_hi);
  double t = t_seed;
  for (int it = 0; it < niter; ++it) {
    double e, cv;
    if (__all(t > tmid_hi)) {
      mix_e_cv_range<1>(t, y, e, cv);
    } else if (__all(!(t > tmid_lo))) {
      mix_e_cv_range<0>(t, y, e, cv);
    } else {
      mix_e_cv(t, y, e, cv);
    }
    t = t + (e_in - e) / cv;
  }
  return t;
}

// Newton solve to a tolerance: iterate T <- T + (e - e(T))/cv(T) until the
// update is at most tol (checked after each update) or max_iter updates;
// *iters receives the number of updates, or -1 when the last update still
// exceeded tol.  The update is the fixed-iteration solve's (mix_e_cv).
__device__ __forceinline__ double mix_temperature_tol(double e_in, double t_seed,
                                                      const double* y, double tol,
                                                      int max_iter, int* iters) {
  double t = t_seed;
  int it = 0;
  bool done = false;
  while (!done && it < max_iter) {
    double e, cv;
    mix_e_cv(t, y, e, cv);
    const double dt = (e_in - e) / cv;
    t = t + dt;
    ++it;
    done = fabs(dt) <= tol;
  }
  *iters = done ? it : -1;
  return t;
}

// Species mass sources W_k omega_k of the uiuc_7sp mechanism at density rho,
// temperature t and mass fractions y (mirgecom/eos.py:858-875 with the
// concentrations of mirgecom/thermochemistry.py:87-94, zero level 0):
//   R1  C2H4 + O2 => 2 CO + 2 H2      r1 = kf1 [C2H4]^0.5 [O2]^0.65
//   R2  CO + 0.5 O2 <=> CO2           r2 = kf2 ([CO][O2]^0.5 - [CO2] / Kc2)
//   R3  H2 + 0.5 O2 => H2O            r3 = kf3 [H2]^0.75 [O2]^0.5
// with kf = exp(log A + b ln T - Ea/(R T)) (SI) and
// 1/Kc2 = exp(dG/RT + 0.5 ln(p0/RT)) (uiuc_7sp.yaml:146-154).
__device__ __forceinline__ void mix_species_source(double rho, double t, const double* y,
                                                   double* src) {
  double conc[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    const double c = c_mech.inv_wt[k] * rho * y[k];
    conc[k] = c < 0.0 ? 0.0 : c;
  }
  const double logt = log(t);
  const double inv_t = 1.0 / t;
  double kf[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    kf[r] = exp(c_mech.log_a[r] + c_mech.b[r] * logt - c_mech.ea_r[r] * inv_t);
  double g_rt[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double cp_r, h_rt;
    nasa7_cp_h(k, t, inv_t, cp_r, h_rt);
    g_rt[k] = h_rt - nasa7_s_r(k, t, logt);
  }
  const double c0 = log(c_mech.one_atm / (c_mech.gas_constant * t));
  const double log_rkc = g_rt[2] - (g_rt[3] + 0.5 * g_rt[1]) + 0.5 * c0;
  const double sq_o2 = sqrt(conc[1]);
  const double r1 = kf[0] * sqrt(conc[0]) * pow(conc[1], 0.65);
  const double r2 = kf[1] * (conc[3] * sq_o2 - exp(log_rkc) * conc[2]);
  const double r3 = kf[2] * pow(conc[5], 0.75) * sq_o2;
  const double w[MIX_NSP] = {-r1, -r1 - 0.5 * r2 - 0.5 * r3, r2, 2 * r1 - r2, r3,
                             2 * r1 - r3, 0.0};
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) src[k] = c_mech.wt[k] * w[k];
}

// Mixture enthalpy and cp (mass basis) at T, and the specific gas constant.
__device__ __forceinline__ void mix_h_cp(double t, const double* y, double& h,
                                         double& cp) {
  double sh = 0.0, sc = 0.0;
  const double inv_t = 1.0 / t;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double cp_r, h_rt;
    nasa7_cp_h(k, t, inv_t, cp_r, h_rt);
    sh += c_mech.inv_wt[k] * h_rt * y[k];
    sc += c_mech.inv_wt[k] * cp_r * y[k];
  }
  h = c_mech.gas_constant * t * sh;
  cp = c_mech.gas_constant * sc;
}

__device__ __forceinline__ double mix_gas_const(const double* y) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) s += c_mech.inv_wt[k] * y[k];
  return c_mech.gas_constant * s;
}

// Pure-species transport fits (degree-4 polynomials in ln T, lowest power
// first), derived on the host from the mechanism's Lennard-Jones parameters:
//   sqrt(mu_k / sqrt(T)),  lambda_k / sqrt(T),  p D_kj / T^1.5
// and the Wilke factors phi_a = (1 + W_k/W_j)^-1/2 / sqrt(8),
// phi_b = (W_j/W_k)^1/4.
constexpr int MIX_NPAIR = MIX_NSP * (MIX_NSP - 1) / 2;

struct TranDev {
  double visc[MIX_NSP][5];
  double cond[MIX_NSP][5];
  double diff[MIX_NPAIR][5];  // pairs k < j, row-major upper triangle
  double dself[MIX_NSP][5];
  double phi_a[MIX_NSP][MIX_NSP];
  double phi_b[MIX_NSP][MIX_NSP];
};

// one copy per translation unit (each .hip file is its own device
// module); mg_load_mechanism / mg_load_transport fill every copy
static __constant__ TranDev c_tran;

// }}}

template <int DIM, int NSP>
struct Node {
  static constexpr int NFLD = 2 + DIM + NSP;
  static constexpr int NS1 = NSP > 0 ? NSP : 1;
  double q[NFLD];
  double v[DIM];
  double y[NS1];
  double hk[NS1];
  double p, t, c, gamma, speed, inv_rho;
  // transport
  double mu, mu_b, kappa;
  double d[NS1];
};

// MixtureAveragedTransport (mirgecom/transport.py:324-502) with the mixture
// rules of the pyrometheus mechanism class: Wilke's viscosity
// mu = sum_k X_k mu_k / sum_j X_j phi_kj, where sqrt(mu_k/mu_j) is the ratio
// of the fitted square roots; kappa = (sum X_k lambda_k + 1/sum X_k/lambda_k)/2;
// d_k = (1 - Y_k) / (p sum_{j != k} X_j / D_kj(p = 1)), the self diffusion
// D_kk/p where that sum vanishes, the singular diffusivity where
// 1 - Y_k < epsilon; or kappa/(rho Le_k cp) with Lewis numbers.
template <int DIM, int NSP>
__device__ __forceinline__ void mixavg_transport(Node<DIM, NSP>& s, const GasDev& g) {
  const double t = s.t;
  const double lt = log(t);
  const double st = sqrt(t);
  double ysw = 0.0;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) ysw += s.y[k] * c_mech.inv_wt[k];
  const double mmw = 1.0 / ysw;
  double x[MIX_NSP], sv[MIX_NSP], isv[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    x[k] = s.y[k] * mmw * c_mech.inv_wt[k];
    sv[k] = horner5(c_tran.visc[k], lt);
    isv[k] = 1.0 / sv[k];
  }
  double mu = 0.0;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double den = 0.0;
#pragma unroll
    for (int j = 0; j < MIX_NSP; ++j) {
      const double r = 1.0 + sv[k] * isv[j] * c_tran.phi_b[k][j];
      den += x[j] * (c_tran.phi_a[k][j] * (r * r));
    }
    mu += x[k] * (st * sv[k] * sv[k]) / den;
  }
  mu *= g.ma_factor;
  double s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    const double lam = st * horner5(c_tran.cond[k], lt);
    s1 += x[k] * lam;
    s2 += x[k] / lam;
  }
  const double kappa = g.ma_factor * (0.5 * (s1 + 1.0 / s2));
  s.mu = mu;
  s.mu_b = g.ma_alpha * mu;
  s.kappa = kappa;
  if (g.has_lewis) {
    // cp of the mixture from gamma and R_mix = p / (rho T)
    const double cp = s.gamma * s.p * s.inv_rho / (t * (s.gamma - 1.0));
#pragma unroll
    for (int k = 0; k < MIX_NSP; ++k) s.d[k] = kappa * s.inv_rho / (g.diff[k] * cp);
    return;
  }
  const double t15 = t * st;
  double dn[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) dn[k] = 0.0;
#pragma unroll
  for (int k = 0, m = 0; k < MIX_NSP; ++k) {
#pragma unroll
    for (int j = k + 1; j < MIX_NSP; ++j, ++m) {
      const double inv_d = 1.0 / (t15 * horner5(c_tran.diff[m], lt));
      dn[k] += x[j] * inv_d;
      dn[j] += x[k] * inv_d;
    }
  }
  const double inv_p = 1.0 / s.p;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double d = dn[k] > 0.0 ? (1.0 - s.y[k]) * inv_p / dn[k]
                           : t15 * horner5(c_tran.dself[k], lt) * inv_p;
    if (1.0 - s.y[k] < g.ma_epsilon) d = g.ma_singular;
    s.d[k] = g.ma_factor * d;
  }
}

// Transport coefficients of a node: SimpleTransport / PowerLawTransport by
// transport_kind, MixtureAveragedTransport in the MIX == 2 instantiations
// (0 ideal gas, 1 mixture, 2 mixture with mixture-averaged transport; kernels
// that evaluate no transport use 1 for both mixture cases).
template <int DIM, int NSP, int MIX>
__device__ __forceinline__ void node_transport(Node<DIM, NSP>& s, const GasDev& g) {
  if constexpr (MIX == 2) {
    mixavg_transport<DIM, NSP>(s, g);
    return;
  }
  const double irho = s.inv_rho;
  if (g.transport_kind == 1) {
    s.mu = g.mu;
    s.mu_b = g.mu_b;
    s.kappa = g.kappa;
#pragma unroll
    for (int k = 0; k < NSP; ++k) s.d[k] = g.diff[k];
  } else if (g.transport_kind == 2) {
    const double mu = g.pl_beta * pow(s.t, g.pl_n);
    s.mu = mu;
    s.mu_b = g.pl_alpha * mu;
    double cv;
    if constexpr (MIX == 0) {
      cv = g.gas_const / (g.gamma - 1.0);
    } else {
      double e;
      mix_e_cv(s.t, s.y, e, cv);
    }
    s.kappa = g.pl_sigma * mu * cv;
#pragma unroll
    for (int k = 0; k < NSP; ++k)
      s.d[k] = g.has_lewis ? g.pl_sigma * mu * irho / (g.diff[k] * s.gamma)
                           : g.diff[k];
  }
}

// Mixture sums of a node at T: sum Y_k/W_k, sum cp_k/R Y_k/W_k and the
// species enthalpies h_k (mass basis).  RANGE 0 / 1: only the low / high
// NASA7 range (every active lane of the wave lies in it); -1: both ranges
// with a per-lane select.
template <int RANGE, int DIM, int NSP>
__device__ __forceinline__ void mix_node_sums(Node<DIM, NSP>& s, double t, double& rsum,
                                              double& cpsum) {
  const double inv_t = 1.0 / t;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double cp_r, h_rt;
    if constexpr (RANGE < 0) {
      nasa7_cp_h(k, t, inv_t, cp_r, h_rt);
    } else {
      const double* h = c_poly.h[RANGE][k];
      cp_r = horner5(c_poly.cp[RANGE][k], t);
      h_rt = h[0] + t * (h[1] + t * (h[2] + t * (h[3] + t * h[4]))) + h[5] * inv_t;
    }
    rsum += s.y[k] * c_mech.inv_wt[k];
    cpsum += c_mech.inv_wt[k] * cp_r * s.y[k];
    s.hk[k] = c_mech.gas_constant * c_mech.inv_wt[k] * t * h_rt;
  }
}

// Fill the dependent variables of a node from q (and T for mixtures); the
// transport coefficients only if TRANSPORT.
template <int DIM, int NSP, int MIX, bool TRANSPORT = true>
__device__ __forceinline__ void node_state(Node<DIM, NSP>& s, const GasDev& g,
                                           double t_mix) {
  const double rho = s.q[0];
  const double irho = 1.0 / rho;
  s.inv_rho = irho;
  double mm = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) mm += s.q[2 + d] * s.q[2 + d];
#pragma unroll
  for (int d = 0; d < DIM; ++d) s.v[d] = s.q[2 + d] * irho;
#pragma unroll
  for (int k = 0; k < NSP; ++k) s.y[k] = s.q[2 + DIM + k] * irho;
  double vv = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) vv += s.v[d] * s.v[d];
  s.speed = sqrt(vv);
  const double e_int = s.q[1] - 0.5 * mm * irho;
  if constexpr (MIX == 0) {
    s.gamma = g.gamma;
    s.p = e_int * (g.gamma - 1.0);
    s.t = ((g.gamma - 1.0) / g.gas_const) * e_int * irho;
    s.c = sqrt(g.gamma * irho * s.p);
  } else {
    static_assert(NSP == MIX_NSP, "mixture path is the 7-species mechanism");
    s.t = t_mix;
    double tmid_lo, tmid_hi;
    nasa7_tmid_range(tmid_lo, tmid_hi);
    double rsum = 0.0, cpsum = 0.0;
    if (__all(t_mix > tmid_hi))
      mix_node_sums<1>(s, t_mix, rsum, cpsum);
    else if (__all(!(t_mix > tmid_lo)))
      mix_node_sums<0>(s, t_mix, rsum, cpsum);
    else
      mix_node_sums<-1>(s, t_mix, rsum, cpsum);
    const double rmix = c_mech.gas_constant * rsum;
    const double cp = c_mech.gas_constant * cpsum;
    s.p = rho * rmix * t_mix;
    s.gamma = cp / (cp - rmix);
    s.c = sqrt(s.gamma * s.p * irho);
  }
  if constexpr (TRANSPORT) node_transport<DIM, NSP, MIX>(s, g);
}

// Inviscid flux F[f][d] (inviscid.py:65-102).
template <int DIM, int NSP>
__device__ __forceinline__ void inviscid_flux(const Node<DIM, NSP>& s,
                                              double (*F)[DIM]) {
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    F[0][d] = s.q[2 + d];
    F[1][d] = s.v[d] * (s.q[1] + s.p);
#pragma unroll
    for (int i = 0; i < DIM; ++i)
      F[2 + i][d] = s.q[0] * (s.v[i] * s.v[d]) + (i == d ? s.p : 0.0);
#pragma unroll
    for (int k = 0; k < NSP; ++k) F[2 + DIM + k][d] = s.v[d] * s.q[2 + DIM + k];
  }
}

// Normal inviscid flux F.n
template <int DIM, int NSP>
__device__ __forceinline__ void inviscid_flux_n(const Node<DIM, NSP>& s,
                                                const double* n, double* fn) {
  double vn = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) vn += s.v[d] * n[d];
  double mn = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) mn += s.q[2 + d] * n[d];
  fn[0] = mn;
  fn[1] = vn * (s.q[1] + s.p);
#pragma unroll
  for (int i = 0; i < DIM; ++i) fn[2 + i] = s.q[0] * s.v[i] * vn + s.p * n[i];
#pragma unroll
  for (int k = 0; k < NSP; ++k) fn[2 + DIM + k] = vn * s.q[2 + DIM + k];
}

// {{{ entropy-stable DG pieces (ideal gas)

// Entropy variables (mirgecom/gas_model.py:1026-1060, Chandrashekar 2013
// eq. 4.5) and their inverse (gas_model.py:1063-1100, Hughes, Franca &
// Mallet 1986).
template <int DIM, int NSP>
__device__ __forceinline__ void cons_to_entropy(const double* q, double gamma,
                                                double* v) {
  const double rho = q[0];
  const double irho = 1.0 / rho;
  double u2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) u2 += (q[2 + d] * irho) * (q[2 + d] * irho);
  const double p = (gamma - 1.0) * (q[1] - 0.5 * rho * u2);
  const double s = log(p) - gamma * log(rho);
  const double rho_p = rho / p;
  v[0] = (gamma - s) / (gamma - 1.0) - 0.5 * rho_p * u2;
  v[1] = -rho_p;
#pragma unroll
  for (int d = 0; d < DIM; ++d) v[2 + d] = rho_p * (q[2 + d] * irho);
#pragma unroll
  for (int k = 0; k < NSP; ++k) v[2 + DIM + k] = q[2 + DIM + k] * irho / (gamma - 1.0);
}

template <int DIM, int NSP>
__device__ __forceinline__ void entropy_to_cons(const double* ev, double gamma,
                                                double* q) {
  const double gm1 = gamma - 1.0;
  const double v1 = ev[0] * gm1, v5 = ev[1] * gm1;
  double vv = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) vv += (ev[2 + d] * gm1) * (ev[2 + d] * gm1);
  const double s = gamma - v1 + vv / (2.0 * v5);
  const double iota = pow(gm1 / pow(-v5, gamma), 1.0 / gm1);
  const double rho_iota = iota * exp(-s / gm1);
  const double mass = -rho_iota * v5;
  q[0] = mass;
  q[1] = rho_iota * (1.0 - vv / (2.0 * v5));
#pragma unroll
  for (int d = 0; d < DIM; ++d) q[2 + d] = rho_iota * (ev[2 + d] * gm1);
#pragma unroll
  for (int k = 0; k < NSP; ++k) q[2 + DIM + k] = mass * (ev[2 + DIM + k] * gm1);
}

// Per-point quantities of the entropy-conserving flux, with the logarithms
// of rho and beta = rho/(2p) taken once per point.
template <int NSP>
struct EcSpecies {
  double yh[NSP];  // Y / 2
};
// Without species one padding double makes a 3-D point 9 doubles: an odd
// stride in 8-byte words, so that lanes reading consecutive points from LDS
// spread over the banks (8 doubles put 64 lanes on 4 bank groups).
template <>
struct EcSpecies<0> {
  double pad_;
};

// What the two-point flux reads of a point (kept in LDS per hybridized point).
template <int DIM, int NSP>
struct EcPair : EcSpecies<NSP> {
  double rho, lrho, beta, lbeta;
  double kh;       // |u|^2 / 4
  double uh[DIM];  // u / 2
};

// ... with the pressure and the wavespeed |u| + c the face flux needs.
template <int DIM, int NSP>
struct EcPrim : EcPair<DIM, NSP> {
  double p, c;
};

template <int DIM, int NSP>
__device__ __forceinline__ void ec_prim(const double* q, double gamma,
                                        EcPrim<DIM, NSP>& a) {
  a.rho = q[0];
  const double irho = 1.0 / q[0];
  double u2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double u = q[2 + d] * irho;
    a.uh[d] = 0.5 * u;
    u2 += u * u;
  }
  a.kh = 0.25 * u2;
  a.p = (gamma - 1.0) * (q[1] - 0.5 * q[0] * u2);
  a.beta = 0.5 * q[0] / a.p;
  a.lrho = log(a.rho);
  a.lbeta = log(a.beta);
  a.c = sqrt(u2) + sqrt(gamma * a.p * irho);   // wavespeed |u| + c
  if constexpr (NSP > 0)
    for (int k = 0; k < NSP; ++k) a.yh[k] = 0.5 * q[2 + DIM + k] * irho;
}

// 1/x from the hardware reciprocal estimate and two Newton steps (a few
// ulp; the two-point flux takes three reciprocals per pair of points).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Logarithmic mean (mirgecom/inviscid.py:470-476) of x, y from the values
// and their logarithms, as the quotient num / den: (y - x) / (ln y - ln x),
// or, when f2 = ((x - y) / (x + y))^2 < 1e-4, the series
// (x + y) / (2 + 2/3 f2 + 2/5 f2^2 + 2/7 f2^3), written with u = (x - y)^2,
// w = (x + y)^2 as (x + y) w^3 / (2 w^3 + 2/3 u w^2 + 2/5 u^2 w + 2/7 u^3).
// Both forms are selected before any division, so a wave whose lanes take
// different branches pays for one reciprocal, and the caller takes either
// the mean or its inverse from the same parts.
__device__ __forceinline__ void ln_mean_parts(double x, double lx, double y, double ly,
                                              double& num, double& den) {
  const double sm = x + y, df = x - y;
  const double u = df * df, w = sm * sm;
  const bool small = u < 1e-4 * w;
  const double ser = fma(fma(fma(2.0, w, (2.0 / 3.0) * u), w, 0.4 * u * u), w,
                         (2.0 / 7.0) * u * u * u);
  num = small ? sm * (w * w * w) : y - x;
  den = small ? ser : ly - lx;
}

// sum_d s_d F_S,d(a, b) of Chandrashekar's entropy-conserving, kinetic-
// energy-preserving flux (mirgecom/inviscid.py:451-518), added to acc (SET:
// stored into acc).  The energy flux needs 1 / ((gamma - 1) beta_mean): it
// is formed as den / num of beta's logarithmic mean; p_mean =
// rho_avg / (2 beta_avg) = (rho_a + rho_b) / (2 (beta_a + beta_b)).
template <int DIM, int NSP, bool SET = false>
__device__ __forceinline__ void ec_flux_dot(const EcPair<DIM, NSP>& a,
                                            const EcPair<DIM, NSP>& b,
                                            const double* s, double gamma,
                                            double* acc) {
  double rn, rd, bn, bd;
  ln_mean_parts(a.rho, a.lrho, b.rho, b.lrho, rn, rd);
  ln_mean_parts(a.beta, a.lbeta, b.beta, b.lbeta, bn, bd);
  const double rho_mean = rn * rcp_nr(rd);
  const double inv_beta_mean = bd * rcp_nr(bn);
  const double half_inv_gm1 = 0.5 / (gamma - 1.0);
  const double p_mean = 0.5 * (a.rho + b.rho) * rcp_nr(a.beta + b.beta);
  double un = 0.0, ua2 = 0.0, ua[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    ua[d] = a.uh[d] + b.uh[d];
    un += s[d] * ua[d];
    ua2 += ua[d] * ua[d];
  }
  const double mflux = rho_mean * un;
  double v[2 + DIM + NSP];
  v[0] = mflux;
  v[1] = mflux * (inv_beta_mean * half_inv_gm1 - (a.kh + b.kh) + ua2) + p_mean * un;
#pragma unroll
  for (int i = 0; i < DIM; ++i) v[2 + i] = mflux * ua[i] + s[i] * p_mean;
  if constexpr (NSP > 0)
    for (int k = 0; k < NSP; ++k) v[2 + DIM + k] = mflux * (a.yh[k] + b.yh[k]);
#pragma unroll
  for (int f = 0; f < 2 + DIM + NSP; ++f) {
    if constexpr (SET)
      acc[f] = v[f];
    else
      acc[f] += v[f];
  }
}

// {{{ Renac's two-point flux (mixtures)

// Mixture c_v (mass basis) at T: sum_k Y_k R / W_k (cp_k / R - 1); one
// NASA7 range when every active lane of the wave lies in it.
template <int RANGE>
__device__ __forceinline__ double mix_cv_range(double t, const double* y) {
  double sc = 0.0;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) {
    double cp;
    if constexpr (RANGE < 0) {
      const bool high = t > c_mech.t_mid[k];
      const double cp_lo = horner5(c_poly.cp[0][k], t);
      const double cp_hi = horner5(c_poly.cp[1][k], t);
      cp = high ? cp_hi : cp_lo;
    } else {
      cp = horner5(c_poly.cp[RANGE][k], t);
    }
    sc += c_mech.inv_wt[k] * (cp - 1.0) * y[k];
  }
  return c_mech.gas_constant * sc;
}

__device__ __forceinline__ double mix_cv(double t, const double* y) {
  double tmid_lo, tmid_hi;
  nasa7_tmid_range(tmid_lo, tmid_hi);
  if (__all(t > tmid_hi)) return mix_cv_range<1>(t, y);
  if (__all(!(t > tmid_lo))) return mix_cv_range<0>(t, y);
  return mix_cv_range<-1>(t, y);
}

// Entropy variables (mirgecom/gas_model.py:1026-1060) of a state with its
// own pressure p and ratio gamma (for a mixture p = rho R T, not
// (gamma - 1) rho e).
template <int DIM, int NSP>
__device__ __forceinline__ void cons_to_entropy_p(const double* q, double p,
                                                  double gamma, double* v) {
  const double rho = q[0];
  const double irho = 1.0 / rho;
  double u2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) u2 += (q[2 + d] * irho) * (q[2 + d] * irho);
  const double s = log(p) - gamma * log(rho);
  const double rho_p = rho / p;
  v[0] = (gamma - s) / (gamma - 1.0) - 0.5 * rho_p * u2;
  v[1] = -rho_p;
#pragma unroll
  for (int d = 0; d < DIM; ++d) v[2 + d] = rho_p * (q[2 + d] * irho);
#pragma unroll
  for (int k = 0; k < NSP; ++k) v[2 + DIM + k] = q[2 + DIM + k] * irho / (gamma - 1.0);
}

// What Renac's flux reads of a point: rho, theta = 1/T (with their
// logarithms for the logarithmic means), p/T, T, u/2, Y/2.
template <int DIM>
struct RnPair {
  double rho, lrho, th, lth, pth, t;
  double uh[DIM];
  double yh[MIX_NSP];
};

// ... with the wavespeed |u| + c the face flux needs.
template <int DIM>
struct RnPrim : RnPair<DIM> {
  double c;
};

template <int DIM>
__device__ __forceinline__ void rn_prim(const Node<DIM, MIX_NSP>& s, RnPrim<DIM>& a) {
  a.rho = s.q[0];
  a.lrho = log(s.q[0]);
  a.t = s.t;
  a.th = 1.0 / s.t;
  a.lth = -log(s.t);
  a.pth = s.p * a.th;
#pragma unroll
  for (int d = 0; d < DIM; ++d) a.uh[d] = 0.5 * s.v[d];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) a.yh[k] = 0.5 * s.y[k];
  a.c = s.speed + s.c;
}

// sum_d s_d F_S,d(a, b) of Renac's entropy-conserving, kinetic-energy-
// preserving flux (mirgecom/inviscid.py:521-598; Renac 2021, eq. 24), added
// to acc: rho_mean = ln-mean(rho), T_mean = 1 / ln-mean(1/T),
// p~ = (p_a/T_a + p_b/T_b)/2 (T_a + T_b)/2, u = (u_a + u_b)/2,
// Y = (Y_a + Y_b)/2; mass rho_mean u.s, momentum rho_mean u (u.s) + s p~,
// energy (c_v(T_mean, Y) T_mean + u_a.u_b / 2) rho_mean u.s + p~ u.s,
// species rho_mean Y u.s.  (The reference's p_mean and gamma_avg enter only
// its c_v evaluation's state, whose c_v depends on T and Y alone.)
template <int DIM>
__device__ __forceinline__ void rn_flux_dot(const RnPair<DIM>& a, const RnPair<DIM>& b,
                                            const double* s, double* acc) {
  double rn, rd, tn, td;
  ln_mean_parts(a.rho, a.lrho, b.rho, b.lrho, rn, rd);
  ln_mean_parts(a.th, a.lth, b.th, b.lth, tn, td);
  const double rho_mean = rn * rcp_nr(rd);
  const double t_mean = td * rcp_nr(tn);
  const double pec = 0.25 * (a.pth + b.pth) * (a.t + b.t);
  double un = 0.0, kin = 0.0, ua[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    ua[d] = a.uh[d] + b.uh[d];
    un += s[d] * ua[d];
    kin += a.uh[d] * b.uh[d];
  }
  double y[MIX_NSP];
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) y[k] = a.yh[k] + b.yh[k];
  const double cvm = mix_cv(t_mean, y);
  const double mflux = rho_mean * un;
  acc[0] += mflux;
  acc[1] += (cvm * t_mean + 2.0 * kin) * mflux + pec * un;
#pragma unroll
  for (int i = 0; i < DIM; ++i) acc[2 + i] += mflux * ua[i] + s[i] * pec;
#pragma unroll
  for (int k = 0; k < MIX_NSP; ++k) acc[2 + DIM + k] += mflux * y[k];
}

// }}}

// }}}

// Numerical inviscid flux: 0 Rusanov, 1 HLL, 2 central.
template <int DIM, int NSP>
__device__ __forceinline__ void inviscid_numflux(const Node<DIM, NSP>& a,
                                                 const Node<DIM, NSP>& b,
                                                 const double* n, int kind,
                                                 double* out) {
  constexpr int NF = 2 + DIM + NSP;
  double fa[NF], fb[NF];
  inviscid_flux_n<DIM, NSP>(a, n, fa);
  inviscid_flux_n<DIM, NSP>(b, n, fb);
  if (kind == 1) {
    double u_a = 0.0, u_b = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      u_a += a.v[d] * n[d];
      u_b += b.v[d] * n[d];
    }
    const double p_star = 0.5 * (a.p + b.p) +
                          (1. / 8.) * (u_a - u_b) * (a.q[0] + b.q[0]) * (a.c + b.c);
    const double qa = 1 + (a.gamma + 1) / (2 * a.gamma) * (p_star / a.p - 1);
    const double qb = 1 + (b.gamma + 1) / (2 * b.gamma) * (p_star / b.p - 1);
    const double qa_rt = sqrt(p_star > a.p ? qa : 1.0);
    const double qb_rt = sqrt(p_star > b.p ? qb : 1.0);
    const double s_m = u_a - a.c * qa_rt;
    const double s_p = u_b + b.c * qb_rt;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double fs = (s_p * fa[f] - s_m * fb[f] + s_p * s_m * (b.q[f] - a.q[f])) /
                  (s_p - s_m);
      if (s_m >= 0.0) fs = fa[f];
      if (s_p <= 0.0) fs = fb[f];
      out[f] = fs;
    }
  } else {
    const double lam = kind == 0 ? fmax(a.speed + a.c, b.speed + b.c) : 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) out[f] = (fa[f] + fb[f] + lam * (a.q[f] - b.q[f])) / 2;
  }
}

// Viscous flux F_V[f][d] (viscous.py:232-293). gq[f][d] = grad of cv field
// f, gt[d] = grad T.
template <int DIM, int NSP, int MIX>
__device__ __forceinline__ void viscous_flux(const Node<DIM, NSP>& s,
                                             const double (*gq)[DIM],
                                             const double* gt, double (*FV)[DIM]) {
  const double rho = s.q[0];
  const double irho = s.inv_rho;
  double gv[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) gv[i][j] = (gq[2 + i][j] - s.v[i] * gq[0][j]) * irho;
  double divv = 0.0;
#pragma unroll
  for (int i = 0; i < DIM; ++i) divv += gv[i][i];
  const double lam = (s.mu_b - 2 * s.mu / 3) * divv;
  double tau[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j)
      tau[i][j] = s.mu * (gv[i][j] + gv[j][i]) + (i == j ? lam : 0.0);
  double jd[Node<DIM, NSP>::NS1][DIM];
  if constexpr (NSP > 0) {
    double gy[NSP][DIM];
#pragma unroll
    for (int k = 0; k < NSP; ++k)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        gy[k][d] = (gq[2 + DIM + k][d] - s.y[k] * gq[0][d]) * irho;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double sum = 0.0;
      if constexpr (MIX != 0) {
#pragma unroll
        for (int k = 0; k < NSP; ++k) sum += s.d[k] * gy[k][d];
      }
#pragma unroll
      for (int k = 0; k < NSP; ++k)
        jd[k][d] = -rho * (s.d[k] * gy[k][d] - s.y[k] * sum);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    FV[0][d] = 0.0;
    double tv = 0.0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) tv += tau[d][j] * s.v[j];
    double qd = 0.0;
    if constexpr (MIX != 0) {
#pragma unroll
      for (int k = 0; k < NSP; ++k) qd += s.hk[k] * jd[k][d];
    }
    FV[1][d] = tv - qd + s.kappa * gt[d];
#pragma unroll
    for (int i = 0; i < DIM; ++i) FV[2 + i][d] = tau[i][d];
#pragma unroll
    for (int k = 0; k < NSP; ++k) FV[2 + DIM + k][d] = -jd[k][d];
  }
}

// Normal viscous flux F_V . n without materialising F_V[f][d].
template <int DIM, int NSP, int MIX>
__device__ __forceinline__ void viscous_flux_n(const Node<DIM, NSP>& s,
                                               const double (*gq)[DIM],
                                               const double* gt, const double* n,
                                               double* out) {
  double fv[2 + DIM + NSP][DIM];
  viscous_flux<DIM, NSP, MIX>(s, gq, gt, fv);
#pragma unroll
  for (int f = 0; f < 2 + DIM + NSP; ++f) {
    double c = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) c += fv[f][d] * n[d];
    out[f] = c;
  }
}

}  // namespace mgdg
