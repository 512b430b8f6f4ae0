// Host-side helpers of the C ABI (included by every .hip file of libmgdg).
#pragma once
#include <string>

#include "mgdg_kernels.h"

// {{{ host side: shared state and kernel dispatch
//
// Every .hip file of the library includes this header.  The error text, the
// mechanism/transport flags and the list of per-file constant uploads are
// shared (defined once, in mgdg.hip); everything else is per file.

namespace mgdg_shared {
std::string& err();
bool& mech_loaded();
bool& tran_loaded();
// Each .hip file is its own device module with its own copies of c_mech,
// c_poly and c_tran: the loaders call every registered upload function.
using UploadFn = hipError_t (*)(const mgdg::MechDev*, const mgdg::PolyDev*,
                                const mgdg::TranDev*);
int register_upload(UploadFn fn);
hipError_t upload_all(const mgdg::MechDev* m, const mgdg::PolyDev* p,
                      const mgdg::TranDev* t);

}  // namespace mgdg_shared

namespace {

// this file's constant copies (NULL pointers: leave that table unchanged)
hipError_t upload_this_module(const mgdg::MechDev* m, const mgdg::PolyDev* p,
                              const mgdg::TranDev* t) {
  hipError_t e = hipSuccess;
  if (m && (e = hipMemcpyToSymbol(HIP_SYMBOL(mgdg::c_mech), m, sizeof(*m), 0,
                                  hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  if (p && (e = hipMemcpyToSymbol(HIP_SYMBOL(mgdg::c_poly), p, sizeof(*p), 0,
                                  hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  if (t && (e = hipMemcpyToSymbol(HIP_SYMBOL(mgdg::c_tran), t, sizeof(*t), 0,
                                  hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  return hipSuccess;
}
[[maybe_unused]] const int g_upload_registered =
    mgdg_shared::register_upload(upload_this_module);

mg_status fail(mg_status code, const std::string& msg) {
  mgdg_shared::err() = msg;
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
  d.ma_alpha = g->ma_alpha;
  d.ma_factor = g->ma_factor;
  d.ma_epsilon = g->ma_epsilon;
  d.ma_singular = g->ma_singular_diffusivity;
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
  if (!m->workspace)
    return fail(MG_ERR_ARG, "mesh workspace not prepared (mg_mesh_prepare)");
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
  a.ops = m->workspace;
  a.border = m->block_order;
  a.K = m->nelements;
  a.nbnd = m->nbnd_nodes;
  a.nnodes = m->nelements * m->np;
  a.gas = to_dev(g);
  return a;
}

template <typename Kern>
mg_status set_lds(Kern kern, size_t lds) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(lds));
  if (e != hipSuccess)
    return fail(MG_ERR_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  return MG_OK;
}

template <int DIM, int P, int NSP, int MIX, int MODE, typename R, bool HLL>
mg_status launch_rhs_kernel(const mgdg::KArgs& a, hipStream_t st) {
  using C = mgdg::RhsCfg<DIM, P, NSP, MIX, MODE, R>;
  auto kern = mgdg::rhs_kernel<DIM, P, NSP, MIX, MODE, R, HLL>;
  static bool attr = false;
  if (!attr) {
    mg_status s = set_lds(kern, C::LDS);
    if (s) return s;
    attr = true;
  }
  const int64_t nblk = (a.K + C::EPB - 1) / C::EPB;
  hipLaunchKernelGGL(kern, dim3(unsigned(nblk)), dim3(mgdg::kBlock), C::LDS, st, a,
                     a.ops, a.out);
  return check_launch("rhs_kernel");
}

template <int DIM, int P, int NSP, int MIX, int MODE, typename R>
mg_status launch_rhs(const mgdg::KArgs& a, hipStream_t st) {
  const bool inv = MODE == mgdg::kEuler || a.inviscid_on;
  if (inv && a.flux_kind == MG_FLUX_HLL)
    return launch_rhs_kernel<DIM, P, NSP, MIX, MODE, R, true>(a, st);
  return launch_rhs_kernel<DIM, P, NSP, MIX, MODE, R, false>(a, st);
}

template <int DIM, int P, int NSP, int MIX, bool FLUX, typename R>
mg_status launch_grad(const mgdg::KArgs& a, hipStream_t st) {
  using C = mgdg::GradCfg<DIM, P, NSP, MIX, R, FLUX>;
  auto kern = mgdg::grad_kernel<DIM, P, NSP, MIX, FLUX, R>;
  static bool attr = false;
  if (!attr) {
    mg_status s = set_lds(kern, C::LDS);
    if (s) return s;
    attr = true;
  }
  const int64_t nblk = (a.K + C::EPB - 1) / C::EPB;
  hipLaunchKernelGGL(kern, dim3(unsigned(nblk)), dim3(mgdg::kBlock), C::LDS, st, a,
                     a.ops, a.out);
  return check_launch("grad_kernel");
}

// compiled (species count, EOS) combinations: (0, ideal), (7, ideal), (7, mixture)
// mix: 0 ideal gas, 1 mixture, 2 mixture with mixture-averaged transport
// (a separate instantiation only where the kernel evaluates transport).
template <int DIM, int P, int MODE, typename R>
mg_status rhs_species(const mgdg::KArgs& a, int nsp, int mix, hipStream_t st) {
  if (nsp == 0 && !mix) return launch_rhs<DIM, P, 0, 0, MODE, R>(a, st);
  if (nsp == 7 && !mix) return launch_rhs<DIM, P, 7, 0, MODE, R>(a, st);
  if constexpr (MODE == mgdg::kGrad) {
    if (nsp == 7 && mix == 2) return launch_rhs<DIM, P, 7, 2, MODE, R>(a, st);
  }
  if (nsp == 7 && mix) return launch_rhs<DIM, P, 7, 1, MODE, R>(a, st);
  return fail(MG_ERR_UNSUPPORTED,
              "unsupported species count / EOS combination (nspecies 0 or 7)");
}

template <int DIM, int P, bool FLUX, typename R>
mg_status grad_species(const mgdg::KArgs& a, int nsp, int mix, hipStream_t st) {
  if (nsp == 0 && !mix) return launch_grad<DIM, P, 0, 0, FLUX, R>(a, st);
  if (nsp == 7 && !mix) return launch_grad<DIM, P, 7, 0, FLUX, R>(a, st);
  if constexpr (FLUX) {
    if (nsp == 7 && mix == 2) return launch_grad<DIM, P, 7, 2, FLUX, R>(a, st);
  }
  if (nsp == 7 && mix) return launch_grad<DIM, P, 7, 1, FLUX, R>(a, st);
  return fail(MG_ERR_UNSUPPORTED,
              "unsupported species count / EOS combination (nspecies 0 or 7)");
}

// Nodal net fluxes of the uiuc_7sp mixture from stored gradients (a.grad).
template <int DIM, int P>
mg_status node_flux_species(const mgdg::KArgs& a, int nsp, int mix, hipStream_t st) {
  if (nsp != 7 || !mix)
    return fail(MG_ERR_UNSUPPORTED, "node-parallel net flux is the mixture path");
  const unsigned nblk = unsigned((a.nnodes + mgdg::kBlock - 1) / mgdg::kBlock);
  // FP32 form when the FP32 net-flux output is set (mixed precision)
  const bool f32 = a.fnet32_out != nullptr;
  if (f32 && !a.grad32) return fail(MG_ERR_ARG, "FP32 net flux needs FP32 gradients");
  if (mix == 2 && f32)
    hipLaunchKernelGGL((mgdg::node_flux_kernel<DIM, 7, 2, true>), dim3(nblk),
                       dim3(mgdg::kBlock), 0, st, a);
  else if (mix == 2)
    hipLaunchKernelGGL((mgdg::node_flux_kernel<DIM, 7, 2>), dim3(nblk),
                       dim3(mgdg::kBlock), 0, st, a);
  else if (f32)
    hipLaunchKernelGGL((mgdg::node_flux_kernel<DIM, 7, 1, true>), dim3(nblk),
                       dim3(mgdg::kBlock), 0, st, a);
  else
    hipLaunchKernelGGL((mgdg::node_flux_kernel<DIM, 7, 1>), dim3(nblk),
                       dim3(mgdg::kBlock), 0, st, a);
  mg_status s = check_launch("node_flux_kernel");
  if (s || a.nbnd == 0) return s;
  const int64_t nt = a.K * mgdg::Elem<DIM, P>::NFN;
  const unsigned nb = unsigned((nt + mgdg::kBlock - 1) / mgdg::kBlock);
  if (mix == 2 && f32)
    hipLaunchKernelGGL((mgdg::bnd_node_flux_kernel<DIM, P, 7, 2, true>), dim3(nb),
                       dim3(mgdg::kBlock), 0, st, a);
  else if (mix == 2)
    hipLaunchKernelGGL((mgdg::bnd_node_flux_kernel<DIM, P, 7, 2>), dim3(nb),
                       dim3(mgdg::kBlock), 0, st, a);
  else if (f32)
    hipLaunchKernelGGL((mgdg::bnd_node_flux_kernel<DIM, P, 7, 1, true>), dim3(nb),
                       dim3(mgdg::kBlock), 0, st, a);
  else
    hipLaunchKernelGGL((mgdg::bnd_node_flux_kernel<DIM, P, 7, 1>), dim3(nb),
                       dim3(mgdg::kBlock), 0, st, a);
  return check_launch("bnd_node_flux_kernel");
}

template <int W, int DIM, int P>
mg_status dispatch_p(const mgdg::KArgs& a, int nsp, int mix, hipStream_t st) {
  if constexpr (W == 0) return rhs_species<DIM, P, mgdg::kEuler, double>(a, nsp, mix, st);
  if constexpr (W == 1) return rhs_species<DIM, P, mgdg::kGrad, double>(a, nsp, mix, st);
  if constexpr (W == 2) return rhs_species<DIM, P, mgdg::kFlux, double>(a, nsp, mix, st);
  if constexpr (W == 3) return grad_species<DIM, P, false, double>(a, nsp, mix, st);
  if constexpr (W == 4) return grad_species<DIM, P, true, double>(a, nsp, mix, st);
  if constexpr (W == 5) {
    // mixtures: FP32 gradients for the node-parallel net flux when an FP32
    // gradient output is set; otherwise the fused gradient + net flux kernel
    if (mix && a.grad32_out) {
      if (nsp != 7) return fail(MG_ERR_UNSUPPORTED, "mixture path is the 7-species mechanism");
      return launch_grad<DIM, P, 7, 1, false, float>(a, st);
    }
    return grad_species<DIM, P, true, float>(a, nsp, mix, st);
  }
  if constexpr (W == 6) return rhs_species<DIM, P, mgdg::kFlux, float>(a, nsp, mix, st);
  if constexpr (W == 7) return node_flux_species<DIM, P>(a, nsp, mix, st);
  if constexpr (W == 8) return rhs_species<DIM, P, mgdg::kEuler, float>(a, nsp, mix, st);
  return fail(MG_ERR_ARG, "kernel selector");
}

}  // namespace

// The instantiations of selector W for every (dim, order); each is compiled
// in a .hip file of its own (MG_DEFINE_DISPATCH) so that the library builds
// in parallel.
#define MG_DISPATCH_DECL(W)                                                        \
  mg_status mgdg_dispatch_##W(const mg_mesh* m, const mgdg::KArgs& a, int nsp,     \
                              int mix, hipStream_t st);
MG_DISPATCH_DECL(0) MG_DISPATCH_DECL(1) MG_DISPATCH_DECL(2) MG_DISPATCH_DECL(3)
MG_DISPATCH_DECL(4) MG_DISPATCH_DECL(5) MG_DISPATCH_DECL(6) MG_DISPATCH_DECL(7)
MG_DISPATCH_DECL(8)
#undef MG_DISPATCH_DECL
#define MG_DEFINE_DISPATCH(W)                                                      \
  mg_status mgdg_dispatch_##W(const mg_mesh* m, const mgdg::KArgs& a, int nsp,     \
                              int mix, hipStream_t st) {                           \
    if (m->dim == 2 && m->order == 1) return dispatch_p<W, 2, 1>(a, nsp, mix, st); \
    if (m->dim == 2 && m->order == 2) return dispatch_p<W, 2, 2>(a, nsp, mix, st); \
    if (m->dim == 2 && m->order == 3) return dispatch_p<W, 2, 3>(a, nsp, mix, st); \
    if (m->dim == 2 && m->order == 4) return dispatch_p<W, 2, 4>(a, nsp, mix, st); \
    if (m->dim == 3 && m->order == 1) return dispatch_p<W, 3, 1>(a, nsp, mix, st); \
    if (m->dim == 3 && m->order == 2) return dispatch_p<W, 3, 2>(a, nsp, mix, st); \
    if (m->dim == 3 && m->order == 3) return dispatch_p<W, 3, 3>(a, nsp, mix, st); \
    if (m->dim == 3 && m->order == 4) return dispatch_p<W, 3, 4>(a, nsp, mix, st); \
    return fail(MG_ERR_UNSUPPORTED, "dim/order");                                  \
  }

namespace {

// what: 0 Euler RHS, 1 NS RHS from gradients, 2 NS RHS from nodal fluxes,
// 3 gradients, 4 gradients + nodal fluxes, 5/6 the mixed-precision 4/2,
// 7 nodal fluxes from stored gradients (mixture), 8 the mixed-precision 0
inline mg_status dispatch(const mg_mesh* m, const mgdg::KArgs& a, int nsp, int mix,
                          int what, hipStream_t st) {
  switch (what) {
    case 0: return mgdg_dispatch_0(m, a, nsp, mix, st);
    case 1: return mgdg_dispatch_1(m, a, nsp, mix, st);
    case 2: return mgdg_dispatch_2(m, a, nsp, mix, st);
    case 3: return mgdg_dispatch_3(m, a, nsp, mix, st);
    case 4: return mgdg_dispatch_4(m, a, nsp, mix, st);
    case 5: return mgdg_dispatch_5(m, a, nsp, mix, st);
    case 6: return mgdg_dispatch_6(m, a, nsp, mix, st);
    case 7: return mgdg_dispatch_7(m, a, nsp, mix, st);
    case 8: return mgdg_dispatch_8(m, a, nsp, mix, st);
  }
  return fail(MG_ERR_ARG, "kernel selector");
}

// Kernel variant of a gas model: 0 ideal gas, 1 uiuc_7sp mixture, 2 mixture
// with mixture-averaged transport.
int gas_variant(const mg_gas* g) {
  return g->eos_kind != 1 ? 0 : (g->transport_kind == 3 ? 2 : 1);
}

mg_status check_gas(const mg_gas* g, bool need_transport) {
  if (!g) return fail(MG_ERR_ARG, "gas is NULL");
  if (g->transport_kind < 0 || g->transport_kind > 3)
    return fail(MG_ERR_ARG, "transport_kind");
  if (g->transport_kind == 3 && g->eos_kind != 1)
    return fail(MG_ERR_ARG, "mixture-averaged transport needs the mixture EOS");
  if (g->transport_kind == 3 && !mgdg_shared::tran_loaded())
    return fail(MG_ERR_ARG, "transport fits not loaded (mg_load_transport)");
  if (g->eos_kind == 1 && g->nspecies != 7)
    return fail(MG_ERR_ARG, "the mixture EOS needs the 7 uiuc_7sp species");
  if (g->eos_kind != 0 && g->eos_kind != 1) return fail(MG_ERR_ARG, "eos_kind");
  if (g->nspecies < 0 || g->nspecies > 8) return fail(MG_ERR_ARG, "nspecies");
  if (need_transport && g->transport_kind == 0)
    return fail(MG_ERR_ARG, "Navier-Stokes operator expects a viscous gas model");
  return MG_OK;
}

mg_status ensure_mech() {
  if (mgdg_shared::mech_loaded()) return MG_OK;
  return fail(MG_ERR_ARG, "mixture mechanism not loaded (mg_load_mechanism)");
}

template <int D, int P>
constexpr int64_t ops_bytes_aligned() {
  return (int64_t(mgdg::Ops<D, P>::BYTES) + 255) / 256 * 256;
}

template <int D, int P>
mg_status pack_ops(mg_mesh* m, double* ws, hipStream_t st) {
  const int n = mgdg::Ops<D, P>::TOTAL + mgdg::Ops<D, P>::MF;
  auto kern = mgdg::pack_ops_kernel<D, P>;
  hipLaunchKernelGGL(kern, dim3((n + 255) / 256), dim3(256), 0, st, m->weak_diff,
                     m->lift, ws);
  mg_status s = check_launch("pack_ops_kernel");
  if (s) return s;
  m->workspace = ws;
  m->block_order = nullptr;
  if (m->elem_rank) {
    const int64_t K = m->nelements;
    int32_t* tab = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(ws) +
                                              ops_bytes_aligned<D, P>());
    for (int i = 0; i < mgdg::kOrderTables; ++i) {
      const int epb = mgdg::kOrderEpb[i];
      const int64_t nb = (K + epb - 1) / epb;
      hipLaunchKernelGGL(mgdg::order_fill_kernel, dim3((K + 255) / 256), dim3(256), 0, st, tab,
                         K);
      hipLaunchKernelGGL(mgdg::order_mark_kernel, dim3((nb + 255) / 256), dim3(256), 0, st,
                         m->elem_rank, tab, K, epb);
      hipLaunchKernelGGL(mgdg::order_compact_kernel, dim3(1), dim3(1024), 0, st, tab,
                         tab + mgdg::order_table_offset(K, i), K, epb);
      if ((s = check_launch("order tables"))) return s;
    }
    m->block_order = tab;
  }
  return MG_OK;
}

}  // namespace

inline mg_status set_wall(mgdg::KArgs& a, const mg_mesh* m, const mg_wall* w, bool mix) {
  if (!w || !w->flags || m->nbnd_nodes == 0) return MG_OK;
  if (!w->q_bc || !w->t_bc || (mix && !w->t_state_bc) || !w->vol_index ||
      !w->normals || !w->work)
    return fail(MG_ERR_ARG,
                "mg_wall: q_bc, t_bc, vol_index, normals, work (and t_state_bc for "
                "mixtures) needed");
  a.wq = w->q_bc;
  a.wt = w->t_bc;
  a.wts = w->t_state_bc;
  a.wflag = w->flags;
  a.wvol = w->vol_index;
  a.wn = w->normals;
  a.wwork = w->work;
  return MG_OK;
}

template <int DIM, int NSP, int MIX, bool FUSED>
mg_status launch_wall(const mgdg::KArgs& a, int shift, hipStream_t st) {
  const unsigned nblk = unsigned((a.nbnd + mgdg::kBlock - 1) / mgdg::kBlock);
  if (FUSED && shift)
    hipLaunchKernelGGL((mgdg::wall_flux_kernel<DIM, NSP, MIX, FUSED, FUSED>), dim3(nblk),
                       dim3(mgdg::kBlock), 0, st, a);
  else
    hipLaunchKernelGGL((mgdg::wall_flux_kernel<DIM, NSP, MIX, FUSED, false>), dim3(nblk),
                       dim3(mgdg::kBlock), 0, st, a);
  return check_launch("wall_flux_kernel");
}

// Mengaldo boundary nodes: after grad_kernel<FLUX> (fused) or before
// rhs_kernel<kGrad>; nothing to do without wall data.
template <bool FUSED>
mg_status wall_pass(const mgdg::KArgs& a, int dim, int nsp, int mix, int shift,
                    hipStream_t st) {
  if (!a.wflag) return MG_OK;
  if (dim == 2) {
    if (nsp == 0 && !mix) return launch_wall<2, 0, 0, FUSED>(a, shift, st);
    if (nsp == 7 && !mix) return launch_wall<2, 7, 0, FUSED>(a, shift, st);
    if (nsp == 7 && mix == 1) return launch_wall<2, 7, 1, FUSED>(a, shift, st);
    if (nsp == 7 && mix == 2) return launch_wall<2, 7, 2, FUSED>(a, shift, st);
  } else {
    if (nsp == 0 && !mix) return launch_wall<3, 0, 0, FUSED>(a, shift, st);
    if (nsp == 7 && !mix) return launch_wall<3, 7, 0, FUSED>(a, shift, st);
    if (nsp == 7 && mix == 1) return launch_wall<3, 7, 1, FUSED>(a, shift, st);
    if (nsp == 7 && mix == 2) return launch_wall<3, 7, 2, FUSED>(a, shift, st);
  }
  return fail(MG_ERR_UNSUPPORTED,
              "unsupported species count / EOS combination (nspecies 0 or 7)");
}

template <int DIM, int NSP, int MIX>
mg_status launch_timestep(const mg_mesh* m, const mgdg::GasDev& g, const double* q,
                          const double* t, const double* h, int viscous, double* dt,
                          hipStream_t st) {
  const int64_t n = m->nelements * m->np;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipLaunchKernelGGL((mgdg::local_timestep_kernel<DIM, NSP, MIX>), dim3(nblk),
                     dim3(mgdg::kBlock), 0, st, n, int(m->np), q, t, h, g, viscous, dt);
  return check_launch("local_timestep_kernel");
}

template <int DIM>
mg_status timestep_species(const mg_mesh* m, const mg_gas* gas, const double* q,
                           const double* t, const double* h, int viscous, double* dt,
                           hipStream_t st) {
  const mgdg::GasDev g = to_dev(gas);
  const int mix = gas_variant(gas);
  if (gas->nspecies == 0 && !mix)
    return launch_timestep<DIM, 0, 0>(m, g, q, t, h, viscous, dt, st);
  if (gas->nspecies == 7 && !mix)
    return launch_timestep<DIM, 7, 0>(m, g, q, t, h, viscous, dt, st);
  if (gas->nspecies == 7 && mix == 1)
    return launch_timestep<DIM, 7, 1>(m, g, q, t, h, viscous, dt, st);
  if (gas->nspecies == 7 && mix == 2)
    return launch_timestep<DIM, 7, 2>(m, g, q, t, h, viscous, dt, st);
  return fail(MG_ERR_UNSUPPORTED,
              "unsupported species count / EOS combination (nspecies 0 or 7)");
}

// }}}
