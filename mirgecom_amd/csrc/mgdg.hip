// MI355X (gfx950) DG right-hand-side library: the C ABI (include/mgdg.h)
// and the pointwise entry points.  The element-batched kernels live in
// mgdg_kernels.h; their instantiations are spread over mg_disp_*.hip and the
// quadrature / entropy-stable entry points over mg_quad.hip, so that the
// files compile in parallel.
#include <mutex>
#include <vector>

#include "mgdg_host.h"

// {{{ state shared by all files of the library

namespace mgdg_shared {

std::string& err() {
  thread_local std::string s;
  return s;
}

bool& mech_loaded() {
  static bool f = false;
  return f;
}

bool& tran_loaded() {
  static bool f = false;
  return f;
}

namespace {
std::vector<UploadFn>& uploads() {
  static std::vector<UploadFn> v;
  return v;
}
std::mutex& uploads_mutex() {
  static std::mutex m;
  return m;
}
}  // namespace

int register_upload(UploadFn fn) {
  std::lock_guard<std::mutex> lk(uploads_mutex());
  uploads().push_back(fn);
  return int(uploads().size());
}

hipError_t upload_all(const mgdg::MechDev* m, const mgdg::PolyDev* p,
                      const mgdg::TranDev* t) {
  std::lock_guard<std::mutex> lk(uploads_mutex());
  for (UploadFn fn : uploads()) {
    hipError_t e = fn(m, p, t);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mgdg_shared

// }}}

extern "C" {

const char* mg_last_error(void) { return mgdg_shared::err().c_str(); }

int32_t mg_abi_version(void) { return 6; }

/* Upload the uiuc_7sp mechanism constants (host array of 7*7*2 + 7 + 7 + 7
 * + 2 + 9 doubles, packed as in mgdg::MechDev). */
mg_status mg_load_mechanism(const double* packed, int64_t count) {
  if (count != int64_t(sizeof(mgdg::MechDev) / sizeof(double)))
    return fail(MG_ERR_ARG, "mechanism size mismatch");
  mgdg::MechDev m;
  std::memcpy(&m, packed, sizeof(m));
  mgdg::PolyDev poly;
  for (int r = 0; r < 2; ++r)
    for (int k = 0; k < mgdg::MIX_NSP; ++k) {
      const double* a = r ? m.nasa_hi[k] : m.nasa_lo[k];
      for (int i = 0; i < 5; ++i) poly.cp[r][k][i] = a[i];
      for (int i = 0; i < 5; ++i) poly.h[r][k][i] = a[i] / double(i + 1);
      poly.h[r][k][5] = a[5];
      poly.s[r][k][0] = a[0];
      for (int i = 1; i < 5; ++i) poly.s[r][k][i] = a[i] / double(i);
      poly.s[r][k][5] = a[6];
    }
  hipError_t e = mgdg_shared::upload_all(&m, &poly, nullptr);
  if (e != hipSuccess) return fail(MG_ERR_HIP, hipGetErrorString(e));
  mgdg_shared::mech_loaded() = true;
  return MG_OK;
}

mg_status mg_load_transport(const double* packed, int64_t count) {
  if (!packed || count != int64_t(sizeof(mgdg::TranDev) / sizeof(double)))
    return fail(MG_ERR_ARG, "transport size mismatch");
  mgdg::TranDev t;
  std::memcpy(&t, packed, sizeof(t));
  hipError_t e = mgdg_shared::upload_all(nullptr, nullptr, &t);
  if (e != hipSuccess) return fail(MG_ERR_HIP, hipGetErrorString(e));
  mgdg_shared::tran_loaded() = true;
  return MG_OK;
}

int64_t mg_mesh_workspace_bytes(const mg_mesh* m) {
  if (!m) return -1;
  int64_t total = -1;
#define MG_WS_CASE(D, P) \
  if (m->dim == D && m->order == P) total = ops_bytes_aligned<D, P>();
  MG_WS_CASE(2, 1) MG_WS_CASE(2, 2) MG_WS_CASE(2, 3) MG_WS_CASE(2, 4)
  MG_WS_CASE(3, 1) MG_WS_CASE(3, 2) MG_WS_CASE(3, 3) MG_WS_CASE(3, 4)
#undef MG_WS_CASE
  if (total < 0) return -1;
  if (m->elem_rank) {
    if (m->nelements <= 0) return -1;
    total += 4 * mgdg::order_table_offset(m->nelements, mgdg::kOrderTables);
  }
  return total;
}

mg_status mg_mesh_prepare(mg_mesh* m, double* workspace, void* stream) {
  if (!m || !workspace) return fail(MG_ERR_ARG, "mesh/workspace NULL");
  if (!m->weak_diff || !m->lift) return fail(MG_ERR_ARG, "mesh matrices NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  mg_status s = MG_ERR_UNSUPPORTED;
#define MG_PACK_CASE(D, P) \
  if (m->dim == D && m->order == P) s = pack_ops<D, P>(m, workspace, st);
  MG_PACK_CASE(2, 1) MG_PACK_CASE(2, 2) MG_PACK_CASE(2, 3) MG_PACK_CASE(2, 4)
  MG_PACK_CASE(3, 1) MG_PACK_CASE(3, 2) MG_PACK_CASE(3, 3) MG_PACK_CASE(3, 4)
#undef MG_PACK_CASE
  if (s == MG_ERR_UNSUPPORTED) return fail(s, "dim/order");
  return s;
}

mg_status mg_euler_rhs(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                       const double* temperature, const double* q_bnd,
                       const double* t_bnd, int32_t flux_kind, double* rhs,
                       void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  const int mix = gas_variant(gas);
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
  return dispatch(mesh, a, gas->nspecies, mix, 0, static_cast<hipStream_t>(stream));
}

mg_status mg_euler_rhs_mixed(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                             const double* temperature, const double* q_bnd,
                             const double* t_bnd, int32_t flux_kind, double* rhs,
                             void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  const int mix = gas_variant(gas);
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
  return dispatch(mesh, a, gas->nspecies, mix, 8, static_cast<hipStream_t>(stream));
}

mg_status mg_ns_gradient(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                         const double* temperature, const double* q_bnd,
                         const double* t_bnd, const mg_wall* wall, double* grad,
                         void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  const int mix = gas_variant(gas);
  if (!q || !grad) return fail(MG_ERR_ARG, "q/grad NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.qb = q_bnd;
  a.tb = t_bnd;
  if ((s = set_wall(a, mesh, wall, mix))) return s;
  a.out = grad;
  return dispatch(mesh, a, gas->nspecies, mix, 3, static_cast<hipStream_t>(stream));
}

mg_status mg_ns_rhs(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                    const double* temperature, const double* grad,
                    const double* q_bnd, const double* t_bnd, const mg_wall* wall,
                    int32_t flux_kind, int32_t inviscid_on, double* rhs, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const int mix = gas_variant(gas);
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
  if ((s = set_wall(a, mesh, wall, mix))) return s;
  a.out = rhs;
  a.flux_kind = flux_kind;
  a.inviscid_on = inviscid_on ? 1 : 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if ((s = wall_pass<false>(a, mesh->dim, gas->nspecies, mix, 0, st))) return s;
  return dispatch(mesh, a, gas->nspecies, mix, 1, st);
}

mg_status mg_ns_flux(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                     const double* temperature, const double* q_bnd,
                     const double* t_bnd, const mg_wall* wall, int32_t inviscid_on,
                     double* fnet, double* fnet_bnd, double* grad, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const int mix = gas_variant(gas);
  if (mix && (s = ensure_mech())) return s;
  if (!q || !fnet) return fail(MG_ERR_ARG, "q/fnet NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || !fnet_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state / flux buffer missing");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.qb = q_bnd;
  a.tb = t_bnd;
  if ((s = set_wall(a, mesh, wall, mix))) return s;
  a.out = fnet;
  a.grad_out = grad;
  a.fnb_out = fnet_bnd;
  a.inviscid_on = inviscid_on ? 1 : 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (mix) {
    // mixture: gradients through `grad`, then node-parallel net fluxes
    if (!grad) return fail(MG_ERR_ARG, "the mixture path needs grad (gradient staging)");
    a.out = grad;
    a.grad_out = nullptr;
    if ((s = dispatch(mesh, a, gas->nspecies, mix, 3, st))) return s;
    a.out = fnet;
    a.grad = grad;
    if ((s = dispatch(mesh, a, gas->nspecies, mix, 7, st))) return s;
  } else if ((s = dispatch(mesh, a, gas->nspecies, mix, 4, st))) {
    return s;
  }
  return wall_pass<true>(a, mesh->dim, gas->nspecies, mix, 0, st);
}

mg_status mg_ns_div(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                    const double* temperature, const double* fnet,
                    const double* q_bnd, const double* t_bnd,
                    const double* fnet_bnd, int32_t flux_kind,
                    int32_t inviscid_on, double* rhs, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const int mix = gas_variant(gas);
  if (mix && (s = ensure_mech())) return s;
  if (!q || !fnet || !rhs) return fail(MG_ERR_ARG, "q/fnet/rhs NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || !fnet_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state / flux missing");
  if (flux_kind < 0 || flux_kind > 2) return fail(MG_ERR_ARG, "flux_kind");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.fnet = fnet;
  a.fnb = fnet_bnd;
  a.qb = q_bnd;
  a.tb = t_bnd;
  a.out = rhs;
  a.flux_kind = flux_kind;
  a.inviscid_on = inviscid_on ? 1 : 0;
  return dispatch(mesh, a, gas->nspecies, mix, 2, static_cast<hipStream_t>(stream));
}

mg_status mg_ns_flux_mixed(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                           const double* temperature, const double* q_bnd,
                           const double* t_bnd, const mg_wall* wall,
                           int32_t inviscid_on, float* fnet, double* fnet_bnd,
                           float* grad, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const int mix = gas_variant(gas);
  if (mix && (s = ensure_mech())) return s;
  if (!q || !fnet) return fail(MG_ERR_ARG, "q/fnet NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || !fnet_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state / flux buffer missing");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.qb = q_bnd;
  a.tb = t_bnd;
  if ((s = set_wall(a, mesh, wall, mix))) return s;
  a.fnet32_out = fnet;
  a.fnb_out = fnet_bnd;
  a.inviscid_on = inviscid_on ? 1 : 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (mix) {
    // mixture: FP32 gradients through `grad`, then node-parallel FP32 net
    // fluxes (the fused kernel holds four 12-field elements per workgroup,
    // too few nodes for the pointwise transport and NASA7 work)
    if (!grad) return fail(MG_ERR_ARG, "the mixture path needs grad (FP32 gradient staging)");
    a.grad32_out = grad;
    if ((s = dispatch(mesh, a, gas->nspecies, mix, 5, st))) return s;
    a.grad32_out = nullptr;
    a.grad32 = grad;
    if ((s = dispatch(mesh, a, gas->nspecies, mix, 7, st))) return s;
  } else if ((s = dispatch(mesh, a, gas->nspecies, mix, 5, st))) {
    return s;
  }
  return wall_pass<true>(a, mesh->dim, gas->nspecies, mix, 1, st);
}

mg_status mg_ns_div_mixed(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                          const double* temperature, const float* fnet,
                          const double* q_bnd, const double* t_bnd,
                          const double* fnet_bnd, int32_t flux_kind,
                          int32_t inviscid_on, double* rhs, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const int mix = gas_variant(gas);
  if (mix && (s = ensure_mech())) return s;
  if (!q || !fnet || !rhs) return fail(MG_ERR_ARG, "q/fnet/rhs NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  if (mesh->nbnd_nodes > 0 && (!q_bnd || !fnet_bnd || (mix && !t_bnd)))
    return fail(MG_ERR_ARG, "boundary exterior state / flux missing");
  if (flux_kind < 0 || flux_kind > 2) return fail(MG_ERR_ARG, "flux_kind");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.fnet32 = fnet;
  a.fnb = fnet_bnd;
  a.qb = q_bnd;
  a.tb = t_bnd;
  a.out = rhs;
  a.flux_kind = flux_kind;
  a.inviscid_on = inviscid_on ? 1 : 0;
  return dispatch(mesh, a, gas->nspecies, mix, 6, static_cast<hipStream_t>(stream));
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

mg_status mg_mixture_temperature_tol(const mg_gas* gas, int32_t dim, int64_t n,
                                     const double* q, const double* t_seed, double tol,
                                     int32_t max_iter, double* t_out, int32_t* iters,
                                     void* stream) {
  mg_status s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 1) return fail(MG_ERR_ARG, "mixture EOS required");
  if ((s = ensure_mech())) return s;
  if (!q || !t_seed || !t_out || n < 0 || max_iter < 1 || !(tol > 0.0))
    return fail(MG_ERR_ARG, "bad arguments (tol > 0, max_iter >= 1)");
  if (dim != 2 && dim != 3) return fail(MG_ERR_ARG, "dim");
  if (n == 0) return MG_OK;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dim == 2)
    hipLaunchKernelGGL(mgdg::mix_temperature_tol_kernel<2>, dim3(nblk), dim3(mgdg::kBlock), 0,
                       st, n, q, t_seed, tol, int(max_iter), t_out, iters);
  else
    hipLaunchKernelGGL(mgdg::mix_temperature_tol_kernel<3>, dim3(nblk), dim3(mgdg::kBlock), 0,
                       st, n, q, t_seed, tol, int(max_iter), t_out, iters);
  return check_launch("mix_temperature_tol_kernel");
}

}  // extern "C"

namespace {

mg_status species_source_launch(const mg_gas* gas, int32_t dim, int64_t n,
                                const double* q, const double* temperature, double* out,
                                bool add, void* stream) {
  mg_status s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 1) return fail(MG_ERR_ARG, "mixture EOS required");
  if ((s = ensure_mech())) return s;
  if (!q || !temperature || !out || n < 0) return fail(MG_ERR_ARG, "bad arguments");
  if (dim != 2 && dim != 3) return fail(MG_ERR_ARG, "dim");
  if (n == 0) return MG_OK;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define MG_SRC(D, A)                                                                 \
  hipLaunchKernelGGL((mgdg::species_source_kernel<D, A>), dim3(nblk), dim3(mgdg::kBlock), \
                     0, st, n, q, temperature, out);
  if (dim == 2) {
    if (add) MG_SRC(2, true) else MG_SRC(2, false)
  } else {
    if (add) MG_SRC(3, true) else MG_SRC(3, false)
  }
#undef MG_SRC
  return check_launch("species_source_kernel");
}

}  // namespace

extern "C" {

mg_status mg_species_source(const mg_gas* gas, int32_t dim, int64_t n,
                            const double* q, const double* temperature,
                            double* src, void* stream) {
  return species_source_launch(gas, dim, n, q, temperature, src, false, stream);
}

mg_status mg_species_source_add(const mg_gas* gas, int32_t dim, int64_t n,
                                const double* q, const double* temperature,
                                double* rhs, void* stream) {
  return species_source_launch(gas, dim, n, q, temperature, rhs, true, stream);
}

mg_status mg_local_timestep(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                            const double* temperature, const double* lengthscale,
                            int32_t viscous, double* dt, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, viscous != 0))) return s;
  const int mix = gas_variant(gas);
  if (mix && (s = ensure_mech())) return s;
  if (!q || !lengthscale || !dt) return fail(MG_ERR_ARG, "q/lengthscale/dt NULL");
  if (mix && !temperature) return fail(MG_ERR_ARG, "mixture needs temperature");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (mesh->dim == 2)
    return timestep_species<2>(mesh, gas, q, temperature, lengthscale, viscous ? 1 : 0,
                               dt, st);
  return timestep_species<3>(mesh, gas, q, temperature, lengthscale, viscous ? 1 : 0,
                             dt, st);
}

mg_status mg_transport_vars(const mg_gas* gas, int32_t dim, int64_t n,
                            const double* q, const double* temperature, double* mu,
                            double* mu_b, double* kappa, double* diff, void* stream) {
  mg_status s;
  if ((s = check_gas(gas, true))) return s;
  const int mix = gas_variant(gas);
  if (mix && (s = ensure_mech())) return s;
  if (!mix && gas->nspecies != 0)
    return fail(MG_ERR_UNSUPPORTED, "ideal gas with species");
  if (!q || !mu || !mu_b || !kappa || n < 0 || (mix && (!temperature || !diff)))
    return fail(MG_ERR_ARG, "bad arguments");
  if (dim != 2 && dim != 3) return fail(MG_ERR_ARG, "dim");
  if (n == 0) return MG_OK;
  const unsigned nblk = unsigned((n + mgdg::kBlock - 1) / mgdg::kBlock);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const mgdg::GasDev g = to_dev(gas);
#define MG_TV(D, NS, MX)                                                          \
  hipLaunchKernelGGL((mgdg::transport_vars_kernel<D, NS, MX>), dim3(nblk),        \
                     dim3(mgdg::kBlock), 0, st, g, n, q, temperature, mu, mu_b, \
                     kappa, diff)
  if (mix == 2) {
    if (dim == 2) MG_TV(2, 7, 2); else MG_TV(3, 7, 2);
  } else if (mix) {
    if (dim == 2) MG_TV(2, 7, 1); else MG_TV(3, 7, 1);
  } else {
    if (dim == 2) MG_TV(2, 0, 0); else MG_TV(3, 0, 0);
  }
#undef MG_TV
  return check_launch("transport_vars_kernel");
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


mg_status mg_bound_limiter(int32_t np, int64_t nelements, int32_t nfields,
                           const double* weights, const double* field, double* out,
                           double mmin, double mmax, int32_t flags, void* stream) {
  if (!weights || !field || !out || nelements < 0 || nfields < 0)
    return fail(MG_ERR_ARG, "bad arguments");
  if (flags & ~(mgdg::kLimFlagMax | mgdg::kLimFlagModifyAvg))
    return fail(MG_ERR_ARG, "unknown limiter flags");
  if (nfields > 65535) return fail(MG_ERR_ARG, "nfields > 65535");
  if (nelements == 0 || nfields == 0) return MG_OK;
  const int64_t nblk64 = (nelements + mgdg::kLimElems - 1) / mgdg::kLimElems;
  if (nblk64 > 0x7fffffff) return fail(MG_ERR_ARG, "too many elements");
  const dim3 grid{unsigned(nblk64), unsigned(nfields), 1u};
  hipStream_t st = static_cast<hipStream_t>(stream);
#define MG_LIM(N)                                                                   \
  case N:                                                                           \
    hipLaunchKernelGGL(mgdg::bound_limiter_kernel<N>, grid, dim3(mgdg::kLimElems), 0, \
                       st, nelements, weights, field, out, mmin, mmax, flags);      \
    break;
  switch (np) {
    MG_LIM(3) MG_LIM(4) MG_LIM(6) MG_LIM(10) MG_LIM(15) MG_LIM(20) MG_LIM(21)
    MG_LIM(35) MG_LIM(56)
    default:
      return fail(MG_ERR_UNSUPPORTED, "np must be a simplex node count (3..56)");
  }
#undef MG_LIM
  return check_launch("bound_limiter_kernel");
}

}  // extern "C"

// }}}
