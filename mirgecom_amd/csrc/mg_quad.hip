// Quadrature (overintegration) and entropy-stable Euler entry points of
// libmgdg (include/mgdg.h): mg_euler_rhs_quad, mg_euler_rhs_esdg,
// mg_euler_rhs_esdg_mixture, mg_euler_rhs_esdg_noalloc.
#include <algorithm>
#include <mutex>
#include <string>

#include "mgdg_host.h"

extern "C" {

mg_status mg_euler_rhs_quad(const mg_mesh* mesh, const mg_gas* gas,
                            const mg_quad* quad, const double* q,
                            const double* temperature, const double* q_bnd_quad,
                            const double* t_bnd_quad, int32_t niter, int32_t flux_kind,
                            double* rhs, void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  const bool mix = gas->eos_kind == 1;
  if (mix && (s = ensure_mech())) return s;
  if (gas->nspecies != 0 && gas->nspecies != 7)
    return fail(MG_ERR_UNSUPPORTED, "nspecies must be 0 or 7");
  if (mix && (!temperature || niter < 0 || (mesh->nbnd_nodes > 0 && !t_bnd_quad)))
    return fail(MG_ERR_ARG, "mixture needs temperature (and t_bnd_quad), niter >= 0");
  if (!quad || !quad->interp || !quad->weak_diff || !quad->face_interp || !quad->lift ||
      quad->nq < 1 || quad->nfq < 1)
    return fail(MG_ERR_ARG, "mg_quad incomplete");
  if (!q || !rhs) return fail(MG_ERR_ARG, "q/rhs NULL");
  if (mesh->nbnd_nodes > 0 && !q_bnd_quad)
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  if (mesh->nbnd_nodes % mesh->nfp) return fail(MG_ERR_ARG, "nbnd_nodes % nfp");
  if (flux_kind < 0 || flux_kind > 2) return fail(MG_ERR_ARG, "flux_kind");
  const int nfld = 2 + mesh->dim + gas->nspecies;
  // LDS row strides = 4 mod 32 doubles (bank spread of the matrix-core reads)
  auto bank_stride = [](int n) { return n + ((4 - n % 32) + 32) % 32; };
  const int gs = bank_stride(mesh->dim * quad->nq);
  const int fs = bank_stride(mesh->nfaces * quad->nfq);
  const size_t per_elem =
      sizeof(double) * (size_t(nfld) * (size_t(mesh->np) + size_t(gs) + size_t(fs) +
                                        size_t(mesh->nfaces) * mesh->nfp) +
                        size_t(mesh->np) + size_t(mesh->nfaces) * mesh->nfp);
  // up to 80 KiB per workgroup: two workgroups per CU, which is what the
  // kernel's registers allow (at least one element; at most 128 KiB)
  constexpr size_t kBudget = 128 * 1024;
  if (per_elem > kBudget) return fail(MG_ERR_UNSUPPORTED, "quadrature too large for LDS");
  int epb = int((80 * 1024) / per_elem);
  if (epb < 1) epb = 1;
  if (epb > 16) epb = 16;
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.out = rhs;
  a.flux_kind = flux_kind;
  mgdg::QuadArgs qa{};
  qa.nq = quad->nq;
  qa.nfq = quad->nfq;
  qa.epb = epb;
  qa.gs = gs;
  qa.fs = fs;
  qa.interp = quad->interp;
  qa.wdiff = quad->weak_diff;
  qa.finterp = quad->face_interp;
  qa.lift = quad->lift;
  qa.qbq = q_bnd_quad;
  qa.nbq = mesh->nbnd_nodes / mesh->nfp * quad->nfq;
  qa.temp = temperature;
  qa.tbq = t_bnd_quad;
  qa.niter = niter;
  const size_t lds = per_elem * epb;
  const unsigned nblk = unsigned((a.K + epb - 1) / epb);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define MG_QCASE(D, P, NS, MX)                                                    \
  if (mesh->dim == D && mesh->order == P && gas->nspecies == NS && int(mix) == MX) { \
    auto kern = mgdg::euler_quad_kernel<D, P, NS, MX>;                            \
    if ((s = set_lds(kern, kBudget))) return s;                                   \
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(mgdg::kBlock), lds, st, a, qa);     \
    return check_launch("euler_quad_kernel");                                     \
  }
  MG_QCASE(2, 1, 0, 0) MG_QCASE(2, 2, 0, 0) MG_QCASE(2, 3, 0, 0) MG_QCASE(2, 4, 0, 0)
  MG_QCASE(3, 1, 0, 0) MG_QCASE(3, 2, 0, 0) MG_QCASE(3, 3, 0, 0) MG_QCASE(3, 4, 0, 0)
  MG_QCASE(2, 1, 7, 0) MG_QCASE(2, 2, 7, 0) MG_QCASE(2, 3, 7, 0) MG_QCASE(2, 4, 7, 0)
  MG_QCASE(3, 1, 7, 0) MG_QCASE(3, 2, 7, 0) MG_QCASE(3, 3, 7, 0) MG_QCASE(3, 4, 7, 0)
  MG_QCASE(2, 1, 7, 1) MG_QCASE(2, 2, 7, 1) MG_QCASE(2, 3, 7, 1) MG_QCASE(2, 4, 7, 1)
  MG_QCASE(3, 1, 7, 1) MG_QCASE(3, 2, 7, 1) MG_QCASE(3, 3, 7, 1) MG_QCASE(3, 4, 7, 1)
#undef MG_QCASE
  return fail(MG_ERR_UNSUPPORTED, "dim/order");
}

}  // extern "C"

namespace {

// Checks shared by both entropy-stable entry points (everything that does
// not depend on where the projected entropy variables live).
mg_status esdg_check(const mg_mesh* mesh, const mg_gas* gas, const mg_quad* quad,
                     const mg_esdg* esdg, const double* q, const double* q_bnd_quad,
                     const double* rhs) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 0)
    return fail(MG_ERR_UNSUPPORTED, "the entropy-stable path is built for the ideal gas");
  if (gas->nspecies != 0 && gas->nspecies != 7)
    return fail(MG_ERR_UNSUPPORTED, "nspecies must be 0 or 7");
  if (!quad || !quad->interp || !quad->face_interp || quad->nq < 1 || quad->nfq < 1)
    return fail(MG_ERR_ARG, "mg_quad incomplete");
  if (!esdg || !esdg->proj || !esdg->vh || !esdg->skew || !esdg->ext || !esdg->ext_col ||
      !esdg->apply || !esdg->face_weights ||
      esdg->nh != quad->nq + mesh->nfaces * quad->nfq)
    return fail(MG_ERR_ARG, "mg_esdg incomplete or nh != nq + nfaces*nfq");
  if (!q || !rhs) return fail(MG_ERR_ARG, "q/rhs NULL");
  if (mesh->nbnd_nodes > 0 && !q_bnd_quad)
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  if (mesh->nbnd_nodes % mesh->nfp) return fail(MG_ERR_ARG, "nbnd_nodes % nfp");
  return MG_OK;
}

mg_status esdg_run(const mg_mesh* mesh, const mg_gas* gas, const mg_quad* quad,
                   const mg_esdg* esdg, const double* q, const double* q_bnd_quad,
                   int32_t dissipation, double* ev_work, double* rhs, void* stream) {
  mg_status s = esdg_check(mesh, gas, quad, esdg, q, q_bnd_quad, rhs);
  if (s) return s;
  if (!ev_work) return fail(MG_ERR_ARG, "ev_work NULL");
  const int nfld = 2 + mesh->dim + gas->nspecies;
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.out = rhs;
  mgdg::EsdgArgs ea{};
  ea.nq = quad->nq;
  ea.nfq = quad->nfq;
  ea.nh = esdg->nh;
  ea.interp = quad->interp;
  ea.finterp = quad->face_interp;
  ea.proj = esdg->proj;
  ea.vh = esdg->vh;
  ea.skew = esdg->skew;
  ea.ext = esdg->ext;
  ea.ext_col = esdg->ext_col;
  ea.apply = esdg->apply;
  ea.wf = esdg->face_weights;
  ea.qbq = q_bnd_quad;
  ea.nbq = mesh->nbnd_nodes / mesh->nfp * quad->nfq;
  ea.ev = ev_work;
  ea.dissipation = dissipation ? 1 : 0;
  ea.skew_diag = esdg->skew_diag;
  ea.ext_diag = esdg->ext_diag;
  ea.paired = (mesh->dim == 3 && quad->nq == 64 && mesh->nfaces * quad->nfq == 192 &&
               esdg->skew_diag && esdg->ext_diag && mgdg::kBlock == 256)
                  ? 1
                  : 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  constexpr size_t kBudget = 128 * 1024;
  const size_t nft = size_t(mesh->nfaces) * quad->nfq;
#define MG_ECASE(D, P, NS)                                                              \
  if (mesh->dim == D && mesh->order == P && gas->nspecies == NS) {                      \
    const size_t per_p = sizeof(double) * nfld * quad->nq;                              \
    if (per_p > kBudget) return fail(MG_ERR_UNSUPPORTED, "quadrature too large for LDS"); \
    int epb_p = int(std::max<size_t>(1, std::min<size_t>(8, 256 / quad->nq)));        \
    ea.epb = epb_p;                                                                     \
    auto kp = mgdg::esdg_project_kernel<D, P, NS>;                                      \
    if ((s = set_lds(kp, kBudget))) return s;                                           \
    hipLaunchKernelGGL(kp, dim3(unsigned((a.K + epb_p - 1) / epb_p)),                   \
                       dim3(mgdg::kBlock), per_p * epb_p, st, a, ea);                  \
    if ((s = check_launch("esdg_project_kernel"))) return s;                            \
    const size_t nchunk = (size_t(esdg->nh) + quad->nq - 1) / quad->nq;               \
    const size_t nacc = std::max(nchunk * quad->nq + nft, nft);                         \
    const size_t per_e = sizeof(mgdg::EcPair<D, NS>) * esdg->nh +                       \
        sizeof(double) * nfld *                                                         \
            (2 * size_t(mesh->np) + size_t(mesh->nfaces) * mesh->nfp +                  \
             (ea.paired ? 0 : nft) + nacc);                                             \
    if (per_e > kBudget) return fail(MG_ERR_UNSUPPORTED, "quadrature too large for LDS"); \
    int epb = ea.paired ? 1                                                             \
                        : int(std::max<size_t>(1, std::min<size_t>(16, (64 * 1024) / per_e))); \
    ea.epb = epb;                                                                       \
    auto kern = mgdg::esdg_kernel<D, P, NS>;                                            \
    if ((s = set_lds(kern, kBudget))) return s;                                         \
    hipLaunchKernelGGL(kern, dim3(unsigned((a.K + epb - 1) / epb)), dim3(mgdg::kBlock), \
                       per_e * epb, st, a, ea);                                         \
    return check_launch("esdg_kernel");                                                 \
  }
  MG_ECASE(2, 1, 0) MG_ECASE(2, 2, 0) MG_ECASE(2, 3, 0) MG_ECASE(2, 4, 0)
  MG_ECASE(3, 1, 0) MG_ECASE(3, 2, 0) MG_ECASE(3, 3, 0)
  MG_ECASE(2, 1, 7) MG_ECASE(2, 2, 7) MG_ECASE(2, 3, 7) MG_ECASE(2, 4, 7)
  MG_ECASE(3, 1, 7) MG_ECASE(3, 2, 7) MG_ECASE(3, 3, 7)
#undef MG_ECASE
  // 3-D p = 4 (nq = 125, nh = 425 points against 256 threads, ~80 KB of LDS
  // per element) is not instantiated: no parity test covers it
  if (mesh->dim == 3 && mesh->order == 4)
    return fail(MG_ERR_UNSUPPORTED, "entropy-stable path: 3-D p = 4 is not supported");
  return fail(MG_ERR_UNSUPPORTED, "dim/order");
}

// The uiuc_7sp mixture (Renac's two-point flux): esdg_project_mix_kernel
// then esdg_mix_kernel (row-by-row phase 4).
mg_status esdg_mix_run(const mg_mesh* mesh, const mg_gas* gas, const mg_quad* quad,
                       const mg_esdg* esdg, const double* q, const double* temperature,
                       const double* q_bnd_quad, const double* t_bnd_quad, int32_t niter,
                       int32_t dissipation, int32_t es_volume, double* ev_work, double* rhs,
                       void* stream) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, false))) return s;
  if (gas->eos_kind != 1)
    return fail(MG_ERR_ARG, "mg_euler_rhs_esdg_mixture needs the mixture EOS");
  if ((s = ensure_mech())) return s;
  if (!quad || !quad->interp || !quad->face_interp || quad->nq < 1 || quad->nfq < 1)
    return fail(MG_ERR_ARG, "mg_quad incomplete");
  if (!esdg || !esdg->proj || !esdg->vh || !esdg->skew || !esdg->ext || !esdg->ext_col ||
      !esdg->apply || !esdg->face_weights ||
      esdg->nh != quad->nq + mesh->nfaces * quad->nfq)
    return fail(MG_ERR_ARG, "mg_esdg incomplete or nh != nq + nfaces*nfq");
  if (!q || !rhs || !temperature || !ev_work)
    return fail(MG_ERR_ARG, "q/temperature/ev_work/rhs NULL");
  if (niter < 0) return fail(MG_ERR_ARG, "niter < 0");
  if (mesh->nbnd_nodes > 0 && (!q_bnd_quad || !t_bnd_quad))
    return fail(MG_ERR_ARG, "boundary exterior state / temperature missing");
  if (mesh->nbnd_nodes % mesh->nfp) return fail(MG_ERR_ARG, "nbnd_nodes % nfp");
  const int nfld = 2 + mesh->dim + gas->nspecies;
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.temp = temperature;
  a.out = rhs;
  mgdg::EsdgArgs ea{};
  ea.nq = quad->nq;
  ea.nfq = quad->nfq;
  ea.nh = esdg->nh;
  ea.interp = quad->interp;
  ea.finterp = quad->face_interp;
  ea.proj = esdg->proj;
  ea.vh = esdg->vh;
  ea.skew = esdg->skew;
  ea.ext = esdg->ext;
  ea.ext_col = esdg->ext_col;
  ea.apply = esdg->apply;
  ea.wf = esdg->face_weights;
  ea.qbq = q_bnd_quad;
  ea.tbq = t_bnd_quad;
  ea.nbq = mesh->nbnd_nodes / mesh->nfp * quad->nfq;
  ea.ev = ev_work;
  ea.dissipation = dissipation ? 1 : 0;
  ea.niter = niter;
  ea.es_volume = es_volume ? 1 : 0;
  ea.paired = 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  constexpr size_t kBudget = 128 * 1024;
  const size_t nft = size_t(mesh->nfaces) * quad->nfq;
  const size_t nfn = size_t(mesh->nfaces) * mesh->nfp;
#define MG_MCASE(D, P)                                                                  \
  if (mesh->dim == D && mesh->order == P) {                                             \
    const size_t per_p = sizeof(double) * nfld * quad->nq;                              \
    if (per_p > kBudget) return fail(MG_ERR_UNSUPPORTED, "quadrature too large for LDS"); \
    int epb_p = int(std::max<size_t>(1, std::min<size_t>(8, 256 / quad->nq)));        \
    ea.epb = epb_p;                                                                     \
    auto kp = mgdg::esdg_project_mix_kernel<D, P>;                                      \
    if ((s = set_lds(kp, kBudget))) return s;                                           \
    hipLaunchKernelGGL(kp, dim3(unsigned((a.K + epb_p - 1) / epb_p)),                   \
                       dim3(mgdg::kBlock), per_p * epb_p, st, a, ea);                  \
    if ((s = check_launch("esdg_project_mix_kernel"))) return s;                        \
    const size_t nchunk = (size_t(esdg->nh) + quad->nq - 1) / quad->nq;               \
    const size_t nacc = nchunk * quad->nq + nft;                                        \
    const size_t per_e = sizeof(mgdg::RnPair<D>) * esdg->nh +                           \
        sizeof(double) * (nfld * (2 * size_t(mesh->np) + nfn + nft + nacc) +            \
                          2 * size_t(mesh->np) + nfn + 2 * nft);                        \
    if (per_e > kBudget) return fail(MG_ERR_UNSUPPORTED, "quadrature too large for LDS"); \
    int epb = int(std::max<size_t>(1, std::min<size_t>(16, (64 * 1024) / per_e)));     \
    ea.epb = epb;                                                                       \
    auto kern = mgdg::esdg_mix_kernel<D, P>;                                            \
    if ((s = set_lds(kern, kBudget))) return s;                                         \
    hipLaunchKernelGGL(kern, dim3(unsigned((a.K + epb - 1) / epb)), dim3(mgdg::kBlock), \
                       per_e * epb, st, a, ea);                                         \
    return check_launch("esdg_mix_kernel");                                             \
  }
  if (gas->nspecies != 7) return fail(MG_ERR_UNSUPPORTED, "the mixture has 7 species");
  MG_MCASE(2, 1) MG_MCASE(2, 2) MG_MCASE(2, 3) MG_MCASE(2, 4)
  MG_MCASE(3, 1) MG_MCASE(3, 2) MG_MCASE(3, 3)
#undef MG_MCASE
  if (mesh->dim == 3 && mesh->order == 4)
    return fail(MG_ERR_UNSUPPORTED, "entropy-stable path: 3-D p = 4 is not supported");
  return fail(MG_ERR_UNSUPPORTED, "dim/order");
}

// {{{ allocation-free variant: the projected entropy variables in `scratch`
//
// `scratch` is a file-scope device array: each GPU that loads the library
// holds its own copy (64 MiB) until the process ends.  Host code never uses
// the array's name as an address (on the host it is only a stand-in for the
// symbol): the device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept in g_scratch_ptr.
// Every call on a GPU shares that GPU's copy, so calls that overlap on one
// GPU (different streams, or host threads) race; the projection kernel
// overwrites every value the flux-differencing kernel reads, so nothing needs
// clearing between calls.

constexpr int64_t kEsdgScratchDoubles = int64_t(1) << 23;   // 64 MiB
__device__ __align__(256) double scratch[kEsdgScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

mg_status device_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return fail(MG_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(MG_ERR_UNSUPPORTED, "device index beyond the 64 the scratch table holds");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MG_OK;
}

// }}}

}  // namespace

extern "C" {

mg_status mg_euler_rhs_esdg(const mg_mesh* mesh, const mg_gas* gas,
                            const mg_quad* quad, const mg_esdg* esdg, const double* q,
                            const double* q_bnd_quad, int32_t dissipation,
                            double* ev_work, double* rhs, void* stream) {
  return esdg_run(mesh, gas, quad, esdg, q, q_bnd_quad, dissipation, ev_work, rhs,
                  stream);
}

mg_status mg_euler_rhs_esdg_mixture(const mg_mesh* mesh, const mg_gas* gas,
                                    const mg_quad* quad, const mg_esdg* esdg,
                                    const double* q, const double* temperature,
                                    const double* q_bnd_quad, const double* t_bnd_quad,
                                    int32_t niter, int32_t dissipation, int32_t es_volume,
                                    double* ev_work, double* rhs, void* stream) {
  return esdg_mix_run(mesh, gas, quad, esdg, q, temperature, q_bnd_quad, t_bnd_quad,
                      niter, dissipation, es_volume, ev_work, rhs, stream);
}

mg_status mg_euler_rhs_esdg_noalloc(const mg_mesh* mesh, const mg_gas* gas,
                                    const mg_quad* quad, const mg_esdg* esdg,
                                    const double* q, const double* q_bnd_quad,
                                    int32_t dissipation, double* rhs, void* stream) {
  mg_status s = esdg_check(mesh, gas, quad, esdg, q, q_bnd_quad, rhs);
  if (s) return s;
  const int64_t need = int64_t(2 + mesh->dim + gas->nspecies) * mesh->nelements * mesh->np;
  if (need > kEsdgScratchDoubles)
    return fail(MG_ERR_ARG, "state of " + std::to_string(need) +
                                " values exceeds the allocation-free scratch of " +
                                std::to_string(kEsdgScratchDoubles) +
                                " doubles (fields x nodes); use mg_euler_rhs_esdg");
  double* ev = nullptr;
  if ((s = device_scratch(&ev))) return s;
  return esdg_run(mesh, gas, quad, esdg, q, q_bnd_quad, dissipation, ev, rhs, stream);
}

}  // extern "C"
