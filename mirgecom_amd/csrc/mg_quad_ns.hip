// Overintegrated Navier-Stokes gradients (include/mgdg.h):
// mg_ns_gradient_quad.
#include "mg_quad_ns.h"

extern "C" {

mg_status mg_ns_gradient_quad(const mg_mesh* mesh, const mg_gas* gas, const mg_quad* quad,
                              const double* q, const double* temperature,
                              const double* q_bnd_quad, const double* t_bnd_quad,
                              int32_t niter, double* grad, void* stream) {
  mg_status s = quad_ns_check(mesh, gas, quad, q, temperature, q_bnd_quad, t_bnd_quad,
                              niter, grad);
  if (s) return s;
  const int var = gas_variant(gas);
  const int nfld = 2 + mesh->dim + gas->nspecies, ns = nfld + 1;
  const int gs = bank_stride(quad->nq);
  const int fs = bank_stride(mesh->nfaces * quad->nfq);
  const size_t nfn = size_t(mesh->nfaces) * mesh->nfp;
  const size_t per_elem =
      sizeof(double) * (size_t(nfld) * (mesh->np + nfn) + size_t(ns) * (gs + fs) +
                        size_t(ns) * mesh->dim * mesh->np +
                        (var != 0 ? mesh->np + nfn : 0));
  if (per_elem > kQuadBudget) return fail(MG_ERR_UNSUPPORTED, "quadrature too large for LDS");
  mgdg::KArgs a = make_args(mesh, gas);
  a.q = q;
  a.out = grad;
  mgdg::QuadArgs qa{};
  quad_args(qa, mesh, quad, temperature, q_bnd_quad, t_bnd_quad, niter);
  qa.epb = quad_epb(per_elem);
  qa.gs = gs;
  qa.fs = fs;
  const size_t lds = per_elem * qa.epb;
  const unsigned nblk = unsigned((a.K + qa.epb - 1) / qa.epb);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define MG_GCASE(D, P, NS, MX)                                                      \
  if (mesh->dim == D && mesh->order == P && gas->nspecies == NS && var == MX) {     \
    auto kern = mgdg::ns_grad_quad_kernel<D, P, NS, MX>;                            \
    if ((s = set_lds(kern, kQuadBudget))) return s;                                 \
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(mgdg::kBlock), lds, st, a, qa);       \
    return check_launch("ns_grad_quad_kernel");                                     \
  }
  MG_NSQ_CASES(MG_GCASE)
#undef MG_GCASE
  return fail(MG_ERR_UNSUPPORTED, "dim/order/gas");
}

}  // extern "C"
