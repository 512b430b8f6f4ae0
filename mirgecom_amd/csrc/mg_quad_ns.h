// Shared host code of the overintegrated Navier-Stokes entry points
// (mg_quad_ns.hip, mg_quad_ns_div.hip).
#pragma once
#include <algorithm>
#include <string>

#include "mgdg_host.h"

namespace {

mg_status quad_ns_check(const mg_mesh* mesh, const mg_gas* gas, const mg_quad* quad,
                        const double* q, const double* temperature,
                        const double* q_bnd_quad, const double* t_bnd_quad, int32_t niter,
                        const double* out) {
  mg_status s = check_mesh(mesh);
  if (s) return s;
  if ((s = check_gas(gas, true))) return s;
  const bool mix = gas->eos_kind == 1;
  if (mix && (s = ensure_mech())) return s;
  if (gas->nspecies != 0 && gas->nspecies != 7)
    return fail(MG_ERR_UNSUPPORTED, "nspecies must be 0 or 7");
  if (mix && (!temperature || niter < 0 || (mesh->nbnd_nodes > 0 && !t_bnd_quad)))
    return fail(MG_ERR_ARG, "mixture needs temperature (and t_bnd_quad), niter >= 0");
  if (!quad || !quad->interp || !quad->weak_diff || !quad->face_interp || !quad->lift ||
      quad->nq < 1 || quad->nfq < 1)
    return fail(MG_ERR_ARG, "mg_quad incomplete");
  if (!q || !out) return fail(MG_ERR_ARG, "q/output NULL");
  if (mesh->nbnd_nodes > 0 && !q_bnd_quad)
    return fail(MG_ERR_ARG, "boundary exterior state missing");
  if (mesh->nbnd_nodes % mesh->nfp) return fail(MG_ERR_ARG, "nbnd_nodes % nfp");
  return MG_OK;
}

// LDS row strides = 4 mod 32 doubles (bank spread of the matrix-core reads)
int bank_stride(int n) { return n + ((4 - n % 32) + 32) % 32; }

// up to 80 KiB per workgroup (at least one element, at most 128 KiB)
constexpr size_t kQuadBudget = 128 * 1024;

int quad_epb(size_t per_elem) {
  int epb = int((80 * 1024) / per_elem);
  return std::max(1, std::min(epb, 16));
}

void quad_args(mgdg::QuadArgs& qa, const mg_mesh* mesh, const mg_quad* quad,
               const double* temperature, const double* q_bnd_quad,
               const double* t_bnd_quad, int32_t niter) {
  qa.nq = quad->nq;
  qa.nfq = quad->nfq;
  qa.interp = quad->interp;
  qa.wdiff = quad->weak_diff;
  qa.finterp = quad->face_interp;
  qa.lift = quad->lift;
  qa.qbq = q_bnd_quad;
  qa.nbq = mesh->nbnd_nodes / mesh->nfp * quad->nfq;
  qa.temp = temperature;
  qa.tbq = t_bnd_quad;
  qa.niter = niter;
}

}  // namespace

#define MG_NSQ_CASES(X)                                                              \
  X(2, 1, 0, 0) X(2, 2, 0, 0) X(2, 3, 0, 0) X(2, 4, 0, 0)                           \
  X(3, 1, 0, 0) X(3, 2, 0, 0) X(3, 3, 0, 0) X(3, 4, 0, 0)                           \
  X(2, 1, 7, 0) X(2, 2, 7, 0) X(2, 3, 7, 0) X(2, 4, 7, 0)                           \
  X(3, 1, 7, 0) X(3, 2, 7, 0) X(3, 3, 7, 0) X(3, 4, 7, 0)                           \
  X(2, 1, 7, 1) X(2, 2, 7, 1) X(2, 3, 7, 1) X(2, 4, 7, 1)                           \
  X(3, 1, 7, 1) X(3, 2, 7, 1) X(3, 3, 7, 1) X(3, 4, 7, 1)                           \
  X(2, 1, 7, 2) X(2, 2, 7, 2) X(2, 3, 7, 2) X(2, 4, 7, 2)                           \
  X(3, 1, 7, 2) X(3, 2, 7, 2) X(3, 3, 7, 2) X(3, 4, 7, 2)

