// Instantiations of kernel selector 2: Navier-Stokes divergence of nodal net fluxes (rhs_kernel<kFlux>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(2)
