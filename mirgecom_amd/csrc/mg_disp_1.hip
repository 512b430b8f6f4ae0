// Instantiations of kernel selector 1: Navier-Stokes RHS from given gradients (rhs_kernel<kGrad>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(1)
