// Instantiations of kernel selector 8: mixed-precision Euler RHS (rhs_kernel<kEuler, float>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(8)
