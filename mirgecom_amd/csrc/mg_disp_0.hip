// Instantiations of kernel selector 0: Euler RHS (rhs_kernel<kEuler>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(0)
