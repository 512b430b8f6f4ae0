// Instantiations of kernel selector 6: mixed-precision divergence (rhs_kernel<kFlux, float>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(6)
