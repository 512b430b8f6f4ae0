// Instantiations of kernel selector 5: mixed-precision fused gradients + nodal net flux (grad_kernel<FLUX, float>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(5)
