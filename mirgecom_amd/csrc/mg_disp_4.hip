// Instantiations of kernel selector 4: fused gradients + nodal net flux (grad_kernel<FLUX>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(4)
