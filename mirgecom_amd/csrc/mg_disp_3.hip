// Instantiations of kernel selector 3: DG gradients (grad_kernel<FLUX=false>), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(3)
