// Instantiations of kernel selector 7: mixture nodal net flux from stored gradients (node_flux_kernel), every (dim, order).
#include "mgdg_host.h"

MG_DEFINE_DISPATCH(7)
