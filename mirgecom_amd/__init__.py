"""MI355X-native DG right-hand side for compressible (reacting) flow.

A drop-in for the right-hand-side evaluation of mirgecom's
``euler_operator`` (mirgecom/euler.py:313-416) and ``ns_operator``
(mirgecom/navierstokes.py:350-575): the same call surface -- discretization,
``ConservedVars`` state, gas model, boundaries, time in, ``ConservedVars`` RHS
out -- with the work done by hand-written HIP kernels for gfx950 behind the C
ABI of ``include/mgdg.h`` (``libmgdg.so``, bound with ctypes).

Modules mirror the reference's: ``discretization``, ``fluid``, ``eos``,
``transport``, ``gas_model``, ``boundary``, ``inviscid``, ``viscous``,
``flux``, ``euler``, ``navierstokes``, ``integrators``, ``initializers``.
"""

__version__ = "0.1.0"
