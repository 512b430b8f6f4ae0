"""CPU FP64 oracle for the DG right-hand side -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy FP64, the algorithm of the reference's
``euler_operator`` (mirgecom/euler.py:313-416) and ``ns_operator``
(mirgecom/navierstokes.py:350-575) together with the pieces they call:
fluxes (mirgecom/inviscid.py, mirgecom/viscous.py, mirgecom/flux.py), the
gas model (mirgecom/gas_model.py, mirgecom/eos.py, mirgecom/transport.py),
the uiuc_7sp thermochemistry (mirgecom/thermochemistry.py,
mirgecom/mechanisms/uiuc_7sp.yaml) and RK4 (mirgecom/integrators/
explicit_rk.py:31-38).

It is the checker, never the product: only ``tests/``, the smoke entry point
and the CPU-baseline leg of the benchmark may import it.  The product path
(``mirgecom_amd``) never imports anything from here.

To be an independent check of the product's discretization, the oracle builds
its own element matrices (Lagrange basis through monomials, integrated with
collapsed-coordinate Gauss-Jacobi quadrature exact to degree 2p), its own
affine geometry and its own face connectivity (matching face-node
coordinates, with periodic wrap).  Only the mesh vertices/elements and the
reference node coordinates are taken as inputs.

Parity status per function is listed in the module docstrings: the reference
itself cannot be imported in this environment (grudge/meshmode/pyopencl/
pyrometheus/cantera absent), so every function is pinned by closed forms,
invariants and convergence rates in ``tests/``.
"""
