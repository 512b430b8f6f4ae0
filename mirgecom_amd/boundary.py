"""Fluid boundary conditions for the GPU operators.

Call surface of mirgecom/boundary.py:216-388 (``FluidBoundary``),
:796-1009 (``PrescribedFluidBoundary``) and :1012-1049 (``DummyBoundary``).

On the device every boundary node carries an exterior conserved state (and,
for mixtures, an exterior temperature).  The HIP kernels then apply the
reference's boundary treatment to it: the inviscid numerical flux between
interior and exterior states, the exterior value as the gradient flux
(boundary.py:70-72, 895-924) and the central viscous flux with identical
gradients on both sides (boundary.py:944-970).  ``DummyBoundary`` is the
exterior state equal to the interior one, for which these reduce to
``F(q-).n``, ``q- n`` and ``F_V(q-).n`` (boundary.py:1024-1049).
"""

from __future__ import annotations

from warnings import warn


class FluidBoundary:
    """Base class; subclasses provide ``exterior_state``."""

    def exterior_state(self, dcoll, dd_bdry, gas_model, state_minus, time=0.0):
        raise NotImplementedError


class PrescribedFluidBoundary(FluidBoundary):
    """Boundary with a user ``boundary_state_func(dcoll, dd_bdry, gas_model,
    state_minus, time=...) -> FluidState`` evaluated on the boundary nodes.

    The other hooks of the reference class (custom inviscid/viscous/gradient
    flux functions, boundary temperature or gradient functions) change the
    flux formulas themselves; they are not available on the device path and
    raise ``NotImplementedError``."""

    def __init__(self, inviscid_flux_func=None, boundary_state_func=None,
                 temperature_gradient_flux_func=None,
                 boundary_temperature_func=None, cv_gradient_flux_func=None,
                 gradient_numerical_flux_func=None, viscous_flux_func=None,
                 boundary_gradient_cv_func=None,
                 boundary_gradient_temperature_func=None):
        custom = [inviscid_flux_func, temperature_gradient_flux_func,
                  boundary_temperature_func, cv_gradient_flux_func,
                  gradient_numerical_flux_func, viscous_flux_func,
                  boundary_gradient_cv_func, boundary_gradient_temperature_func]
        if any(f is not None for f in custom):
            raise NotImplementedError(
                "only boundary_state_func is supported on the device path")
        if boundary_state_func is None:
            warn("Using dummy boundary: copies interior solution.", stacklevel=2)
        self._bnd_state_func = boundary_state_func

    def exterior_state(self, dcoll, dd_bdry, gas_model, state_minus, time=0.0):
        if self._bnd_state_func is None:
            return state_minus
        return self._bnd_state_func(dcoll=dcoll, dd_bdry=dd_bdry,
                                    gas_model=gas_model,
                                    state_minus=state_minus, time=time)


class DummyBoundary(FluidBoundary):
    """Exterior state = interior state (mirgecom/boundary.py:1012-1049)."""

    def __init__(self):
        warn("Using dummy boundary: copies interior solution.", stacklevel=2)

    def exterior_state(self, dcoll, dd_bdry, gas_model, state_minus, time=0.0):
        return state_minus
