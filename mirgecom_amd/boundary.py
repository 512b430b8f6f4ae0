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


# {{{ Mengaldo family (mirgecom/boundary.py:391-1857)

# per-boundary-node flags of the device boundary data (include/mgdg.h)
BC_MENGALDO = 1       # CV_bc / T_bc gradient values, F_V(bc) . n viscous flux
BC_SLIP = 2           # slip velocity gradient (boundary.py:159-173)
BC_IMPERMEABLE = 4    # no normal species-fraction gradient (193-213)
BC_ADIABATIC = 8      # no normal temperature gradient (186-190)


def _replace_state(state, gas_model, energy=None, momentum=None):
    """``replace_fluid_state`` (mirgecom/gas_model.py:905-995): new conserved
    fields, temperature seeded with the old one."""
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import make_fluid_state
    cv = state.cv
    q = cv.q.clone()
    if energy is not None:
        q[1] = energy
    if momentum is not None:
        q[2:2 + cv.dim] = momentum
    return make_fluid_state(ConservedVars(cv.dim, q), gas_model,
                            temperature_seed=state.temperature)


def _slip_momentum_plus(mom, n):
    return mom - 2.0 * (mom * n).sum(dim=0) * n


def _slip_momentum_bc(mom, n):
    return mom - (mom * n).sum(dim=0) * n


class MengaldoBoundaryCondition(FluidBoundary):
    """mirgecom/boundary.py:391-793: the inviscid flux takes ``state_plus``,
    the gradient fluxes ``CV_bc n`` and ``T_bc n``, the viscous flux
    ``F_V(state_bc, grad_cv_bc, grad_t_bc) . n``.

    ``state_plus`` / ``state_bc`` / ``temperature_bc`` are evaluated here on
    the boundary nodes; the gradient conditions, which need the interior
    gradients, run inside the HIP kernels, selected by ``device_flags``."""

    device_flags = BC_MENGALDO

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        raise NotImplementedError

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        raise NotImplementedError

    def temperature_bc(self, dcoll, dd_bdry, state_minus, **kwargs):
        return state_minus.temperature

    def exterior_state(self, dcoll, dd_bdry, gas_model, state_minus, time=0.0):
        return self.state_plus(dcoll, dd_bdry, gas_model, state_minus,
                               time=time)


class AdiabaticSlipBoundary(MengaldoBoundaryCondition):
    """mirgecom/boundary.py:1052-1143"""

    device_flags = BC_MENGALDO | BC_SLIP | BC_IMPERMEABLE | BC_ADIABATIC

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        n = dcoll.normal(dd_bdry)
        return _replace_state(state_minus, gas_model, momentum=_slip_momentum_plus(
            state_minus.momentum_density, n))

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        n = dcoll.normal(dd_bdry)
        mom_bc = _slip_momentum_bc(state_minus.momentum_density, n)
        energy = (gas_model.eos.internal_energy(state_minus.cv)
                  + 0.5 * (mom_bc * mom_bc).sum(dim=0) / state_minus.mass_density)
        return _replace_state(state_minus, gas_model, energy=energy,
                              momentum=mom_bc)


class FarfieldBoundary(MengaldoBoundaryCondition):
    """mirgecom/boundary.py:1146-1228"""

    def __init__(self, free_stream_pressure, free_stream_velocity,
                 free_stream_temperature, free_stream_mass_fractions=None):
        self._pressure = free_stream_pressure
        self._velocity = free_stream_velocity
        self._temperature = free_stream_temperature
        self._species_mass_fractions = free_stream_mass_fractions

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        import torch

        from mirgecom_amd.fluid import make_conserved
        from mirgecom_amd.gas_model import make_fluid_state
        rho_m = state_minus.mass_density
        one = torch.ones_like(rho_m)
        nsp = state_minus.cv.nspecies
        y = one.new_zeros((0,) + tuple(one.shape))
        if nsp > 0:
            y = torch.as_tensor(self._species_mass_fractions, dtype=one.dtype,
                                device=one.device).reshape(-1, *([1] * one.dim())) * one
        t = self._temperature * one
        p = self._pressure * one
        v = torch.as_tensor(self._velocity, dtype=one.dtype,
                            device=one.device).reshape(-1, *([1] * one.dim())) * one
        eos = gas_model.eos
        rho = eos.get_density(p, t, y if nsp > 0 else None)
        eint = eos.get_internal_energy(t, y if nsp > 0 else None)
        energy = rho * (eint + 0.5 * (v * v).sum(dim=0))
        cv = make_conserved(state_minus.cv.dim, mass=rho, energy=energy,
                            momentum=rho * v,
                            species_mass=rho * y if nsp > 0 else None)
        return make_fluid_state(cv, gas_model, temperature_seed=t)

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        return self.state_plus(dcoll, dd_bdry, gas_model, state_minus, **kwargs)

    def temperature_bc(self, dcoll, dd_bdry, state_minus, **kwargs):
        import torch
        return torch.zeros_like(state_minus.temperature) + self._temperature


class PressureOutflowBoundary(MengaldoBoundaryCondition):
    """mirgecom/boundary.py:1231-1405"""

    def __init__(self, boundary_pressure=101325):
        self._pressure = boundary_pressure

    def _outflow_state(self, dcoll, dd_bdry, gas_model, state_minus, p_plus):
        import torch

        from mirgecom_amd.fluid import ConservedVars
        from mirgecom_amd.gas_model import make_fluid_state
        n = dcoll.normal(dd_bdry)
        sm = state_minus
        vn = (sm.velocity * n).sum(dim=0)
        bvel = vn * n
        bspeed = torch.sqrt((bvel * bvel).sum(dim=0))
        supersonic = bspeed > sm.speed_of_sound
        eos = gas_model.eos
        ke = eos.kinetic_energy(sm.cv)
        if sm.is_mixture:
            y = sm.species_mass_fractions
            rmix = eos.gas_const(species_mass_fractions=y)
            t_plus = torch.where(supersonic, sm.temperature,
                                 p_plus / (sm.mass_density * rmix))
            eint = sm.mass_density * eos.get_internal_energy(t_plus, y)
        else:
            gamma = eos.gamma(sm.cv, sm.temperature)
            p_b = torch.where(supersonic, sm.pressure, p_plus)
            eint = p_b / (gamma - 1.0)
        q = sm.cv.q.clone()
        q[1] = eint + ke
        return make_fluid_state(ConservedVars(sm.cv.dim, q), gas_model,
                                temperature_seed=sm.temperature)

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        import torch
        pm = state_minus.pressure
        p_plus = torch.where(pm > 2.0 * self._pressure, pm,
                             2.0 * self._pressure - pm)
        return self._outflow_state(dcoll, dd_bdry, gas_model, state_minus,
                                   p_plus)

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        import torch
        p_b = self._pressure + torch.zeros_like(state_minus.pressure)
        return self._outflow_state(dcoll, dd_bdry, gas_model, state_minus, p_b)


class IsothermalSlipWallBoundary(MengaldoBoundaryCondition):
    """mirgecom/boundary.py:1632-1724"""

    device_flags = BC_MENGALDO | BC_SLIP | BC_IMPERMEABLE

    def __init__(self, wall_temperature=300):
        self._wall_temp = wall_temperature

    def temperature_bc(self, dcoll, dd_bdry, state_minus, **kwargs):
        import torch
        return torch.zeros_like(state_minus.temperature) + self._wall_temp

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        n = dcoll.normal(dd_bdry)
        sm = state_minus
        mom_bc = _slip_momentum_bc(sm.momentum_density, n)
        t_bc = self.temperature_bc(dcoll, dd_bdry, sm)
        eint = gas_model.eos.get_internal_energy(
            t_bc, sm.species_mass_fractions if sm.cv.nspecies else None)
        energy = (sm.mass_density * eint
                  + 0.5 * (mom_bc * mom_bc).sum(dim=0) / sm.mass_density)
        return _replace_state(sm, gas_model, energy=energy, momentum=mom_bc)

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        n = dcoll.normal(dd_bdry)
        return _replace_state(state_minus, gas_model, momentum=_slip_momentum_plus(
            state_minus.momentum_density, n))


class IsothermalWallBoundary(MengaldoBoundaryCondition):
    """mirgecom/boundary.py:1727-1791"""

    device_flags = BC_MENGALDO | BC_IMPERMEABLE

    def __init__(self, wall_temperature=300):
        self._wall_temp = wall_temperature

    def temperature_bc(self, dcoll, dd_bdry, state_minus, **kwargs):
        import torch
        return torch.zeros_like(state_minus.temperature) + self._wall_temp

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        sm = state_minus
        t_bc = self.temperature_bc(dcoll, dd_bdry, sm)
        eint = gas_model.eos.get_internal_energy(
            t_bc, sm.species_mass_fractions if sm.cv.nspecies else None)
        return _replace_state(sm, gas_model, energy=sm.mass_density * eint,
                              momentum=0.0 * sm.momentum_density)

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        return _replace_state(state_minus, gas_model,
                              momentum=-state_minus.momentum_density)


class AdiabaticNoslipWallBoundary(MengaldoBoundaryCondition):
    """mirgecom/boundary.py:1794-1857"""

    device_flags = BC_MENGALDO | BC_IMPERMEABLE | BC_ADIABATIC

    def state_plus(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        return _replace_state(state_minus, gas_model,
                              momentum=-state_minus.momentum_density)

    def state_bc(self, dcoll, dd_bdry, gas_model, state_minus, **kwargs):
        sm = state_minus
        eint = gas_model.eos.get_internal_energy(
            sm.temperature, sm.species_mass_fractions if sm.cv.nspecies else None)
        return _replace_state(sm, gas_model, energy=sm.mass_density * eint,
                              momentum=0.0 * sm.momentum_density)

# }}}
