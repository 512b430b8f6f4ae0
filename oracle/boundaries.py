"""Oracle restatement of the Mengaldo boundary family (test infrastructure:
only tests/ and the benchmark's CPU baseline use this package).

mirgecom/boundary.py, restated in NumPy FP64:

* components: ``_SlipBoundaryComponent`` (149-173), ``_NoSlipBoundaryComponent``
  (176-183), ``_AdiabaticBoundaryComponent`` (186-190),
  ``_ImpermeableBoundaryComponent`` (193-213), ``_get_normal_axes`` /
  ``_get_rotation_matrix`` (75-129)
* ``MengaldoBoundaryCondition`` (391-793): inviscid flux = numerical flux with
  ``state_plus``; gradient fluxes ``CV_bc n`` and ``T_bc n``; viscous flux
  ``F_V(state_bc, grad_cv_bc, grad_t_bc) . n``
* ``AdiabaticSlipBoundary`` (1052-1143), ``FarfieldBoundary`` (1146-1228),
  ``PressureOutflowBoundary`` (1231-1405), ``IsothermalSlipWallBoundary``
  (1632-1724), ``IsothermalWallBoundary`` (1727-1791),
  ``AdiabaticNoslipWallBoundary`` (1794-1857)

States are ``oracle.physics.State`` on boundary nodes ``(..., n)``; normals
``(dim, n)``.  Pinned by the closed-form boundary states of
test/test_bc.py:136-972 and the wall invariants in tests/test_oracle_bc.py.
"""

from __future__ import annotations

import numpy as np

from oracle import physics as ph


# {{{ EOS helpers the boundary states need

def _is_mixture(eos):
    return hasattr(eos, "species_enthalpies")


def internal_energy_mass(eos, temperature, y):
    """``e(T, Y)`` per unit mass (mirgecom/eos.py:281-299, 697-716)."""
    if _is_mixture(eos):
        from oracle import thermochem as tc
        return tc.mixture_energy_mass(temperature, y)
    return eos.gas_const() * temperature / (eos.gamma() - 1.0)


def density(eos, pressure, temperature, y):
    """``rho(p, T, Y)`` (mirgecom/eos.py:264-279, 675-695)."""
    if _is_mixture(eos):
        from oracle import thermochem as tc
        return tc.density(pressure, temperature, y)
    return pressure / (eos.gas_const() * temperature)


def gas_const(eos, y):
    if _is_mixture(eos):
        from oracle import thermochem as tc
        return tc.specific_gas_constant(y)
    return eos.gas_const()


def replace_state(sm: ph.State, gas_model, mass=None, energy=None,
                  momentum=None):
    """``replace_fluid_state`` (mirgecom/gas_model.py:905-995): new conserved
    fields, dependent variables recomputed with the old temperature as seed."""
    dim = sm.dim
    cv = sm.cv.copy()
    if mass is not None:
        cv[0] = mass
    if energy is not None:
        cv[1] = energy
    if momentum is not None:
        cv[2:2 + dim] = momentum
    return ph.make_fluid_state(cv, gas_model, dim,
                               temperature_seed=sm.temperature)

# }}}


# {{{ components

def normal_axes(n):
    """mirgecom/boundary.py:75-115: an orthonormal set whose first member is
    the unit normal (2-D: its rotation; 3-D: Householder-like construction,
    undefined where n_z == 0)."""
    dim = n.shape[0]
    mag = np.sqrt(np.sum(n * n, axis=0))
    n = n / mag
    if dim == 2:
        return n, np.stack([-n[1], n[0]])
    x, y, z = n
    zsign = z / np.abs(z)
    a = mag * zsign
    b = z + a
    v2 = np.stack([a * b - x * x, -x * y, -x * b])
    v2 = v2 / np.sqrt(np.sum(v2 * v2, axis=0))
    v3 = np.stack([y * v2[2] - v2[1] * z, v2[0] * z - x * v2[2],
                   x * v2[1] - y * v2[0]])
    return n, v2, v3


def rotation_matrix(n):
    """mirgecom/boundary.py:118-129: rows are the normal axes; (dim, dim, n)."""
    return np.stack(normal_axes(n))


def slip_grad_velocity_bc_rotated(state_minus, grad_cv_minus, n):
    """mirgecom/boundary.py:159-173 literally: rotate grad v into the normal
    frame, zero the normal derivative of the tangential velocity, rotate
    back.  Undefined in 3-D where n_z == 0 (the frame divides by |n_z|)."""
    dim = state_minus.dim
    gv = ph.velocity_gradient(state_minus.cv, grad_cv_minus, dim)
    rot = rotation_matrix(n)
    gn = np.einsum("ia...,ab...,jb...->ij...", rot, gv, rot)
    for i in range(dim - 1):
        gn[i + 1, 0] = 0.0 * gn[i + 1, 0]
    return np.einsum("ai...,ab...,bj...->ij...", rot, gn, rot)


def slip_grad_velocity_bc(state_minus, grad_cv_minus, n):
    """The same boundary velocity gradient in frame-free form,
    ``G - (G n) n^T + n (n^T G n) n^T``: zeroing the tangential rows of the
    normal column in any orthonormal frame ``(n, t_i)`` removes
    ``sum_i t_i (t_i^T G n) n^T = (I - n n^T) G n n^T``.  It equals the
    rotated form wherever that is defined and stays defined on faces with
    ``n_z == 0``, which axis-aligned box walls have."""
    dim = state_minus.dim
    gv = ph.velocity_gradient(state_minus.cv, grad_cv_minus, dim)
    gn = np.einsum("ij...,j...->i...", gv, n)
    ngn = np.einsum("i...,i...->...", n, gn)
    return gv - gn[:, None] * n[None] + n[:, None] * n[None] * ngn


def impermeable_grad_species_mass_bc(state_minus, grad_cv_minus, n):
    """mirgecom/boundary.py:196-213"""
    dim = state_minus.dim
    gsm = 1.0 * grad_cv_minus[2 + dim:]
    nsp = gsm.shape[0]
    if nsp == 0:
        return gsm
    gy = ph.species_mass_fraction_gradient(state_minus.cv, grad_cv_minus, dim)
    gyn = np.einsum("ki...,i...->k...", gy, n)
    gy_bc = gy - gyn[:, None] * n[None]
    y = state_minus.species_mass_fractions
    return state_minus.mass * gy_bc + y[:, None] * grad_cv_minus[0][None]


def adiabatic_grad_t_bc(grad_t_minus, n):
    """mirgecom/boundary.py:189-190"""
    return grad_t_minus - np.sum(grad_t_minus * n, axis=0) * n

# }}}


class MengaldoBoundary:
    """mirgecom/boundary.py:391-793 (the oracle's operators dispatch on
    ``kind``)."""

    kind = "mengaldo"

    def temperature_bc(self, sm):
        return sm.temperature

    def grad_cv_bc(self, sm, state_bc, grad_cv_minus, n, gas_model):
        return grad_cv_minus

    def grad_temperature_bc(self, grad_t_minus, n):
        return grad_t_minus

    def viscous_flux_n(self, sm, grad_cv_minus, grad_t_minus, n, gas_model):
        """mirgecom/boundary.py:630-710: ``F_V(bc) . n``."""
        sbc = self.state_bc(sm, n, gas_model)
        gcv = self.grad_cv_bc(sm, sbc, grad_cv_minus, n, gas_model)
        gt = self.grad_temperature_bc(grad_t_minus, n)
        return ph._dot_normal(ph.viscous_flux(sbc, gcv, gt), n)


class AdiabaticSlipBoundary(MengaldoBoundary):
    """mirgecom/boundary.py:1052-1143"""

    def state_plus(self, sm, n, gas_model):
        mom = sm.momentum
        mn = np.sum(mom * n, axis=0)
        return replace_state(sm, gas_model, momentum=mom - 2.0 * mn * n)

    def state_bc(self, sm, n, gas_model):
        mom = sm.momentum
        mom_bc = mom - np.sum(mom * n, axis=0) * n
        eint = sm.energy - 0.5 * np.sum(mom * mom, axis=0) / sm.mass
        energy = eint + 0.5 * np.sum(mom_bc * mom_bc, axis=0) / sm.mass
        return replace_state(sm, gas_model, energy=energy, momentum=mom_bc)

    def grad_temperature_bc(self, grad_t_minus, n):
        return adiabatic_grad_t_bc(grad_t_minus, n)

    def grad_cv_bc(self, sm, state_bc, grad_cv_minus, n, gas_model):
        dim = sm.dim
        gv_bc = slip_grad_velocity_bc(sm, grad_cv_minus, n)
        out = grad_cv_minus.copy()
        out[2:2 + dim] = (state_bc.mass * gv_bc
                          + state_bc.velocity[:, None] * grad_cv_minus[0][None])
        out[2 + dim:] = impermeable_grad_species_mass_bc(sm, grad_cv_minus, n)
        return out


class FarfieldBoundary(MengaldoBoundary):
    """mirgecom/boundary.py:1146-1228"""

    def __init__(self, free_stream_pressure, free_stream_velocity,
                 free_stream_temperature, free_stream_mass_fractions=None):
        self.p = free_stream_pressure
        self.v = np.asarray(free_stream_velocity, dtype=np.float64)
        self.t = free_stream_temperature
        self.y = (None if free_stream_mass_fractions is None
                  else np.asarray(free_stream_mass_fractions, dtype=np.float64))

    def state_plus(self, sm, n, gas_model):
        eos = gas_model.eos
        one = 0.0 * sm.mass + 1.0
        nsp = sm.cv.shape[0] - 2 - sm.dim
        y = (np.zeros((0,) + one.shape) if nsp == 0
             else self.y.reshape((-1,) + (1,) * one.ndim) * one)
        t = self.t * one
        p = self.p * one
        v = self.v.reshape((-1,) + (1,) * one.ndim) * one
        rho = density(eos, p, t, y)
        eint = internal_energy_mass(eos, t, y)
        energy = rho * (eint + 0.5 * np.sum(v * v, axis=0))
        cv = np.concatenate([rho[None], energy[None], rho * v, rho * y])
        return ph.make_fluid_state(cv, gas_model, sm.dim, temperature_seed=t)

    def state_bc(self, sm, n, gas_model):
        return self.state_plus(sm, n, gas_model)

    def temperature_bc(self, sm):
        return 0.0 * sm.temperature + self.t


class PressureOutflowBoundary(MengaldoBoundary):
    """mirgecom/boundary.py:1231-1405"""

    def __init__(self, boundary_pressure=101325):
        self.p = boundary_pressure

    def _state(self, sm, n, gas_model, pressure_plus):
        eos = gas_model.eos
        vn = np.sum(sm.velocity * n, axis=0)
        bspeed = np.sqrt(np.sum((vn * n) ** 2, axis=0))
        ke = 0.5 * np.sum(sm.momentum ** 2, axis=0) / sm.mass
        supersonic = bspeed > sm.speed_of_sound
        if sm.is_mixture:
            y = sm.species_mass_fractions
            t_plus = np.where(supersonic, sm.temperature,
                              pressure_plus / (sm.mass * gas_const(eos, y)))
            eint = sm.mass * internal_energy_mass(eos, t_plus, y)
        else:
            gamma = eos.gamma(sm.cv, sm.temperature)
            p_b = np.where(supersonic, sm.pressure, pressure_plus)
            eint = p_b / (gamma - 1.0)
        cv = sm.cv.copy()
        cv[1] = eint + ke
        return ph.make_fluid_state(cv, gas_model, sm.dim,
                                   temperature_seed=sm.temperature)

    def state_plus(self, sm, n, gas_model):
        p_plus = np.where(sm.pressure > 2.0 * self.p, sm.pressure,
                          2.0 * self.p - sm.pressure)
        return self._state(sm, n, gas_model, p_plus)

    def state_bc(self, sm, n, gas_model):
        return self._state(sm, n, gas_model, self.p + 0.0 * sm.pressure)


class IsothermalSlipWallBoundary(MengaldoBoundary):
    """mirgecom/boundary.py:1632-1724"""

    def __init__(self, wall_temperature=300):
        self.tw = wall_temperature

    def temperature_bc(self, sm):
        return 0.0 * sm.temperature + self.tw

    def state_bc(self, sm, n, gas_model):
        mom = sm.momentum
        mom_bc = mom - np.sum(mom * n, axis=0) * n
        eint = internal_energy_mass(gas_model.eos, self.temperature_bc(sm),
                                    sm.species_mass_fractions)
        energy = sm.mass * eint + 0.5 * np.sum(mom_bc * mom_bc, axis=0) / sm.mass
        return replace_state(sm, gas_model, energy=energy, momentum=mom_bc)

    def grad_cv_bc(self, sm, state_bc, grad_cv_minus, n, gas_model):
        return AdiabaticSlipBoundary.grad_cv_bc(self, sm, state_bc,
                                                grad_cv_minus, n, gas_model)

    def state_plus(self, sm, n, gas_model):
        return AdiabaticSlipBoundary.state_plus(self, sm, n, gas_model)


class IsothermalWallBoundary(MengaldoBoundary):
    """mirgecom/boundary.py:1727-1791"""

    def __init__(self, wall_temperature=300):
        self.tw = wall_temperature

    def temperature_bc(self, sm):
        return 0.0 * sm.temperature + self.tw

    def state_bc(self, sm, n, gas_model):
        eint = internal_energy_mass(gas_model.eos, self.temperature_bc(sm),
                                    sm.species_mass_fractions)
        return replace_state(sm, gas_model, energy=sm.mass * eint,
                             momentum=0.0 * sm.momentum)

    def grad_cv_bc(self, sm, state_bc, grad_cv_minus, n, gas_model):
        out = grad_cv_minus.copy()
        out[2 + sm.dim:] = impermeable_grad_species_mass_bc(sm, grad_cv_minus, n)
        return out

    def state_plus(self, sm, n, gas_model):
        return replace_state(sm, gas_model, momentum=-sm.momentum)


class AdiabaticNoslipWallBoundary(MengaldoBoundary):
    """mirgecom/boundary.py:1794-1857"""

    def state_plus(self, sm, n, gas_model):
        return replace_state(sm, gas_model, momentum=-sm.momentum)

    def state_bc(self, sm, n, gas_model):
        eint = internal_energy_mass(gas_model.eos, sm.temperature,
                                    sm.species_mass_fractions)
        return replace_state(sm, gas_model, energy=sm.mass * eint,
                             momentum=0.0 * sm.momentum)

    def grad_cv_bc(self, sm, state_bc, grad_cv_minus, n, gas_model):
        out = grad_cv_minus.copy()
        out[2 + sm.dim:] = impermeable_grad_species_mass_bc(sm, grad_cv_minus, n)
        return out

    def grad_temperature_bc(self, grad_t_minus, n):
        return adiabatic_grad_t_bc(grad_t_minus, n)
