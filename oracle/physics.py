"""Oracle pointwise physics: EOS, transport, fluid states and fluxes.

Restated from the reference (NumPy FP64, arrays broadcast over nodes):

* ``IdealSingleGas`` -- mirgecom/eos.py:302-537 (p 429-444, c 446-462,
  T 464-483)
* ``SimpleTransport`` / ``PowerLawTransport`` -- mirgecom/transport.py:143-321
* ``make_fluid_state`` -- mirgecom/gas_model.py:302-431
* ``inviscid_flux`` -- mirgecom/inviscid.py:65-102
* Rusanov / central / HLL facial fluxes -- mirgecom/inviscid.py:105-287 with
  ``num_flux_lfr``/``num_flux_central``/``num_flux_hll`` mirgecom/flux.py:45-177
* ``velocity_gradient``, ``species_mass_fraction_gradient`` --
  mirgecom/fluid.py:355-427
* ``viscous_flux`` and its parts -- mirgecom/viscous.py:72-293
* ``viscous_facial_flux_central`` -- mirgecom/viscous.py:296-349

Conserved-variable arrays have the field axis first, in the reference order
``[mass, energy, momentum_0..d-1, species_0..ns-1]``
(mirgecom/fluid.py:311-333).

Parity: closed-form known-answer tests (tests/test_oracle_flux.py) follow the
reference's own tests test/test_flux.py:57-407, test/test_inviscid.py:64-273,
test/test_viscous.py:63-355 and test/test_eos.py:528-624.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Optional

import numpy as np


# {{{ gas models

class IdealSingleGas:
    def __init__(self, gamma=1.4, gas_const=287.1):
        self.gamma_ = gamma
        self.gas_const_ = gas_const

    def gamma(self, cv=None, temperature=None):
        return self.gamma_

    def gas_const(self, cv=None):
        return self.gas_const_

    def heat_capacity_cv(self, cv=None, temperature=None):
        return self.gas_const_ / (self.gamma_ - 1)

    def heat_capacity_cp(self, cv=None, temperature=None):
        return self.gas_const_ * self.gamma_ / (self.gamma_ - 1)

    def kinetic_energy(self, cv, dim):
        mom = cv[2:2 + dim]
        return 0.5 * np.sum(mom * mom, axis=0) / cv[0]

    def internal_energy(self, cv, dim):
        return cv[1] - self.kinetic_energy(cv, dim)

    def pressure(self, cv, dim, temperature=None):
        return self.internal_energy(cv, dim) * (self.gamma_ - 1.0)

    def sound_speed(self, cv, dim, temperature=None):
        return np.sqrt(self.gamma_ / cv[0] * self.pressure(cv, dim))

    def temperature(self, cv, dim, temperature_seed=None):
        return (((self.gamma_ - 1.0) / self.gas_const_)
                * self.internal_energy(cv, dim) / cv[0])


class SimpleTransport:
    def __init__(self, bulk_viscosity=0.0, viscosity=0.0,
                 thermal_conductivity=0.0, species_diffusivity=None):
        self.mu_b = bulk_viscosity
        self.mu = viscosity
        self.kappa = thermal_conductivity
        self.d = (np.zeros(0) if species_diffusivity is None
                  else np.asarray(species_diffusivity, dtype=np.float64))

    def transport_vars(self, cv, temperature, eos, dim):
        one = 0 * cv[0] + 1.0
        nsp = cv.shape[0] - 2 - dim
        d = self.d if len(self.d) == nsp else np.zeros(nsp)
        return TransportVars(
            bulk_viscosity=self.mu_b * one, viscosity=self.mu * one,
            thermal_conductivity=self.kappa * one,
            species_diffusivity=d.reshape((-1,) + (1,) * cv[0].ndim) * one)


class PowerLawTransport:
    def __init__(self, alpha=0.6, beta=4.093e-7, sigma=2.5, n=.666,
                 species_diffusivity=None, lewis=None):
        self.alpha, self.beta, self.sigma, self.n = alpha, beta, sigma, n
        self.d = species_diffusivity
        self.lewis = lewis

    def transport_vars(self, cv, temperature, eos, dim):
        mu = self.beta * temperature**self.n
        kappa = self.sigma * mu * eos.heat_capacity_cv(cv, temperature)
        nsp = cv.shape[0] - 2 - dim
        if self.lewis is not None:
            lew = np.asarray(self.lewis).reshape((-1,) + (1,) * cv[0].ndim)
            d = self.sigma * mu / (cv[0] * lew * eos.gamma(cv, temperature))
        else:
            dd = np.zeros(nsp) if self.d is None else np.asarray(self.d)
            d = dd.reshape((-1,) + (1,) * cv[0].ndim) * (0 * cv[0] + 1.)
        return TransportVars(bulk_viscosity=self.alpha * mu, viscosity=mu,
                             thermal_conductivity=kappa, species_diffusivity=d)


@dataclass
class TransportVars:
    bulk_viscosity: Any
    viscosity: Any
    thermal_conductivity: Any
    species_diffusivity: Any


@dataclass
class State:
    """A model-consistent fluid state (cv + dependent + transport vars)."""

    cv: np.ndarray
    dim: int
    temperature: np.ndarray
    pressure: np.ndarray
    speed_of_sound: np.ndarray
    species_enthalpies: Optional[np.ndarray] = None
    tv: Optional[TransportVars] = None

    @property
    def mass(self):
        return self.cv[0]

    @property
    def energy(self):
        return self.cv[1]

    @property
    def momentum(self):
        return self.cv[2:2 + self.dim]

    @property
    def species_mass(self):
        return self.cv[2 + self.dim:]

    @property
    def velocity(self):
        return self.momentum / self.mass

    @property
    def species_mass_fractions(self):
        return self.species_mass / self.mass

    @property
    def speed(self):
        v = self.velocity
        return np.sqrt(np.sum(v * v, axis=0))

    @property
    def wavespeed(self):
        return self.speed + self.speed_of_sound

    @property
    def is_mixture(self):
        return self.species_enthalpies is not None


class GasModel:
    def __init__(self, eos, transport=None):
        self.eos = eos
        self.transport = transport


def make_fluid_state(cv, gas_model, dim, temperature_seed=None):
    """mirgecom/gas_model.py:302-431 (no limiter, no smoothness fields)."""
    eos = gas_model.eos
    temperature = eos.temperature(cv, dim, temperature_seed=temperature_seed)
    pressure = eos.pressure(cv, dim, temperature)
    c = eos.sound_speed(cv, dim, temperature)
    hk = None
    if hasattr(eos, "species_enthalpies"):
        hk = eos.species_enthalpies(cv, temperature)
    tv = None
    if gas_model.transport is not None:
        tv = gas_model.transport.transport_vars(cv, temperature, eos, dim)
    return State(cv=cv, dim=dim, temperature=temperature, pressure=pressure,
                 speed_of_sound=c, species_enthalpies=hk, tv=tv)

# }}}


# {{{ inviscid

def inviscid_flux(state: State):
    """mirgecom/inviscid.py:65-102; returns ``(F, dim, ...)``."""
    dim = state.dim
    v = state.velocity
    p = state.pressure
    mass_flux = state.momentum
    energy_flux = v * (state.energy + p)
    mom_flux = np.stack([
        np.stack([state.mass * v[i] * v[j] + (p if i == j else 0 * p)
                  for j in range(dim)]) for i in range(dim)])
    rows = [mass_flux, energy_flux] + [mom_flux[i] for i in range(dim)]
    ys = state.species_mass
    for k in range(ys.shape[0]):
        rows.append(v * ys[k])
    return np.stack(rows)


def _dot_normal(flux, normal):
    """``flux @ normal`` over the dim axis (axis 1)."""
    return np.einsum("fi...,i...->f...", flux, normal)


def num_flux_lfr(f_minus_normal, f_plus_normal, q_minus, q_plus, lam):
    """mirgecom/flux.py:45-84"""
    return (f_minus_normal + f_plus_normal + lam * (q_minus - q_plus)) / 2


def num_flux_central(f_minus_normal, f_plus_normal):
    """mirgecom/flux.py:87-110"""
    return (f_plus_normal + f_minus_normal) / 2


def num_flux_hll(f_minus_normal, f_plus_normal, q_minus, q_plus, s_minus,
                 s_plus):
    """mirgecom/flux.py:113-177"""
    f_star = (s_plus * f_minus_normal - s_minus * f_plus_normal
              + s_plus * s_minus * (q_plus - q_minus)) / (s_plus - s_minus)
    f = f_star
    f = np.where(s_minus >= 0, f_minus_normal, f)
    f = np.where(s_plus <= 0, f_plus_normal, f)
    return f


def inviscid_facial_flux_rusanov(s_int: State, s_ext: State, gas_model, normal):
    """mirgecom/inviscid.py:105-152"""
    lam = np.maximum(s_int.wavespeed, s_ext.wavespeed)
    return num_flux_lfr(_dot_normal(inviscid_flux(s_int), normal),
                        _dot_normal(inviscid_flux(s_ext), normal),
                        s_int.cv, s_ext.cv, lam)


def inviscid_facial_flux_central(s_int: State, s_ext: State, gas_model, normal):
    """mirgecom/inviscid.py:155-199"""
    return num_flux_lfr(_dot_normal(inviscid_flux(s_int), normal),
                        _dot_normal(inviscid_flux(s_ext), normal),
                        s_int.cv, s_ext.cv, 0)


def inviscid_facial_flux_hll(s_int: State, s_ext: State, gas_model, normal):
    """mirgecom/inviscid.py:202-287"""
    ones = 0. * s_int.mass + 1.
    p_int, p_ext = s_int.pressure, s_ext.pressure
    u_int = np.einsum("i...,i...->...", s_int.velocity, normal)
    u_ext = np.einsum("i...,i...->...", s_ext.velocity, normal)
    rho_int, rho_ext = s_int.mass, s_ext.mass
    c_int, c_ext = s_int.speed_of_sound, s_ext.speed_of_sound
    p_star = (0.5 * (p_int + p_ext) + (1. / 8.) * (u_int - u_ext)
              * (rho_int + rho_ext) * (c_int + c_ext))
    gamma_int = gas_model.eos.gamma(s_int.cv, s_int.temperature)
    gamma_ext = gas_model.eos.gamma(s_ext.cv, s_ext.temperature)
    q_int = 1 + (gamma_int + 1) / (2 * gamma_int) * (p_star / p_int - 1)
    q_ext = 1 + (gamma_ext + 1) / (2 * gamma_ext) * (p_star / p_ext - 1)
    q_int_rt = np.sqrt(np.where(p_star > p_int, q_int, ones))
    q_ext_rt = np.sqrt(np.where(p_star > p_ext, q_ext, ones))
    s_minus = u_int - c_int * q_int_rt
    s_plus = u_ext + c_ext * q_ext_rt
    return num_flux_hll(_dot_normal(inviscid_flux(s_int), normal),
                        _dot_normal(inviscid_flux(s_ext), normal),
                        s_int.cv, s_ext.cv, s_minus, s_plus)

# }}}


# {{{ viscous

def velocity_gradient(cv, grad_cv, dim):
    """mirgecom/fluid.py:355-397: ``(dim_v, dim_x, ...)``."""
    rho = cv[0]
    v = cv[2:2 + dim] / rho
    return (grad_cv[2:2 + dim] - v[:, None] * grad_cv[0][None]) / rho


def species_mass_fraction_gradient(cv, grad_cv, dim):
    """mirgecom/fluid.py:400-427"""
    rho = cv[0]
    y = cv[2 + dim:] / rho
    return (grad_cv[2 + dim:] - y[:, None] * grad_cv[0][None]) / rho


def viscous_stress_tensor(state: State, grad_cv):
    """mirgecom/viscous.py:68-105"""
    dim = state.dim
    mu = state.tv.viscosity
    mu_b = state.tv.bulk_viscosity
    gv = velocity_gradient(state.cv, grad_cv, dim)
    div_v = sum(gv[i, i] for i in range(dim))
    eye = np.eye(dim).reshape((dim, dim) + (1,) * div_v.ndim)
    return mu * (gv + np.swapaxes(gv, 0, 1)) + (mu_b - 2 * mu / 3) * div_v * eye


def diffusive_flux(state: State, grad_cv):
    """mirgecom/viscous.py:108-153: ``(nspecies, dim, ...)``."""
    dim = state.dim
    grad_y = species_mass_fraction_gradient(state.cv, grad_cv, dim)
    rho = state.mass
    d = state.tv.species_diffusivity
    y = state.species_mass_fractions
    if grad_y.shape[0] == 0:
        return grad_y
    if state.is_mixture:
        dgy = d[:, None] * grad_y
        return -rho * (dgy - y[:, None] * np.sum(dgy, axis=0)[None])
    return -rho * (d[:, None] * grad_y)


def viscous_flux(state: State, grad_cv, grad_t):
    """mirgecom/viscous.py:232-293: ``(F, dim, ...)``."""
    dim = state.dim
    tau = viscous_stress_tensor(state, grad_cv)
    j = diffusive_flux(state, grad_cv)
    v = state.velocity
    tau_v = np.einsum("ij...,j...->i...", tau, v)
    q_d = 0
    if state.is_mixture and j.shape[0] > 0:
        q_d = np.sum(state.species_enthalpies[:, None] * j, axis=0)
    q_c = -state.tv.thermal_conductivity * grad_t
    energy = tau_v - q_d - q_c
    mass = 0 * state.momentum
    rows = [mass, energy] + [tau[i] for i in range(dim)] \
        + [-j[k] for k in range(j.shape[0])]
    return np.stack(rows)


def viscous_facial_flux_central(s_int, s_ext, grad_cv_int, grad_cv_ext,
                                grad_t_int, grad_t_ext, normal):
    """mirgecom/viscous.py:296-349"""
    f_int = viscous_flux(s_int, grad_cv_int, grad_t_int)
    f_ext = viscous_flux(s_ext, grad_cv_ext, grad_t_ext)
    return _dot_normal(num_flux_central(f_int, f_ext), normal)

# }}}
