"""Viscous fluxes (call surface of mirgecom/viscous.py:72-349) on tensors.

Inside ``ns_operator`` the HIP kernels evaluate these pointwise; the tensor
versions serve direct callers."""

from __future__ import annotations

import torch

from mirgecom_amd.dof_desc import DD_VOLUME_ALL
from mirgecom_amd.fluid import (
    ConservedVars, species_mass_fraction_gradient, velocity_gradient)
from mirgecom_amd.flux import num_flux_central


def viscous_stress_tensor(state, grad_cv):
    """mirgecom/viscous.py:68-105"""
    dim = state.dim
    mu = state.viscosity
    mu_b = state.bulk_viscosity
    gv = velocity_gradient(state.cv, grad_cv)
    div_v = sum(gv[i, i] for i in range(dim))
    eye = torch.eye(dim, dtype=gv.dtype, device=gv.device).reshape(
        (dim, dim) + (1,) * div_v.dim())
    return mu * (gv + gv.transpose(0, 1)) + (mu_b - 2 * mu / 3) * div_v * eye


def diffusive_flux(state, grad_cv):
    """mirgecom/viscous.py:108-153"""
    grad_y = species_mass_fraction_gradient(state.cv, grad_cv)
    rho = state.mass_density
    d = state.species_diffusivity
    y = state.species_mass_fractions
    if state.is_mixture:
        dgy = d[:, None] * grad_y
        return -rho * (dgy - y[:, None] * dgy.sum(dim=0)[None])
    return -rho * (d[:, None] * grad_y)


def viscous_flux(state, grad_cv, grad_t):
    """mirgecom/viscous.py:232-293: ConservedVars of ``(F, dim, ...)``."""
    dim = state.dim
    tau = viscous_stress_tensor(state, grad_cv)
    j = diffusive_flux(state, grad_cv)
    v = state.velocity
    energy = torch.einsum("ij...,j...->i...", tau, v) \
        + state.thermal_conductivity * grad_t
    if state.is_mixture and j.shape[0] > 0:
        energy = energy - (state.species_enthalpies[:, None] * j).sum(dim=0)
    rows = [torch.zeros_like(v), energy] + [tau[i] for i in range(dim)] \
        + [-j[k] for k in range(j.shape[0])]
    return ConservedVars(dim, torch.stack(rows))


def viscous_facial_flux_central(dcoll, state_pair, grad_cv_pair, grad_t_pair,
                                gas_model=None, normal=None):
    """mirgecom/viscous.py:296-349 (``normal`` given explicitly)."""
    f_int = viscous_flux(state_pair.int, grad_cv_pair.int, grad_t_pair.int)
    f_ext = viscous_flux(state_pair.ext, grad_cv_pair.ext, grad_t_pair.ext)
    return ConservedVars(state_pair.int.dim, torch.einsum(
        "fi...,i...->f...", num_flux_central(f_int.q, f_ext.q), normal))


def get_local_max_species_diffusivity(actx, d_alpha):
    """mirgecom/viscous.py:660-680: node-wise maximum over species of the
    diffusivities (0 without species; ``actx`` is unused)."""
    if len(d_alpha) == 0:
        return 0
    if not isinstance(d_alpha[0], torch.Tensor):
        return max(d_alpha)
    out = d_alpha[0]
    for d in d_alpha[1:]:
        out = torch.maximum(out, d)
    return out


def get_viscous_timestep(dcoll, state, *, dd=DD_VOLUME_ALL):
    """mirgecom/viscous.py:547-607: node-local
    ``h_e / (|v| + c + (mu / rho + max_k d_k) / h_e)`` for a viscous state,
    ``h_e / (|v| + c)`` otherwise; shape ``(K, Np)``, one pass of
    ``mg_local_timestep``."""
    from mirgecom_amd._opcommon import check_dd, local_timestep
    check_dd(dd)
    return local_timestep(dcoll, state, viscous=state.is_viscous)


def get_viscous_cfl(dcoll, dt, state, *, dd=DD_VOLUME_ALL):
    """mirgecom/viscous.py:610-636: ``dt / get_viscous_timestep``."""
    return dt / get_viscous_timestep(dcoll, state=state, dd=dd)
