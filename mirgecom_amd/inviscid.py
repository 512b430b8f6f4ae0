"""Inviscid fluxes (call surface of mirgecom/inviscid.py:65-376).

``inviscid_flux`` and the facial fluxes are tensor functions of fluid states
(with a TracePair-like ``state_pair.int``/``state_pair.ext``); inside
``euler_operator``/``ns_operator`` the same fluxes are evaluated by the HIP
kernels, selected by identity of the function passed in."""

from __future__ import annotations

import torch

from mirgecom_amd import _lib
from mirgecom_amd.dof_desc import DD_VOLUME_ALL
from mirgecom_amd.flux import num_flux_hll, num_flux_lfr
from mirgecom_amd.fluid import ConservedVars


def inviscid_flux(state):
    """mirgecom/inviscid.py:65-102: ``(F, dim, ...)`` as ConservedVars."""
    dim = state.dim
    v = state.velocity
    p = state.pressure
    rows = [state.momentum_density, v * (state.energy_density + p)]
    rho = state.mass_density
    for i in range(dim):
        rows.append(torch.stack([rho * (v[i] * v[j]) + (p if i == j else 0 * p)
                                 for j in range(dim)]))
    for k in range(state.nspecies):
        rows.append(v * state.species_mass_density[k])
    return ConservedVars(dim, torch.stack(rows))


def _fn(flux, normal):
    return torch.einsum("fi...,i...->f...", flux.q, normal)


def inviscid_facial_flux_rusanov(state_pair, gas_model, normal):
    """mirgecom/inviscid.py:105-152"""
    s_i, s_e = state_pair.int, state_pair.ext
    lam = torch.maximum(s_i.wavespeed, s_e.wavespeed)
    return ConservedVars(s_i.dim, num_flux_lfr(
        _fn(inviscid_flux(s_i), normal), _fn(inviscid_flux(s_e), normal),
        s_i.cv.q, s_e.cv.q, lam))


def inviscid_facial_flux_central(state_pair, gas_model, normal):
    """mirgecom/inviscid.py:155-199"""
    s_i, s_e = state_pair.int, state_pair.ext
    return ConservedVars(s_i.dim, num_flux_lfr(
        _fn(inviscid_flux(s_i), normal), _fn(inviscid_flux(s_e), normal),
        s_i.cv.q, s_e.cv.q, 0))


def inviscid_facial_flux_hll(state_pair, gas_model, normal):
    """mirgecom/inviscid.py:202-287"""
    s_i, s_e = state_pair.int, state_pair.ext
    p_i, p_e = s_i.pressure, s_e.pressure
    u_i = torch.einsum("i...,i...->...", s_i.velocity, normal)
    u_e = torch.einsum("i...,i...->...", s_e.velocity, normal)
    c_i, c_e = s_i.speed_of_sound, s_e.speed_of_sound
    p_star = (0.5 * (p_i + p_e) + (1. / 8.) * (u_i - u_e)
              * (s_i.mass_density + s_e.mass_density) * (c_i + c_e))
    g_i = gas_model.eos.gamma(s_i.cv, s_i.temperature)
    g_e = gas_model.eos.gamma(s_e.cv, s_e.temperature)
    q_i = 1 + (g_i + 1) / (2 * g_i) * (p_star / p_i - 1)
    q_e = 1 + (g_e + 1) / (2 * g_e) * (p_star / p_e - 1)
    one = torch.ones_like(p_i)
    s_minus = u_i - c_i * torch.sqrt(torch.where(p_star > p_i, q_i, one))
    s_plus = u_e + c_e * torch.sqrt(torch.where(p_star > p_e, q_e, one))
    return ConservedVars(s_i.dim, num_flux_hll(
        _fn(inviscid_flux(s_i), normal), _fn(inviscid_flux(s_e), normal),
        s_i.cv.q, s_e.cv.q, s_minus, s_plus))


def _ln_mean(x, y, epsilon=1e-4):
    """mirgecom/inviscid.py:470-476"""
    f2 = (x * (x - 2 * y) + y * y) / (x * (x + 2 * y) + y * y)
    series = (x + y) / (2 + f2 * 2 / 3 + f2 * f2 * 2 / 5 + f2 * f2 * f2 * 2 / 7)
    safe = torch.where(f2 < epsilon, torch.full_like(x, 2.0), y / x)
    return torch.where(f2 < epsilon, series, (y - x) / torch.log(safe))


def entropy_conserving_flux_chandrashekar(gas_model, state_ll, state_rr):
    """mirgecom/inviscid.py:451-518: the two-point flux ``(F, dim, ...)``
    (Chandrashekar 2013, eqs. 4.12-4.14)."""
    dim = state_ll.dim
    g_ll = gas_model.eos.gamma(state_ll.cv, state_ll.temperature)
    g_rr = gas_model.eos.gamma(state_rr.cv, state_rr.temperature)
    rho_l, rho_r = state_ll.mass_density, state_rr.mass_density
    u_l, u_r = state_ll.velocity, state_rr.velocity
    beta_l = 0.5 * rho_l / state_ll.pressure
    beta_r = 0.5 * rho_r / state_rr.pressure
    rho_avg = 0.5 * (rho_l + rho_r)
    rho_mean = _ln_mean(rho_l, rho_r)
    beta_mean = _ln_mean(beta_l, beta_r)
    beta_avg = 0.5 * (beta_l + beta_r)
    u_avg = 0.5 * (u_l + u_r)
    p_mean = 0.5 * rho_avg / beta_avg
    vsq_avg = 0.5 * (u_l * u_l).sum(dim=0) + 0.5 * (u_r * u_r).sum(dim=0)
    mass = rho_mean * u_avg
    mom = torch.stack([mass[i] * u_avg + torch.stack(
        [p_mean if j == i else 0 * p_mean for j in range(dim)])
        for i in range(dim)])
    gamma = 0.5 * (g_ll + g_rr)
    energy = (mass * 0.5 * (1 / (gamma - 1) / beta_mean - vsq_avg)
              + torch.einsum("ij...,j...->i...", mom, u_avg))
    rows = [mass, energy] + [mom[i] for i in range(dim)]
    y_mean = 0.5 * (state_ll.species_mass_fractions
                    + state_rr.species_mass_fractions)
    rows += [rho_mean * y_mean[k] * u_avg for k in range(state_ll.nspecies)]
    return ConservedVars(dim, torch.stack(rows))


def entropy_conserving_flux_renac(gas_model, state_ll, state_rr):
    """mirgecom/inviscid.py:521-598: Renac's entropy-conserving,
    kinetic-energy-preserving two-point flux ``(F, dim, ...)`` (Renac 2021,
    eq. 24), the mixtures' flux of the entropy-stable operator."""
    dim = state_ll.dim
    t_l, t_r = state_ll.temperature, state_rr.temperature
    p_l, p_r = state_ll.pressure, state_rr.pressure
    g_l = gas_model.eos.gamma(state_ll.cv, t_l)
    g_r = gas_model.eos.gamma(state_rr.cv, t_r)
    th_l, th_r = 1.0 / t_l, 1.0 / t_r
    t_avg = 0.5 * (t_l + t_r)
    pot_avg = 0.5 * (p_l * th_l + p_r * th_r)
    t_mean = 1.0 / _ln_mean(th_l, th_r)
    pec_avg = pot_avg * t_avg
    p_mean = _ln_mean(p_l, p_r)
    rho_l, rho_r = state_ll.mass_density, state_rr.mass_density
    u_l, u_r = state_ll.velocity, state_rr.velocity
    kin_comb = 0.5 * (u_l * u_r).sum(dim=0)
    rho_mean = _ln_mean(rho_l, rho_r)
    y_avg = 0.5 * (state_ll.species_mass_fractions
                   + state_rr.species_mass_fractions)
    spec_mean = rho_mean * y_avg
    u_avg = 0.5 * (u_l + u_r)
    mass = rho_mean * u_avg
    mom = torch.stack([mass[i] * u_avg + torch.stack(
        [pec_avg if j == i else 0 * pec_avg for j in range(dim)])
        for i in range(dim)])
    gamma_avg = 0.5 * (g_l + g_r)
    ener_es = p_mean / (gamma_avg - 1) + 0.5 * rho_mean * (u_avg * u_avg).sum(dim=0)
    cv_es = ConservedVars(dim, torch.cat([rho_mean[None], ener_es[None],
                                          rho_mean * u_avg, spec_mean]))
    cvm = gas_model.eos.heat_capacity_cv(cv_es, t_mean)
    energy = (cvm * t_mean + kin_comb) * mass + pec_avg * u_avg
    rows = [mass, energy] + [mom[i] for i in range(dim)]
    rows += [spec_mean[k] * u_avg for k in range(state_ll.nspecies)]
    return ConservedVars(dim, torch.stack(rows))


def entropy_stable_inviscid_facial_flux(state_pair, gas_model, normal,
                                        entropy_conserving_flux_func=None,
                                        alpha=None):
    """mirgecom/inviscid.py:601-662: ``F_S(q-, q+) . n``, penalised by
    ``alpha/2 (q+ - q-)`` when ``alpha`` is given."""
    if entropy_conserving_flux_func is None:
        entropy_conserving_flux_func = (
            entropy_conserving_flux_renac if state_pair.int.is_mixture
            else entropy_conserving_flux_chandrashekar)
    flux = entropy_conserving_flux_func(gas_model, state_pair.int,
                                        state_pair.ext).q
    if alpha is not None:
        jump = state_pair.ext.cv.q - state_pair.int.cv.q
        flux = flux - 0.5 * alpha * jump[:, None] * normal[None]
    return ConservedVars(state_pair.int.dim,
                         torch.einsum("fi...,i...->f...", flux, normal))


def entropy_stable_inviscid_facial_flux_rusanov(state_pair, gas_model, normal,
                                                entropy_conserving_flux_func=None,
                                                **kwargs):
    """mirgecom/inviscid.py:665-711: the entropy-stable flux with the
    Rusanov penalty ``alpha = max(|v| + c)``."""
    alpha = torch.maximum(state_pair.int.wavespeed, state_pair.ext.wavespeed)
    return entropy_stable_inviscid_facial_flux(
        state_pair, gas_model, normal,
        entropy_conserving_flux_func=entropy_conserving_flux_func, alpha=alpha)


_FLUX_KIND = {
    inviscid_facial_flux_rusanov: _lib.MG_FLUX_RUSANOV,
    inviscid_facial_flux_hll: _lib.MG_FLUX_HLL,
    inviscid_facial_flux_central: _lib.MG_FLUX_CENTRAL,
}


def flux_kind(numerical_flux_func):
    """Device flux selector of a facial flux function."""
    try:
        return _FLUX_KIND[numerical_flux_func]
    except (KeyError, TypeError):
        raise NotImplementedError(
            f"numerical flux {numerical_flux_func!r} has no device kernel; "
            "use inviscid_facial_flux_rusanov, _hll or _central") from None


def get_inviscid_timestep(dcoll, state, dd=DD_VOLUME_ALL):
    """mirgecom/inviscid.py:379-422: node-local ``h_e / (|v| + c)``, shape
    ``(K, Np)``, with ``h_e`` from
    ``mirgecom_amd.discretization.characteristic_lengthscales`` (one pass of
    ``mg_local_timestep``)."""
    from mirgecom_amd._opcommon import check_dd, local_timestep
    check_dd(dd)
    return local_timestep(dcoll, state, viscous=False)


def get_inviscid_cfl(dcoll, state, dt):
    """mirgecom/inviscid.py:425-448: ``dt / get_inviscid_timestep``."""
    return dt / get_inviscid_timestep(dcoll, state=state)
