"""Oracle mixture-averaged transport (NumPy FP64; test infrastructure only).

Restates mirgecom's ``MixtureAveragedTransport`` (mirgecom/transport.py:
324-502) on top of the transport functions of pyrometheus' generated
mechanism class (pyrometheus @ tulio/entropy):

* species properties from the fitted polynomials in ``ln T``:
  ``mu_k = sqrt(T) (sum_i a_ki ln^i T)^2``, ``lambda_k = sqrt(T) sum_i b_ki
  ln^i T``, ``D_kj = T^1.5 sum_i c_kji ln^i T`` (at unit pressure);
* ``get_mixture_viscosity_mixavg``: Wilke's rule
  ``sum_k X_k mu_k / sum_j X_j phi_kj``,
  ``phi_kj = (1 + W_k/W_j)^-1/2 (1 + sqrt(mu_k/mu_j) (W_j/W_k)^1/4)^2 / sqrt 8``;
* ``get_mixture_thermal_conductivity_mixavg``:
  ``(sum X_k lambda_k + 1 / sum X_k/lambda_k) / 2``;
* ``get_species_mass_diffusivities_mixavg``: ``denom_k = sum_j X_j/D_kj -
  X_k/D_kk``; ``(W - X_k W_k) / (p W denom_k)`` where ``denom_k > 0``, else
  ``D_kk / p``;
* mirgecom's own rules on top: ``mu_B = alpha mu``, the singular diffusivity
  where ``1 - Y_k < epsilon``, ``kappa / (rho Le_k c_p)`` with Lewis numbers,
  and the scaling ``factor`` (not applied twice in the Lewis branch).

The fit coefficients are mechanism data: they come from
``mirgecom_amd.transport_fits`` (the reference takes them from Cantera at
code-generation time; Cantera is absent, so the fitted values themselves are
parity unpinned against the reference and are checked against kinetic-theory
closed forms and measured gas data instead).  The mixture rules here are
pinned by the single-species limits (test/test_transport.py:119-159), the
singular-diffusivity rule, the trace-species limit ``d_j = D_jk / p``,
Wilke's symmetry ``phi_jk = phi_kj (mu_j/mu_k)(W_k/W_j)`` with
``phi_kk = 1``, the binary-mixture closed form and the Lewis-number branch
(tests/test_oracle_transport.py).
"""

from __future__ import annotations

import math

import numpy as np

from oracle import thermochem as tc
from oracle.physics import TransportVars


def _poly(c, lt):
    """``sum_i c[i] lt^i`` (coefficients lowest power first)."""
    out = np.zeros_like(lt)
    for i in range(len(c) - 1, -1, -1):
        out = out * lt + c[i]
    return out


def species_viscosities(t, visc):
    lt = np.log(t)
    return np.stack([np.sqrt(t) * _poly(visc[k], lt) ** 2
                     for k in range(len(visc))])


def species_thermal_conductivities(t, cond):
    lt = np.log(t)
    return np.stack([np.sqrt(t) * _poly(cond[k], lt) for k in range(len(cond))])


def species_binary_diffusivities(t, diff):
    """``D_kj`` at unit pressure, ``(nsp, nsp, ...)``."""
    lt = np.log(t)
    n = diff.shape[0]
    return np.stack([np.stack([t ** 1.5 * _poly(diff[i, j], lt)
                               for j in range(n)]) for i in range(n)])


def mix_molecular_weight(y):
    return 1.0 / np.einsum("k,k...->...", 1.0 / tc.W, y)


def mole_fractions(y):
    mmw = mix_molecular_weight(y)
    return mmw * y / tc.W.reshape((-1,) + (1,) * (y.ndim - 1))


def wilke_phi(mu, w):
    """Wilke's interaction factors ``phi[k][j]`` for species viscosities
    ``mu (nsp, ...)`` and molecular weights ``w``."""
    n = len(w)
    return [[(1.0 + w[k] / w[j]) ** -0.5 / math.sqrt(8.0)
             * (1.0 + np.sqrt(mu[k] / mu[j]) * (w[j] / w[k]) ** 0.25) ** 2
             for j in range(n)] for k in range(n)]


def mixture_viscosity(t, y, visc):
    x = mole_fractions(y)
    mu = species_viscosities(t, visc)
    phi = wilke_phi(mu, tc.W)
    n = len(tc.W)
    out = 0.0
    for k in range(n):
        den = 0.0
        for j in range(n):
            den = den + x[j] * phi[k][j]
        out = out + x[k] * mu[k] / den
    return out


def mixture_thermal_conductivity(t, y, cond):
    x = mole_fractions(y)
    lam = species_thermal_conductivities(t, cond)
    return 0.5 * (np.sum(x * lam, axis=0) + 1.0 / np.sum(x / lam, axis=0))


def species_mass_diffusivities(p, t, y, diff):
    x = mole_fractions(y)
    mmw = mix_molecular_weight(y)
    bd = species_binary_diffusivities(t, diff)
    n = len(tc.W)
    out = []
    for k in range(n):
        x_sum = sum(x[j] / bd[k, j] for j in range(n))
        denom = x_sum - x[k] / bd[k, k]
        safe = np.where(denom > 0.0, denom, 1.0)
        out.append(np.where(denom > 0.0,
                            (mmw - x[k] * tc.W[k]) / (p * mmw * safe),
                            bd[k, k] / p))
    return np.stack(out)


class MixtureAveragedTransport:
    """Oracle counterpart of mirgecom's MixtureAveragedTransport with the
    ``oracle.physics`` transport interface."""

    def __init__(self, fits, alpha=0.6, factor=1.0, lewis=None, epsilon=1e-4,
                 singular_diffusivity=1e-6):
        self.visc, self.cond, self.diff = fits
        self.alpha = alpha
        self.factor = factor
        self.lewis = lewis
        self.epsilon = epsilon
        self.singular = singular_diffusivity

    def transport_vars(self, cv, temperature, eos, dim):
        rho = cv[0]
        y = cv[2 + dim:] / rho
        mu = self.factor * mixture_viscosity(temperature, y, self.visc)
        kappa = self.factor * mixture_thermal_conductivity(temperature, y,
                                                           self.cond)
        if self.lewis is not None:
            lew = np.asarray(self.lewis).reshape((-1,) + (1,) * rho.ndim)
            cp = tc.mixture_cp_mass(temperature, y)
            d = kappa / (rho * lew * cp)
        else:
            p = eos.pressure(cv, dim, temperature)
            d = species_mass_diffusivities(p, temperature, y, self.diff)
            d = np.where(1.0 - y < self.epsilon, self.singular, d)
            d = self.factor * d
        return TransportVars(bulk_viscosity=self.alpha * mu, viscosity=mu,
                             thermal_conductivity=kappa, species_diffusivity=d)
