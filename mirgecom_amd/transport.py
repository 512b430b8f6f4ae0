"""Transport models: ``SimpleTransport`` and ``PowerLawTransport``.

Call surface and formulas of mirgecom/transport.py:143-321.  The DG kernels
evaluate them pointwise in registers from these parameters; the tensor
methods here serve callers that want the transport variables themselves.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class GasTransportVars:
    """mirgecom/transport.py:66-83"""

    bulk_viscosity: torch.Tensor
    viscosity: torch.Tensor
    thermal_conductivity: torch.Tensor
    species_diffusivity: torch.Tensor


class SimpleTransport:
    """Constant mu, mu_B, kappa and species diffusivities
    (mirgecom/transport.py:143-202)."""

    kind = 1

    def __init__(self, bulk_viscosity=0, viscosity=0, thermal_conductivity=0,
                 species_diffusivity=None):
        self._mu_bulk = float(bulk_viscosity)
        self._mu = float(viscosity)
        self._kappa = float(thermal_conductivity)
        self._d_alpha = (np.zeros(0) if species_diffusivity is None
                         else np.asarray(species_diffusivity, dtype=np.float64))

    def device_params(self, nspecies):
        d = np.zeros(8)
        if len(self._d_alpha):
            d[:nspecies] = self._d_alpha[:nspecies]
        return dict(transport_kind=1, mu=self._mu, mu_b=self._mu_bulk,
                    kappa=self._kappa, has_lewis=0, diff=d)

    def transport_vars(self, cv, dv, eos=None):
        one = torch.ones_like(cv.mass)
        d = torch.as_tensor(self._d_alpha, dtype=one.dtype, device=one.device)
        return GasTransportVars(
            bulk_viscosity=self._mu_bulk * one, viscosity=self._mu * one,
            thermal_conductivity=self._kappa * one,
            species_diffusivity=d.reshape((-1,) + (1,) * one.dim()) * one)


class PowerLawTransport:
    """``mu = beta T^n``, ``mu_B = alpha mu``, ``kappa = sigma mu c_v``, species
    diffusivities given or from Lewis numbers (mirgecom/transport.py:205-321)."""

    kind = 2

    def __init__(self, alpha=0.6, beta=4.093e-7, sigma=2.5, n=.666,
                 species_diffusivity=None, lewis=None):
        self._alpha, self._beta, self._sigma, self._n = alpha, beta, sigma, n
        self._d_alpha = species_diffusivity
        self._lewis = lewis

    def device_params(self, nspecies):
        d = np.zeros(8)
        src = self._lewis if self._lewis is not None else self._d_alpha
        if src is not None:
            d[:nspecies] = np.asarray(src, dtype=np.float64)[:nspecies]
        return dict(transport_kind=2, pl_alpha=self._alpha, pl_beta=self._beta,
                    pl_sigma=self._sigma, pl_n=self._n,
                    has_lewis=int(self._lewis is not None), diff=d)

    def transport_vars(self, cv, dv, eos):
        mu = self._beta * dv.temperature**self._n
        kappa = self._sigma * mu * eos.heat_capacity_cv(cv, dv.temperature)
        if self._lewis is not None:
            lew = torch.as_tensor(np.asarray(self._lewis), dtype=mu.dtype,
                                  device=mu.device).reshape((-1,) + (1,) * mu.dim())
            d = self._sigma * mu / (cv.mass * lew * eos.gamma(cv, dv.temperature))
        else:
            dd = torch.as_tensor(np.zeros(cv.nspecies) if self._d_alpha is None
                                 else np.asarray(self._d_alpha), dtype=mu.dtype,
                                 device=mu.device)
            d = dd.reshape((-1,) + (1,) * mu.dim()) * torch.ones_like(mu)
        return GasTransportVars(bulk_viscosity=self._alpha * mu, viscosity=mu,
                                thermal_conductivity=kappa,
                                species_diffusivity=d)
