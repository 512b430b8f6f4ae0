"""Initial and exact states on device node coordinates.

* ``Vortex2D`` -- mirgecom/initializers.py:143-225
* ``Uniform`` -- mirgecom/initializers.py:989-1088 (ideal gas)
* ``TaylorGreenVortex`` -- textbook compressible Taylor-Green vortex (the
  reference's InviscidTaylorGreenVortex has a z-coordinate slip and another
  pressure, mirgecom/initializers.py:1566-1576)
"""

from __future__ import annotations

import numpy as np
import torch

from mirgecom_amd.eos import IdealSingleGas
from mirgecom_amd.fluid import make_conserved


class Vortex2D:
    """Isentropic vortex (mirgecom/initializers.py:143-225)."""

    def __init__(self, *, beta=5, center=(0, 0), velocity=(0, 0)):
        self._beta = beta
        self._center = np.array(center, dtype=np.float64)
        self._velocity = np.array(velocity, dtype=np.float64)

    def __call__(self, x_vec, *, time=0, eos=None, **kwargs):
        if eos is None:
            eos = IdealSingleGas()
        loc = self._center + time * self._velocity
        x_rel = x_vec[0] - loc[0]
        y_rel = x_vec[1] - loc[1]
        gamma = eos.gamma()
        r = torch.sqrt(x_rel**2 + y_rel**2)
        expterm = self._beta * torch.exp(1 - r**2)
        u = self._velocity[0] - expterm * y_rel / (2 * np.pi)
        v = self._velocity[1] + expterm * x_rel / (2 * np.pi)
        mass = (1 - (gamma - 1) / (16 * gamma * np.pi**2)
                * expterm**2)**(1 / (gamma - 1))
        p = mass**gamma
        energy = p / (gamma - 1) + mass / 2 * (u**2 + v**2)
        return make_conserved(dim=2, mass=mass, energy=energy,
                              momentum=torch.stack([mass * u, mass * v]))


class Uniform:
    """Uniform state (mirgecom/initializers.py:989-1088)."""

    def __init__(self, dim, rho=1.0, pressure=1.0, velocity=None,
                 species_mass_fractions=None):
        self._dim = dim
        self._rho = rho
        self._p = pressure
        self._vel = np.zeros(dim) if velocity is None else np.asarray(velocity)
        self._y = (np.zeros(0) if species_mass_fractions is None
                   else np.asarray(species_mass_fractions))

    def __call__(self, x_vec, eos=None, **kwargs):
        if eos is None:
            eos = IdealSingleGas()
        one = torch.ones_like(x_vec[0])
        mass = self._rho * one
        mom = torch.stack([self._rho * float(v) * one for v in self._vel])
        energy = (self._p / (eos.gamma() - 1.0)
                  + 0.5 * self._rho * float(np.dot(self._vel, self._vel))) * one
        species = (torch.stack([self._rho * float(y) * one for y in self._y])
                   if len(self._y) else None)
        return make_conserved(self._dim, mass=mass, energy=energy,
                              momentum=mom, species_mass=species)


class TaylorGreenVortex:
    """Compressible Taylor-Green vortex on [0, 2 pi]^3 (textbook form):
    ``u = V0 sin x cos y cos z``, ``v = -V0 cos x sin y cos z``, ``w = 0``,
    ``p = p0 + rho0 V0^2/16 (cos 2x + cos 2y)(cos 2z + 2)``,
    ``p0 = rho0 V0^2/(gamma Ma^2)``, uniform density ``rho0``."""

    def __init__(self, rho0=1.0, v0=1.0, mach=0.1):
        self.rho0, self.v0, self.mach = rho0, v0, mach

    def p0(self, gamma):
        return self.rho0 * self.v0**2 / (gamma * self.mach**2)

    def __call__(self, x_vec, eos=None, **kwargs):
        if eos is None:
            eos = IdealSingleGas()
        gamma = eos.gamma()
        x, y, z = x_vec[0], x_vec[1], x_vec[2]
        u = self.v0 * torch.sin(x) * torch.cos(y) * torch.cos(z)
        v = -self.v0 * torch.cos(x) * torch.sin(y) * torch.cos(z)
        w = torch.zeros_like(x)
        p = self.p0(gamma) + self.rho0 * self.v0**2 / 16 * (
            torch.cos(2 * x) + torch.cos(2 * y)) * (torch.cos(2 * z) + 2)
        rho = self.rho0 * torch.ones_like(x)
        energy = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
        return make_conserved(3, mass=rho, energy=energy,
                              momentum=torch.stack([rho * u, rho * v, rho * w]))

    def transport_for(self, reynolds=1600.0, prandtl=0.71, gamma=1.4,
                      gas_const=287.1, length=1.0):
        """``(mu, kappa)`` for ``Re = rho0 V0 L / mu`` and ``Pr = mu cp/kappa``."""
        mu = self.rho0 * self.v0 * length / reynolds
        cp = gas_const * gamma / (gamma - 1)
        return mu, mu * cp / prandtl


__all__ = ["Vortex2D", "Uniform", "TaylorGreenVortex"]
