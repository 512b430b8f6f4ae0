"""Oracle initial/exact states.

* ``vortex2d`` -- mirgecom/initializers.py:143-225 (isentropic vortex, exact
  solution at time t for the advected vortex)
* ``uniform`` -- mirgecom/initializers.py:989-1088 (constant state)
* ``taylor_green`` -- the textbook compressible Taylor-Green vortex (not the
  reference's ``InviscidTaylorGreenVortex``, whose z coordinate and pressure
  differ, mirgecom/initializers.py:1566-1576)

``x`` has the coordinate axis first: ``(dim, ...)``; the result is a
conserved array ``(F, ...)``.
"""

from __future__ import annotations

import numpy as np


def vortex2d(x, time=0.0, beta=5.0, center=(0.0, 0.0), velocity=(0.0, 0.0),
             gamma=1.4):
    """mirgecom/initializers.py:185-225"""
    center = np.asarray(center, dtype=np.float64)
    velocity = np.asarray(velocity, dtype=np.float64)
    loc = center + time * velocity
    x_rel = x[0] - loc[0]
    y_rel = x[1] - loc[1]
    r = np.sqrt(x_rel**2 + y_rel**2)
    expterm = beta * np.exp(1 - r**2)
    u = velocity[0] - expterm * y_rel / (2 * np.pi)
    v = velocity[1] + expterm * x_rel / (2 * np.pi)
    mass = (1 - (gamma - 1) / (16 * gamma * np.pi**2)
            * expterm**2)**(1 / (gamma - 1))
    p = mass**gamma
    energy = p / (gamma - 1) + mass / 2 * (u**2 + v**2)
    return np.stack([mass, energy, mass * u, mass * v])


def uniform(x, rho=1.0, p=1.0, velocity=None, mass_fractions=None, gamma=1.4):
    """mirgecom/initializers.py:989-1088 (ideal gas)."""
    dim = x.shape[0]
    one = 0 * x[0] + 1.0
    vel = np.zeros(dim) if velocity is None else np.asarray(velocity)
    mom = [rho * vel[i] * one for i in range(dim)]
    energy = p / (gamma - 1) + 0.5 * rho * np.dot(vel, vel)
    rows = [rho * one, energy * one] + mom
    if mass_fractions is not None:
        rows += [rho * y * one for y in mass_fractions]
    return np.stack(rows)


def taylor_green(x, rho0=1.0, v0=1.0, mach=0.1, gamma=1.4):
    """Textbook 3-D compressible Taylor-Green vortex on [0, 2 pi]^3:
    ``u = V0 sin x cos y cos z``, ``v = -V0 cos x sin y cos z``, ``w = 0``,
    ``p = p0 + rho0 V0^2/16 (cos 2x + cos 2y)(cos 2z + 2)`` with
    ``p0 = rho0 V0^2 / (gamma Ma^2)``, constant density."""
    p0 = rho0 * v0**2 / (gamma * mach**2)
    u = v0 * np.sin(x[0]) * np.cos(x[1]) * np.cos(x[2])
    v = -v0 * np.cos(x[0]) * np.sin(x[1]) * np.cos(x[2])
    w = 0 * x[0]
    p = p0 + rho0 * v0**2 / 16 * (np.cos(2 * x[0]) + np.cos(2 * x[1])) \
        * (np.cos(2 * x[2]) + 2)
    rho = rho0 + 0 * x[0]
    energy = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, energy, rho * u, rho * v, rho * w])
