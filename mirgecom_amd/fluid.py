"""``ConservedVars``: the conserved state as one contiguous HBM tensor.

Same field order and meaning as the reference's ConservedVars
(mirgecom/fluid.py:49-275, order at :311-333): ``[mass, energy,
momentum_0..dim-1, species_mass_0..nspecies-1]``.  Instead of an object array
of per-field DOFArrays, the state is a single ``(F, *shape)`` tensor (on the
volume ``shape = (K, Np)``), so that the kernels read one base pointer and
field ``f`` starts at ``f * N``.  The field attributes are views into it.
"""

from __future__ import annotations

import torch


class ConservedVars:
    """Conserved fluid state; ``q`` has shape ``(2 + dim + nspecies, ...)``."""

    __array_priority__ = 1000

    def __init__(self, dim: int, q: torch.Tensor):
        if q.shape[0] < 2 + dim:
            raise ValueError("state has fewer fields than 2 + dim")
        self.dim = dim
        self.q = q

    # {{{ field views (mirgecom/fluid.py:234-275)

    @property
    def mass(self):
        return self.q[0]

    @property
    def energy(self):
        return self.q[1]

    @property
    def momentum(self):
        return self.q[2:2 + self.dim]

    @property
    def species_mass(self):
        return self.q[2 + self.dim:]

    @property
    def nspecies(self):
        return self.q.shape[0] - 2 - self.dim

    @property
    def velocity(self):
        return self.momentum / self.mass

    @property
    def speed(self):
        v = self.velocity
        return torch.sqrt(torch.sum(v * v, dim=0))

    @property
    def species_mass_fractions(self):
        return self.species_mass / self.mass

    def join(self):
        return self.q

    # }}}

    # {{{ arithmetic (what time integrators need)

    def _wrap(self, q):
        return ConservedVars(self.dim, q)

    @staticmethod
    def _val(x):
        return x.q if isinstance(x, ConservedVars) else x

    def __add__(self, other):
        return self._wrap(self.q + self._val(other))

    __radd__ = __add__

    def __sub__(self, other):
        return self._wrap(self.q - self._val(other))

    def __rsub__(self, other):
        return self._wrap(self._val(other) - self.q)

    def __mul__(self, other):
        return self._wrap(self.q * self._val(other))

    __rmul__ = __mul__

    def __truediv__(self, other):
        return self._wrap(self.q / self._val(other))

    def __neg__(self):
        return self._wrap(-self.q)

    # }}}

    def __repr__(self):
        return (f"ConservedVars(dim={self.dim}, nspecies={self.nspecies}, "
                f"shape={tuple(self.q.shape)}, dtype={self.q.dtype})")


def make_conserved(dim, mass=None, energy=None, momentum=None,
                   species_mass=None, q=None, scalar_quantities=None,
                   vector_quantities=None):
    """mirgecom/fluid.py:338-352"""
    for joined in (scalar_quantities, vector_quantities, q):
        if joined is not None:
            return ConservedVars(dim, joined)
    if mass is None or energy is None or momentum is None:
        raise ValueError("Must have one of *q* or *mass, energy, momentum*.")
    rows = [mass[None], energy[None], momentum]
    if species_mass is not None and len(species_mass) > 0:
        rows.append(species_mass)
    return ConservedVars(dim, torch.cat(rows, dim=0).contiguous())


def velocity_gradient(cv, grad_cv):
    """mirgecom/fluid.py:355-397: ``grad_cv`` is ``(F, dim, ...)``; returns
    ``(dim_v, dim_x, ...)``."""
    return (grad_cv.momentum - cv.velocity[:, None] * grad_cv.mass[None]) \
        / cv.mass


def species_mass_fraction_gradient(cv, grad_cv):
    """mirgecom/fluid.py:400-427"""
    return (grad_cv.species_mass
            - cv.species_mass_fractions[:, None] * grad_cv.mass[None]) / cv.mass
