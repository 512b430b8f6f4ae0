"""Equations of state: ``IdealSingleGas`` and the uiuc_7sp ``PyrometheusMixture``.

Call surface of mirgecom/eos.py:302-888.  The ideal-gas relations are cheap
pointwise tensor expressions (the DG kernels recompute them in registers);
the mixture temperature -- the Newton inversion of the internal energy --
runs in the HIP kernel ``mg_mixture_temperature`` in FP64.
"""

from __future__ import annotations

import torch

from mirgecom_amd import mechanisms as mech


class TemperatureSeedMissingError(Exception):
    """mirgecom/eos.py:60-63"""


class IdealSingleGas:
    """mirgecom/eos.py:302-537; ``p = (gamma - 1) (rhoE - |rho v|^2/(2 rho))``,
    ``T = (gamma - 1)/R * e / rho``, ``c = sqrt(gamma p / rho)``."""

    is_mixture = False

    def __init__(self, gamma=1.4, gas_const=287.1):
        self._gamma = float(gamma)
        self._gas_const = float(gas_const)

    def gamma(self, cv=None, temperature=None):
        return self._gamma

    def gas_const(self, cv=None, temperature=None, species_mass_fractions=None):
        return self._gas_const

    def heat_capacity_cp(self, cv=None, temperature=None):
        return self._gas_const * self._gamma / (self._gamma - 1)

    def heat_capacity_cv(self, cv=None, temperature=None):
        return self._gas_const / (self._gamma - 1)

    def get_density(self, pressure, temperature, species_mass_fractions=None):
        return pressure / (self._gas_const * temperature)

    def kinetic_energy(self, cv):
        mom = cv.momentum
        return 0.5 * torch.sum(mom * mom, dim=0) / cv.mass

    def internal_energy(self, cv):
        return cv.energy - self.kinetic_energy(cv)

    def pressure(self, cv, temperature=None):
        return self.internal_energy(cv) * (self._gamma - 1.0)

    def sound_speed(self, cv, temperature=None):
        return torch.sqrt(self._gamma / cv.mass * self.pressure(cv))

    def temperature(self, cv, temperature_seed=None):
        return (((self._gamma - 1.0) / self._gas_const)
                * self.internal_energy(cv) / cv.mass)

    def total_energy(self, cv, pressure, temperature=None):
        return pressure / (self._gamma - 1.0) + self.kinetic_energy(cv)

    def get_internal_energy(self, temperature, species_mass_fractions=None):
        return self._gas_const * temperature / (self._gamma - 1)


def _nasa7(t, coef_lo, coef_hi):
    """``(cp/R, h/RT, s/R)`` per species, shape ``(7, ...)``."""
    lo = torch.as_tensor(coef_lo, dtype=t.dtype, device=t.device)
    hi = torch.as_tensor(coef_hi, dtype=t.dtype, device=t.device)
    shape = (-1,) + (1,) * t.dim()
    high = (t > mech.T_MID)[None]
    a = [torch.where(high, hi[:, i].reshape(shape), lo[:, i].reshape(shape))
         for i in range(7)]
    tt = t[None]
    cp = a[0] + a[1] * tt + a[2] * tt**2 + a[3] * tt**3 + a[4] * tt**4
    h = (a[0] + a[1] * 0.5 * tt + a[2] / 3.0 * tt**2 + a[3] * 0.25 * tt**3
         + a[4] * 0.2 * tt**4 + a[5] / tt)
    s = (a[0] * torch.log(tt) + a[1] * tt + a[2] * 0.5 * tt**2
         + a[3] / 3.0 * tt**3 + a[4] * 0.25 * tt**4 + a[6])
    return cp, h, s


class PyrometheusMixture:
    """The uiuc_7sp mixture EOS (call surface of mirgecom/eos.py:540-888,
    thermochemistry of mirgecom/thermochemistry.py:87-167).

    ``temperature_niter`` fixed Newton iterations from a required seed, as
    in the reference's PyroWrapper (default 5)."""

    is_mixture = True
    nspecies = mech.NSPECIES
    species_names = mech.SPECIES

    def __init__(self, pyrometheus_mech=None, temperature_guess=300.0,
                 temperature_niter=5, temperature_tolerance=None,
                 temperature_max_iter=50):
        if pyrometheus_mech not in (None, "uiuc_7sp"):
            raise NotImplementedError("only the uiuc_7sp mechanism is built in")
        self._tguess = temperature_guess
        self.temperature_niter = int(temperature_niter)
        # tolerance mode (not in the reference): iterate each node until its
        # Newton update is at most temperature_tolerance K
        self.temperature_tolerance = (None if temperature_tolerance is None
                                      else float(temperature_tolerance))
        self.temperature_max_iter = int(temperature_max_iter)
        self.last_temperature_iterations = None
        self._wt = mech.molecular_weights()
        self._lo, self._hi = mech.nasa7_tables()

    def _w(self, ref):
        return torch.as_tensor(self._wt, dtype=ref.dtype,
                               device=ref.device).reshape((-1,) + (1,) * ref.dim())

    def get_species_molecular_weights(self):
        return self._wt.copy()

    def gas_const(self, cv=None, temperature=None, species_mass_fractions=None):
        y = species_mass_fractions if cv is None else cv.species_mass_fractions
        return mech.GAS_CONSTANT * torch.sum(y / self._w(y[0]), dim=0)

    def heat_capacity_cp(self, cv, temperature):
        cp_r, _, _ = _nasa7(temperature, self._lo, self._hi)
        y = cv.species_mass_fractions
        return mech.GAS_CONSTANT * torch.sum(cp_r * y / self._w(temperature), dim=0)

    def heat_capacity_cv(self, cv, temperature):
        return self.heat_capacity_cp(cv, temperature) - self.gas_const(cv)

    def gamma(self, cv, temperature=None):
        cp = self.heat_capacity_cp(cv, temperature)
        return cp / (cp - self.gas_const(cv))

    def kinetic_energy(self, cv):
        mom = cv.momentum
        return 0.5 * torch.sum(mom * mom, dim=0) / cv.mass

    def internal_energy(self, cv):
        return cv.energy - self.kinetic_energy(cv)

    def pressure(self, cv, temperature):
        return cv.mass * self.gas_const(cv) * temperature

    def sound_speed(self, cv, temperature):
        return torch.sqrt(self.gamma(cv, temperature)
                          * self.pressure(cv, temperature) / cv.mass)

    def species_enthalpies(self, cv, temperature):
        _, h_rt, _ = _nasa7(temperature, self._lo, self._hi)
        return mech.GAS_CONSTANT / self._w(temperature) * temperature * h_rt

    def get_internal_energy(self, temperature, species_mass_fractions):
        _, h_rt, _ = _nasa7(temperature, self._lo, self._hi)
        y = species_mass_fractions
        return mech.GAS_CONSTANT * temperature * torch.sum(
            (h_rt - 1.0) * y / self._w(temperature), dim=0)

    def get_density(self, pressure, temperature, species_mass_fractions):
        y = species_mass_fractions
        rmix = mech.GAS_CONSTANT * torch.sum(y / self._w(temperature), dim=0)
        return pressure / (rmix * temperature)

    def total_energy(self, cv, pressure, temperature):
        return (cv.mass * self.get_internal_energy(
            temperature, cv.species_mass_fractions) + self.kinetic_energy(cv))

    def get_temperature_seed(self, ary, temperature_seed=None):
        tseed = self._tguess if temperature_seed is None else temperature_seed
        if isinstance(tseed, torch.Tensor):
            return tseed
        return torch.full_like(ary, float(tseed))

    def temperature(self, cv, temperature_seed=None):
        """Newton solve on the GPU (mg_mixture_temperature)."""
        if temperature_seed is None:
            raise TemperatureSeedMissingError(
                "MixtureEOS.get_temperature requires a *temperature_seed*.")
        from mirgecom_amd import _lib
        tseed = self.get_temperature_seed(cv.mass, temperature_seed)
        q = cv.q.contiguous()
        tseed = tseed.to(dtype=q.dtype).expand_as(cv.mass).contiguous()
        out = torch.empty_like(cv.mass, memory_format=torch.contiguous_format)
        _lib.ensure_mechanism()
        gas = _lib.MgGas(eos_kind=1, nspecies=7)
        n = cv.mass.numel()
        if self.temperature_tolerance is not None:
            iters = torch.empty(cv.mass.shape, dtype=torch.int32, device=q.device)
            _lib.call("mg_mixture_temperature_tol", gas, cv.dim, n, _lib.ptr(q),
                      _lib.ptr(tseed), self.temperature_tolerance,
                      self.temperature_max_iter, _lib.ptr(out), _lib.ptr(iters),
                      _lib.stream_handle())
            # per-node updates taken (-1: not converged within max_iter)
            self.last_temperature_iterations = iters
            return out
        _lib.call("mg_mixture_temperature", gas, cv.dim, n, _lib.ptr(q),
                  _lib.ptr(tseed), self.temperature_niter, _lib.ptr(out),
                  _lib.stream_handle())
        return out

    def get_production_rates(self, cv, temperature):
        src = self.get_species_source_terms(cv, temperature)
        return src.species_mass / self._w(temperature)

    def get_species_source_terms(self, cv, temperature):
        """``W_k omega_k`` on the GPU (mg_species_source); mass, energy and
        momentum sources are zero (mirgecom/eos.py:858-875)."""
        from mirgecom_amd import _lib
        from mirgecom_amd.fluid import ConservedVars
        q = cv.q.contiguous()
        t = temperature.contiguous()
        out = torch.zeros_like(q)
        n = cv.mass.numel()
        _lib.ensure_mechanism()
        gas = _lib.MgGas(eos_kind=1, nspecies=7)
        _lib.call("mg_species_source", gas, cv.dim, n, _lib.ptr(q), _lib.ptr(t),
                  _lib.ptr(out[2 + cv.dim]), _lib.stream_handle())
        return ConservedVars(cv.dim, out)

    def add_species_source_terms(self, cv, temperature, rhs):
        """``rhs + get_species_source_terms(cv, temperature)`` formed in place
        in ``rhs`` (a ``ConservedVars`` shaped like ``cv``) by one kernel
        (mg_species_source_add); returns ``rhs``."""
        from mirgecom_amd import _lib
        q = cv.q.contiguous()
        t = temperature.contiguous()
        r = rhs.q
        if (r.shape != q.shape or r.dtype != torch.float64 or not r.is_contiguous()
                or r.device != q.device):
            raise ValueError("rhs must be a contiguous float64 tensor shaped like "
                             "the state, on its device")
        _lib.ensure_mechanism()
        gas = _lib.MgGas(eos_kind=1, nspecies=7)
        _lib.call("mg_species_source_add", gas, cv.dim, cv.mass.numel(),
                  _lib.ptr(q), _lib.ptr(t), _lib.ptr(r), _lib.stream_handle())
        return rhs

