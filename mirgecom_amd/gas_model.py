"""Gas model and model-consistent fluid states.

Call surface of mirgecom/gas_model.py:82-491: ``GasModel(eos, transport)``,
``FluidState(cv, dv)``, ``ViscousFluidState(cv, dv, tv)`` and
``make_fluid_state(cv, gas_model, temperature_seed=None, ...)``.

The dependent variables are evaluated lazily: the DG kernels recompute
pressure, sound speed and transport coefficients in registers from the
conserved state (and, for mixtures, the temperature), so only the mixture
temperature -- the FP64 Newton solve ``mg_mixture_temperature`` -- is computed
eagerly.  The other fields are produced on first access with the reference's
formulas.
"""

from __future__ import annotations

import numpy as np
import torch

from mirgecom_amd import _lib
from mirgecom_amd.fluid import ConservedVars


class GasModel:
    """mirgecom/gas_model.py:82-97"""

    def __init__(self, eos, transport=None):
        self.eos = eos
        self.transport = transport

    def device_params(self, nspecies: int,
                      with_transport: bool = True) -> "_lib.MgGas":
        """The ``mg_gas`` struct of this model for a state with ``nspecies``
        (without the transport model if ``with_transport`` is false)."""
        eos = self.eos
        g = _lib.MgGas()
        g.nspecies = nspecies
        if getattr(eos, "is_mixture", False):
            if nspecies != 7:
                raise ValueError("the uiuc_7sp mixture has 7 species")
            g.eos_kind = 1
        else:
            g.eos_kind = 0
            g.gamma = eos.gamma()
            g.gas_const = eos.gas_const()
        if nspecies > 8:
            raise NotImplementedError("at most 8 species")
        if self.transport is not None and with_transport:
            if getattr(self.transport, "kind", 0) == 3 and g.eos_kind != 1:
                raise ValueError("mixture-averaged transport needs the "
                                 "mixture EOS")
            prm = self.transport.device_params(nspecies)
            diff = prm.pop("diff")
            for k, v in prm.items():
                setattr(g, k, v)
            for k in range(8):
                g.species_diffusivity[k] = float(diff[k])
        return g


class _LazyDV:
    """Dependent variables (mirgecom/eos.py:66-100) computed on access."""

    def __init__(self, cv, eos, temperature=None):
        self._cv = cv
        self._eos = eos
        self._t = temperature
        self._cache = {}

    def _get(self, name, fn):
        if name not in self._cache:
            self._cache[name] = fn()
        return self._cache[name]

    @property
    def temperature(self):
        if self._t is None:
            self._t = self._eos.temperature(self._cv)
        return self._t

    @property
    def pressure(self):
        return self._get("p", lambda: self._eos.pressure(self._cv, self.temperature))

    @property
    def speed_of_sound(self):
        return self._get("c", lambda: self._eos.sound_speed(self._cv,
                                                            self.temperature))

    @property
    def species_enthalpies(self):
        if not getattr(self._eos, "is_mixture", False):
            raise AttributeError("species enthalpies need a mixture EOS")
        return self._get("h", lambda: self._eos.species_enthalpies(
            self._cv, self.temperature))

    @property
    def smoothness_mu(self):
        return self._get("smu", lambda: torch.zeros_like(self._cv.mass))

    smoothness_kappa = smoothness_mu
    smoothness_d = smoothness_mu
    smoothness_beta = smoothness_mu


class FluidState:
    """mirgecom/gas_model.py:100-248"""

    def __init__(self, cv, dv, gas_model=None):
        self.cv = cv
        self.dv = dv
        self.gas_model = gas_model

    dim = property(lambda self: self.cv.dim)
    nspecies = property(lambda self: self.cv.nspecies)
    pressure = property(lambda self: self.dv.pressure)
    temperature = property(lambda self: self.dv.temperature)
    mass_density = property(lambda self: self.cv.mass)
    momentum_density = property(lambda self: self.cv.momentum)
    energy_density = property(lambda self: self.cv.energy)
    species_mass_density = property(lambda self: self.cv.species_mass)
    velocity = property(lambda self: self.cv.velocity)
    speed = property(lambda self: self.cv.speed)
    species_mass_fractions = property(lambda self: self.cv.species_mass_fractions)
    speed_of_sound = property(lambda self: self.dv.speed_of_sound)
    species_enthalpies = property(lambda self: self.dv.species_enthalpies)

    @property
    def wavespeed(self):
        """mirgecom/gas_model.py:223-226"""
        return self.cv.speed + self.dv.speed_of_sound

    @property
    def is_viscous(self):
        return False

    @property
    def is_mixture(self):
        return bool(getattr(self.gas_model.eos, "is_mixture", False)) \
            if self.gas_model is not None else False


class ViscousFluidState(FluidState):
    """mirgecom/gas_model.py:251-286"""

    def __init__(self, cv, dv, gas_model):
        super().__init__(cv, dv, gas_model)
        self._tv = None

    @property
    def tv(self):
        if self._tv is None:
            self._tv = self.gas_model.transport.transport_vars(
                self.cv, self.dv, self.gas_model.eos)
        return self._tv

    viscosity = property(lambda self: self.tv.viscosity)
    bulk_viscosity = property(lambda self: self.tv.bulk_viscosity)
    thermal_conductivity = property(lambda self: self.tv.thermal_conductivity)
    species_diffusivity = property(lambda self: self.tv.species_diffusivity)

    @property
    def is_viscous(self):
        return True


def make_fluid_state(cv, gas_model, temperature_seed=None, smoothness_mu=None,
                     smoothness_kappa=None, smoothness_d=None,
                     smoothness_beta=None, entropy_min=None,
                     material_densities=None, limiter_func=None,
                     limiter_dd=None):
    """mirgecom/gas_model.py:302-431 (GasModel branch).

    ``limiter_func`` is called as in the reference (gas_model.py:373-383)
    with ``cv``, ``temperature_seed``, ``gas_model``, ``dd`` (and
    ``entropy_min`` when given) and returns either the limited CV or an
    object array ``[cv, pressure, temperature]``; the state is then built on
    what it returned.  Artificial-viscosity smoothness fields and porous media
    are outside this build and raise ``NotImplementedError`` when given."""
    if not isinstance(gas_model, GasModel):
        raise NotImplementedError("only GasModel (no porous media)")
    if any(x is not None for x in (smoothness_mu, smoothness_kappa,
                                   smoothness_d, smoothness_beta)):
        raise NotImplementedError("artificial-viscosity fields not supported")
    pressure = temperature = None
    if limiter_func is not None:
        kw = dict(cv=cv, temperature_seed=temperature_seed,
                  gas_model=gas_model, dd=limiter_dd)
        if entropy_min is not None:
            kw["entropy_min"] = entropy_min
        rv = limiter_func(**kw)
        if isinstance(rv, np.ndarray):
            cv, pressure, temperature = rv
        else:
            cv = rv
    if temperature is None and getattr(gas_model.eos, "is_mixture", False):
        temperature = gas_model.eos.temperature(cv, temperature_seed)
    dv = _LazyDV(cv, gas_model.eos, temperature)
    if pressure is not None:
        dv._cache["p"] = pressure
    if gas_model.transport is not None:
        return ViscousFluidState(cv, dv, gas_model)
    return FluidState(cv, dv, gas_model)



def conservative_to_entropy_vars(gamma, state):
    """mirgecom/gas_model.py:1026-1060: entropy variables of ``state`` (its
    pressure, Chandrashekar 2013 eq. 4.5) with ratio ``gamma``."""
    cv = state.cv
    rho = cv.mass
    u = cv.velocity
    p = state.pressure
    u2 = torch.sum(u * u, dim=0)
    s = torch.log(p) - gamma * torch.log(rho)
    rho_p = rho / p
    rows = [(gamma - s) / (gamma - 1) - 0.5 * rho_p * u2, -rho_p]
    rows += [rho_p * u[d] for d in range(cv.dim)]
    rows += [y / (gamma - 1) for y in cv.species_mass_fractions]
    return ConservedVars(cv.dim, torch.stack(rows))


def entropy_to_conservative_vars(gamma, ev):
    """mirgecom/gas_model.py:1063-1100: conserved variables of the entropy
    variables ``ev`` with ratio ``gamma`` (Hughes, Franca & Mallet 1986)."""
    dim = ev.dim
    v = ev.q * (gamma - 1)
    v1, v5 = v[0], v[1]
    v234 = v[2:2 + dim]
    vv = torch.sum(v234 * v234, dim=0)
    s = gamma - v1 + vv / (2 * v5)
    iota = ((gamma - 1) / (-v5) ** gamma) ** (1 / (gamma - 1))
    rho_iota = iota * torch.exp(-s / (gamma - 1))
    mass = -rho_iota * v5
    rows = [mass, rho_iota * (1 - vv / (2 * v5))]
    rows += [rho_iota * v234[d] for d in range(dim)]
    rows += [mass * v[k] for k in range(2 + dim, v.shape[0])]
    return ConservedVars(dim, torch.stack(rows))


class OperatorFluidStates:
    """What ``make_operator_fluid_states`` returns (mirgecom/gas_model.py:
    728-904); it unpacks as ``(vol_state_quad, inter_elem_bnd_states_quad,
    domain_bnd_states_quad)``.

    The device operators form every face-trace and boundary state inside
    their kernels from the volume state, so an operator given these states
    checks that they were made from its own ``state``, quadrature tag and
    boundary tags, and then evaluates exactly what it evaluates without them.
    ``vol_state_quad`` is the volume state (on ``DISCR_TAG_QUAD``: projected
    to the quadrature points with the EOS evaluated there, mixtures by Newton
    from the interpolated temperature; with ``entropy_stable`` through the
    entropy variables and back, gas_model.py:494-588), built on first access;
    the trace pairs and the boundary dictionary are not materialized
    (``None``)."""

    def __init__(self, dcoll, volume_state, gas_model, boundaries,
                 quadrature_tag, entropy_stable):
        from mirgecom_amd.utils import normalize_boundaries
        self.dcoll = dcoll
        self.volume_state = volume_state
        self.gas_model = gas_model
        self.quadrature_tag = quadrature_tag
        self.entropy_stable = bool(entropy_stable)
        self.boundary_tags = frozenset(
            b.tag for b in normalize_boundaries(boundaries))
        self._vol = None

    @property
    def vol_state_quad(self):
        if self._vol is None:
            self._vol = self._project()
        return self._vol

    def _project(self):
        from mirgecom_amd.dof_desc import DISCR_TAG_BASE
        st = self.volume_state
        if self.quadrature_tag == DISCR_TAG_BASE:
            return st
        from mirgecom_amd.device import get_device_discretization
        q = st.cv.q
        dev = get_device_discretization(self.dcoll, q.device)
        interp = dev._quad()["interp"].to(q.dtype)             # (nq, Np)
        cv_q = ConservedVars(st.dim, torch.einsum("qn,fkn->fkq", interp, q))
        seed = None
        if st.is_mixture:
            seed = torch.einsum("qn,kn->kq", interp, st.temperature)
        s = make_fluid_state(cv_q, self.gas_model, temperature_seed=seed)
        if self.entropy_stable:
            g = self.gas_model.eos.gamma(s.cv, s.temperature)
            cv_q = entropy_to_conservative_vars(g, conservative_to_entropy_vars(g, s))
            s = make_fluid_state(cv_q, self.gas_model, temperature_seed=seed)
        return s

    def __iter__(self):
        return iter((self.vol_state_quad, None, None))

    def __len__(self):
        return 3

    def __getitem__(self, i):
        return tuple(self)[i]

    def check(self, state, quadrature_tag, boundaries):
        """Raise ``ValueError`` unless these states were made for ``state``
        on ``quadrature_tag`` with the tags of ``boundaries``."""
        from mirgecom_amd.utils import normalize_boundaries
        if state is not self.volume_state:
            raise ValueError("operator_states_quad were made from another state")
        if quadrature_tag != self.quadrature_tag:
            raise ValueError("operator_states_quad were made for quadrature tag "
                             f"{self.quadrature_tag!r}, not {quadrature_tag!r}")
        tags = frozenset(b.tag for b in normalize_boundaries(boundaries))
        if tags != self.boundary_tags:
            raise ValueError("operator_states_quad were made for other boundaries")


def make_operator_fluid_states(dcoll, volume_state, gas_model, boundaries,
                               quadrature_tag=None, dd=None, comm_tag=None,
                               limiter_func=None, entropy_min=None,
                               entropy_stable=False):
    """mirgecom/gas_model.py:728-904 for the device operators: the states
    they accept as ``operator_states_quad`` (see ``OperatorFluidStates``).
    Limiting the operator states (``limiter_func``, ``entropy_min``) is
    outside the device path and raises ``NotImplementedError``."""
    from mirgecom_amd.dof_desc import DD_VOLUME_ALL, DISCR_TAG_BASE
    from mirgecom_amd._opcommon import check_dd
    check_dd(DD_VOLUME_ALL if dd is None else dd)
    if limiter_func is not None or entropy_min is not None:
        raise NotImplementedError("limiting of the operator states is not part "
                                  "of the device path")
    return OperatorFluidStates(
        dcoll, volume_state, gas_model, boundaries,
        DISCR_TAG_BASE if quadrature_tag is None else quadrature_tag,
        entropy_stable)
