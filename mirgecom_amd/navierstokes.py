"""``ns_operator``, ``grad_cv_operator``, ``grad_t_operator`` on the GPU.

Call surface of mirgecom/navierstokes.py:129-575.  One RHS is two HIP
launches: ``mg_ns_gradient`` (the DG gradients of all conserved fields and
of T, central flux inside, exterior value on boundaries) and ``mg_ns_rhs``
(viscous and inviscid volume fluxes, face fluxes with the gathered
neighbour traces and gradients, weak divergence, lift, inverse mass).  With
``quadrature_tag=DISCR_TAG_QUAD`` the same two steps run on the quadrature
points (``mg_ns_gradient_quad``, ``mg_ns_rhs_quad``); an
``inviscid_fluid_operator`` (``entropy_stable_euler_operator`` with
``use_esdg=True``) replaces the fused inviscid terms."""

from __future__ import annotations

import torch

from mirgecom_amd import _lib
from mirgecom_amd._opcommon import (
    Prepared, check_dd, check_unsupported, warn_mixture_limiter)
from mirgecom_amd.dof_desc import DD_VOLUME_ALL, DISCR_TAG_BASE, DISCR_TAG_QUAD
from mirgecom_amd.flux import num_flux_central
from mirgecom_amd.fluid import ConservedVars
from mirgecom_amd.inviscid import flux_kind, inviscid_facial_flux_rusanov
from mirgecom_amd.viscous import viscous_facial_flux_central


def _check_grad_flux(numerical_flux_func):
    if numerical_flux_func is not num_flux_central:
        raise NotImplementedError("the device gradient uses num_flux_central")


class _QuadInputs:
    """Device pointers of one call on the quadrature points: the state, the
    nodal temperature (mixtures), the boundaries' exterior states at the
    boundary faces' points (prescribed and dummy boundaries)."""

    def __init__(self, dcoll, gas_model, state, boundaries, time):
        from mirgecom_amd.boundary import MengaldoBoundaryCondition
        from mirgecom_amd._opcommon import quad_boundary_states
        from mirgecom_amd.device import get_device_discretization
        from mirgecom_amd.utils import normalize_boundaries
        if any(isinstance(b, MengaldoBoundaryCondition)
               for b in normalize_boundaries(boundaries).values()):
            raise NotImplementedError(
                "the overintegrated Navier-Stokes operator takes prescribed and "
                "dummy boundaries; the wall family runs on the nodal operator")
        q = state.cv.q
        if not q.is_cuda or q.dtype != torch.float64:
            raise TypeError("the state must be a float64 tensor on the GPU")
        K, npe = dcoll.nelements, dcoll.refel.nnodes
        if tuple(q.shape[1:]) != (K, npe):
            raise ValueError(f"state shape {tuple(q.shape)} does not match the "
                             f"discretization ({K}, {npe})")
        self.dev = get_device_discretization(dcoll, q.device)
        self.q = q.contiguous()
        self.dim = state.dim
        mix = bool(getattr(gas_model.eos, "is_mixture", False))
        self.temp = state.temperature.contiguous() if mix else None
        self.qbq, self.tbq = quad_boundary_states(
            dcoll, self.dev, self.q, state.dim, gas_model, boundaries, time,
            self.temp)
        self.niter = gas_model.eos.temperature_niter if mix else 0
        self.gas = gas_model.device_params(state.cv.nspecies)
        self.stream = _lib.stream_handle()

    def gradients(self):
        nf = self.q.shape[0]
        grad = torch.empty((nf + 1, self.dim) + tuple(self.q.shape[1:]),
                           dtype=self.q.dtype, device=self.q.device)
        _lib.call("mg_ns_gradient_quad", self.dev.mesh, self.gas,
                  self.dev.quad_ops(), _lib.ptr(self.q), _lib.ptr(self.temp),
                  _lib.ptr(self.qbq), _lib.ptr(self.tbq), self.niter,
                  _lib.ptr(grad), self.stream)
        return grad

    def rhs(self, grad, kind, inv):
        rhs = torch.empty_like(self.q)
        _lib.call("mg_ns_rhs_quad", self.dev.mesh, self.gas, self.dev.quad_ops(),
                  _lib.ptr(self.q), _lib.ptr(self.temp), _lib.ptr(grad),
                  _lib.ptr(self.qbq), _lib.ptr(self.tbq), self.niter, kind, inv,
                  _lib.ptr(rhs), self.stream)
        return rhs


def _gradients(dcoll, gas_model, boundaries, state, time,
               quadrature_tag=DISCR_TAG_BASE):
    if quadrature_tag == DISCR_TAG_QUAD:
        qi = _QuadInputs(dcoll, gas_model, state, boundaries, time)
        return qi, qi.gradients()
    prep = Prepared(dcoll, state, gas_model, boundaries, time)
    nf = prep.q.shape[0]
    grad = torch.empty((nf + 1, prep.dim) + tuple(prep.q.shape[1:]),
                       dtype=prep.q.dtype, device=prep.q.device)
    _lib.call("mg_ns_gradient", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
              _lib.ptr(prep.temp), _lib.ptr(prep.qb), _lib.ptr(prep.tb),
              prep.wall, _lib.ptr(grad), prep.stream)
    return prep, grad


def grad_cv_operator(dcoll, gas_model, boundaries, state, *, time=0.0,
                     numerical_flux_func=num_flux_central,
                     quadrature_tag=DISCR_TAG_BASE, dd=DD_VOLUME_ALL,
                     comm_tag=None, operator_states_quad=None,
                     limiter_func=None, entropy_min=None, use_esdg=False):
    """mirgecom/navierstokes.py:129-235: ConservedVars of ``(F, dim, K, Np)``.
    ``quadrature_tag=DISCR_TAG_QUAD``: the weak gradient on the quadrature
    points (``mg_ns_gradient_quad``)."""
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg,
                      operator_states_quad, limiter_func, entropy_min,
                      allow_quad=True,
                      state=state, boundaries=boundaries)
    _check_grad_flux(numerical_flux_func)
    _, grad = _gradients(dcoll, gas_model, boundaries, state, time,
                         quadrature_tag)
    return ConservedVars(state.dim, grad[:-1])


def grad_t_operator(dcoll, gas_model, boundaries, state, *, time=0.0,
                    numerical_flux_func=num_flux_central,
                    quadrature_tag=DISCR_TAG_BASE, dd=DD_VOLUME_ALL,
                    comm_tag=None, operator_states_quad=None,
                    limiter_func=None, entropy_min=None, use_esdg=False):
    """mirgecom/navierstokes.py:238-347: ``(dim, K, Np)``.
    ``quadrature_tag=DISCR_TAG_QUAD``: the weak gradient on the quadrature
    points (``mg_ns_gradient_quad``)."""
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg,
                      operator_states_quad, limiter_func, entropy_min,
                      allow_quad=True,
                      state=state, boundaries=boundaries)
    _check_grad_flux(numerical_flux_func)
    _, grad = _gradients(dcoll, gas_model, boundaries, state, time,
                         quadrature_tag)
    return grad[-1]


def ns_operator(dcoll, gas_model, state, boundaries, *, time=0.0,
                inviscid_fluid_operator=None,
                inviscid_numerical_flux_func=inviscid_facial_flux_rusanov,
                gradient_numerical_flux_func=num_flux_central,
                viscous_numerical_flux_func=viscous_facial_flux_central,
                return_gradients=False, quadrature_tag=DISCR_TAG_BASE,
                dd=DD_VOLUME_ALL, comm_tag=None, limiter_func=None,
                entropy_min=None, operator_states_quad=None, use_esdg=False,
                grad_cv=None, grad_t=None, inviscid_terms_on=True,
                entropy_conserving_flux_func=None, precision="float64"):
    r"""RHS of the Navier-Stokes equations,
    :math:`\partial_t Q = \nabla\cdot(F_V - F_I)` (mirgecom/navierstokes.py:
    350-575).  The inviscid terms are fused into the viscous kernels unless
    ``inviscid_fluid_operator`` is given (or ``use_esdg=True``, which takes
    ``entropy_stable_euler_operator``; ideal gas, ``DISCR_TAG_QUAD``), whose
    RHS is then added.  ``quadrature_tag=DISCR_TAG_QUAD`` evaluates fluxes
    and integrals on the quadrature points (prescribed and dummy
    boundaries).

    ``precision`` (not in the reference signature): ``"float64"`` (default)
    or ``"mixed"`` -- FP32 contractions and FP32 nodal fluxes with FP64
    physics, FP64 state and FP64 RHS (mg_ns_flux_mixed / mg_ns_div_mixed)."""
    if not state.is_viscous:
        raise ValueError("Navier-Stokes operator expects viscous gas model.")
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg,
                      operator_states_quad, limiter_func, entropy_min,
                      allow_quad=True,
                      state=state, boundaries=boundaries)
    from mirgecom_amd.euler import entropy_stable_euler_operator, euler_operator
    if inviscid_fluid_operator is None and use_esdg:
        inviscid_fluid_operator = entropy_stable_euler_operator
    elif use_esdg and inviscid_fluid_operator is euler_operator:
        raise RuntimeError("Standard Euler operator is incompatible with ESDG.")
    if viscous_numerical_flux_func is not viscous_facial_flux_central:
        raise NotImplementedError("the device path uses "
                                  "viscous_facial_flux_central")
    _check_grad_flux(gradient_numerical_flux_func)
    warn_mixture_limiter(state, limiter_func)
    kind = flux_kind(inviscid_numerical_flux_func)
    # the fused inviscid terms, unless a separate inviscid operator adds them
    inv = 1 if (inviscid_terms_on and inviscid_fluid_operator is None) else 0
    if precision not in ("float64", "mixed"):
        raise ValueError("precision must be 'float64' or 'mixed'")
    if quadrature_tag == DISCR_TAG_QUAD:
        if precision == "mixed":
            raise NotImplementedError("the mixed-precision Navier-Stokes path "
                                      "is the nodal operator")
        qi = _QuadInputs(dcoll, gas_model, state, boundaries, time)
        if grad_cv is None or grad_t is None:
            grad = qi.gradients()
            if grad_cv is not None:
                grad[:-1] = grad_cv.q
            if grad_t is not None:
                grad[-1] = grad_t
        else:
            grad = torch.cat([grad_cv.q, grad_t[None]], dim=0).contiguous()
        rhs = qi.rhs(grad, kind, inv)
    elif precision == "mixed":
        if return_gradients or grad_cv is not None or grad_t is not None:
            raise NotImplementedError("the mixed-precision path computes the "
                                      "gradients internally")
        prep = Prepared(dcoll, state, gas_model, boundaries, time)
        nf, dim = prep.q.shape[0], prep.dim
        fnet = torch.empty((nf * dim,) + tuple(prep.q.shape[1:]),
                           dtype=torch.float32, device=prep.q.device)
        fnb = None
        if prep.qb is not None:
            fnb = torch.empty((nf, prep.qb.shape[1]), dtype=prep.q.dtype,
                              device=prep.q.device)
        g32 = None
        if state.is_mixture:
            # the mixture path stages FP32 gradients through this buffer
            g32 = torch.empty((nf + 1, dim) + tuple(prep.q.shape[1:]),
                              dtype=torch.float32, device=prep.q.device)
        _lib.call("mg_ns_flux_mixed", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
                  _lib.ptr(prep.temp), _lib.ptr(prep.qb), _lib.ptr(prep.tb),
                  prep.wall, inv, _lib.ptr(fnet), _lib.ptr(fnb), _lib.ptr(g32),
                  prep.stream)
        rhs = torch.empty_like(prep.q)
        _lib.call("mg_ns_div_mixed", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
                  _lib.ptr(prep.temp), _lib.ptr(fnet), _lib.ptr(prep.qb),
                  _lib.ptr(prep.tb), _lib.ptr(fnb), kind, inv, _lib.ptr(rhs),
                  prep.stream)
    elif grad_cv is None and grad_t is None:
        # fused path: gradients -> nodal net flux (one kernel), divergence
        prep = Prepared(dcoll, state, gas_model, boundaries, time)
        nf, dim = prep.q.shape[0], prep.dim
        shape = tuple(prep.q.shape[1:])
        fnet = torch.empty((nf * dim,) + shape, dtype=prep.q.dtype,
                           device=prep.q.device)
        fnb = None
        if prep.qb is not None:
            fnb = torch.empty((nf, prep.qb.shape[1]), dtype=prep.q.dtype,
                              device=prep.q.device)
        grad = None
        if return_gradients or state.is_mixture:
            # the mixture path stages its gradients through this buffer
            grad = torch.empty((nf + 1, dim) + shape, dtype=prep.q.dtype,
                               device=prep.q.device)
        _lib.call("mg_ns_flux", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
                  _lib.ptr(prep.temp), _lib.ptr(prep.qb), _lib.ptr(prep.tb),
                  prep.wall, inv, _lib.ptr(fnet), _lib.ptr(fnb), _lib.ptr(grad),
                  prep.stream)
        rhs = torch.empty_like(prep.q)
        _lib.call("mg_ns_div", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
                  _lib.ptr(prep.temp), _lib.ptr(fnet), _lib.ptr(prep.qb),
                  _lib.ptr(prep.tb), _lib.ptr(fnb), kind, inv, _lib.ptr(rhs),
                  prep.stream)
    else:
        if grad_cv is None or grad_t is None:
            prep, grad = _gradients(dcoll, gas_model, boundaries, state, time)
            if grad_cv is not None:
                grad[:-1] = grad_cv.q
            if grad_t is not None:
                grad[-1] = grad_t
        else:
            prep = Prepared(dcoll, state, gas_model, boundaries, time)
            grad = torch.cat([grad_cv.q, grad_t[None]], dim=0).contiguous()
        rhs = torch.empty_like(prep.q)
        _lib.call("mg_ns_rhs", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
                  _lib.ptr(prep.temp), _lib.ptr(grad), _lib.ptr(prep.qb),
                  _lib.ptr(prep.tb), prep.wall, kind, inv, _lib.ptr(rhs),
                  prep.stream)
    out = ConservedVars(state.dim, rhs)
    if inviscid_terms_on and inviscid_fluid_operator is not None:
        # mirgecom/navierstokes.py:564-570 (ESDG always takes this branch)
        if (inviscid_fluid_operator is entropy_stable_euler_operator
                and quadrature_tag == DISCR_TAG_QUAD):
            # the reference hands its own (plain projected) operator states
            # to the entropy-stable operator
            from mirgecom_amd.euler import _esdg_for_ns
            out = out + _esdg_for_ns(
                dcoll, state, gas_model, boundaries, time,
                es_volume=(operator_states_quad is not None
                           and operator_states_quad.entropy_stable))
        else:
            out = out + inviscid_fluid_operator(
                dcoll, state=state, gas_model=gas_model, boundaries=boundaries,
                time=time, dd=dd, comm_tag=comm_tag,
                quadrature_tag=quadrature_tag,
                operator_states_quad=operator_states_quad)
    if return_gradients:
        return out, ConservedVars(state.dim, grad[:-1]), grad[-1]
    return out
