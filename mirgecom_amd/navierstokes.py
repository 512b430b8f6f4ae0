"""``ns_operator``, ``grad_cv_operator``, ``grad_t_operator`` on the GPU.

Call surface of mirgecom/navierstokes.py:129-575.  One RHS is two HIP
launches: ``mg_ns_gradient`` (the DG gradients of all conserved fields and
of T, central flux inside, exterior value on boundaries) and ``mg_ns_rhs``
(viscous and inviscid volume fluxes, face fluxes with the gathered
neighbour traces and gradients, weak divergence, lift, inverse mass)."""

from __future__ import annotations

import torch

from mirgecom_amd import _lib
from mirgecom_amd._opcommon import (
    Prepared, check_dd, check_unsupported, warn_mixture_limiter)
from mirgecom_amd.dof_desc import DD_VOLUME_ALL, DISCR_TAG_BASE
from mirgecom_amd.flux import num_flux_central
from mirgecom_amd.fluid import ConservedVars
from mirgecom_amd.inviscid import flux_kind, inviscid_facial_flux_rusanov
from mirgecom_amd.viscous import viscous_facial_flux_central


def _check_grad_flux(numerical_flux_func):
    if numerical_flux_func is not num_flux_central:
        raise NotImplementedError("the device gradient uses num_flux_central")


def _gradients(dcoll, gas_model, boundaries, state, time):
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
    """mirgecom/navierstokes.py:129-235: ConservedVars of ``(F, dim, K, Np)``."""
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg, operator_states_quad,
                      limiter_func, entropy_min)
    _check_grad_flux(numerical_flux_func)
    _, grad = _gradients(dcoll, gas_model, boundaries, state, time)
    return ConservedVars(state.dim, grad[:-1])


def grad_t_operator(dcoll, gas_model, boundaries, state, *, time=0.0,
                    numerical_flux_func=num_flux_central,
                    quadrature_tag=DISCR_TAG_BASE, dd=DD_VOLUME_ALL,
                    comm_tag=None, operator_states_quad=None,
                    limiter_func=None, entropy_min=None, use_esdg=False):
    """mirgecom/navierstokes.py:238-347: ``(dim, K, Np)``."""
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg, operator_states_quad,
                      limiter_func, entropy_min)
    _check_grad_flux(numerical_flux_func)
    _, grad = _gradients(dcoll, gas_model, boundaries, state, time)
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
    350-575, fused inviscid path).

    ``precision`` (not in the reference signature): ``"float64"`` (default)
    or ``"mixed"`` -- FP32 contractions and FP32 nodal fluxes with FP64
    physics, FP64 state and FP64 RHS (mg_ns_flux_mixed / mg_ns_div_mixed)."""
    if not state.is_viscous:
        raise ValueError("Navier-Stokes operator expects viscous gas model.")
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg, operator_states_quad,
                      limiter_func, entropy_min)
    if inviscid_fluid_operator is not None:
        raise NotImplementedError("a separate inviscid operator is not "
                                  "supported; the inviscid terms are fused")
    if viscous_numerical_flux_func is not viscous_facial_flux_central:
        raise NotImplementedError("the device path uses "
                                  "viscous_facial_flux_central")
    _check_grad_flux(gradient_numerical_flux_func)
    warn_mixture_limiter(state, limiter_func)
    kind = flux_kind(inviscid_numerical_flux_func)
    inv = 1 if inviscid_terms_on else 0
    if precision not in ("float64", "mixed"):
        raise ValueError("precision must be 'float64' or 'mixed'")
    if precision == "mixed":
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
    if return_gradients:
        return out, ConservedVars(state.dim, grad[:-1]), grad[-1]
    return out
