"""``euler_operator`` on the GPU (call surface of mirgecom/euler.py:313-416).

The right-hand side ``-div F_I`` -- volume flux, weak divergence, face-trace
gather, numerical flux, lift and inverse mass -- is one HIP kernel launch
(``mg_euler_rhs``)."""

from __future__ import annotations

from warnings import warn

import torch

from mirgecom_amd import _lib
from mirgecom_amd._opcommon import (
    Prepared, check_dd, check_unsupported, warn_mixture_limiter)
from mirgecom_amd.dof_desc import DD_VOLUME_ALL, DISCR_TAG_BASE, DISCR_TAG_QUAD
from mirgecom_amd.fluid import ConservedVars
from mirgecom_amd.inviscid import flux_kind, inviscid_facial_flux_rusanov


def euler_operator(dcoll, state, gas_model, boundaries, time=0.0,
                   inviscid_numerical_flux_func=None,
                   quadrature_tag=DISCR_TAG_BASE, dd=DD_VOLUME_ALL,
                   comm_tag=None, use_esdg=False, operator_states_quad=None,
                   entropy_conserving_flux_func=None, limiter_func=None,
                   entropy_min=None, precision="float64"):
    r"""RHS of the Euler equations, :math:`\dot{q} = -\nabla\cdot F`.

    Same arguments and errors as the reference (mirgecom/euler.py:313-372).
    ``quadrature_tag=DISCR_TAG_QUAD`` (overintegration) runs
    ``mg_euler_rhs_quad``: fluxes formed and integrated at the quadrature
    points of ``dcoll.quadrature``; with ``use_esdg=True`` as well, the
    entropy-stable flux-differencing operator (``mg_euler_rhs_esdg``,
    mirgecom/euler.py:108-310; ideal gas).  Limiters are outside the device
    path and raise ``NotImplementedError``.

    ``precision`` (not in the reference signature): ``"float64"`` (default)
    or ``"mixed"`` -- nodal fluxes held relative to each element's first node
    and contracted in FP32, physics, face fluxes, state and RHS in FP64
    (``mg_euler_rhs_mixed``; nodal operator only)."""
    if precision not in ("float64", "mixed"):
        raise ValueError("precision must be 'float64' or 'mixed'")
    if precision == "mixed" and (use_esdg or quadrature_tag == DISCR_TAG_QUAD):
        raise NotImplementedError("the mixed-precision Euler path is the nodal "
                                  "operator")
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg, operator_states_quad,
                      limiter_func, entropy_min, allow_quad=True,
                      state=state, boundaries=boundaries)
    warn_mixture_limiter(state, limiter_func)
    if use_esdg:
        # the volume points take operator_states_quad's volume state when
        # given (euler.py:159-171): entropy-stable projected only if those
        # states were made with entropy_stable=True
        es_volume = (operator_states_quad.entropy_stable
                     if operator_states_quad is not None else True)
        return _euler_esdg(dcoll, state, gas_model, boundaries, time,
                           inviscid_numerical_flux_func,
                           entropy_conserving_flux_func, es_volume=es_volume)
    if inviscid_numerical_flux_func is None:
        warn("inviscid_numerical_flux_func unspecified, defaulting to "
             "inviscid_facial_flux_rusanov.", stacklevel=2)
        inviscid_numerical_flux_func = inviscid_facial_flux_rusanov
    kind = flux_kind(inviscid_numerical_flux_func)
    if quadrature_tag == DISCR_TAG_QUAD:
        return _euler_quad(dcoll, state, gas_model, boundaries, time, kind)
    prep = Prepared(dcoll, state, gas_model, boundaries, time)
    rhs = torch.empty_like(prep.q)
    entry = "mg_euler_rhs_mixed" if precision == "mixed" else "mg_euler_rhs"
    _lib.call(entry, prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
              _lib.ptr(prep.temp), _lib.ptr(prep.qb), _lib.ptr(prep.tb), kind,
              _lib.ptr(rhs), prep.stream)
    return ConservedVars(state.dim, rhs)


def entropy_stable_euler_operator(
        dcoll, gas_model, state, boundaries, time=0.0,
        inviscid_numerical_flux_func=None, entropy_conserving_flux_func=None,
        operator_states_quad=None, dd=DD_VOLUME_ALL, quadrature_tag=None,
        comm_tag=None, entropy_min=None, limiter_func=None):
    """mirgecom/euler.py:108-310: the flux-differencing RHS, i.e.
    ``euler_operator(..., use_esdg=True)`` (``mg_euler_rhs_esdg``); built for
    the ideal gas with ``quadrature_tag=DISCR_TAG_QUAD``."""
    if quadrature_tag is None:
        quadrature_tag = DISCR_TAG_BASE
    return euler_operator(
        dcoll, state, gas_model, boundaries, time=time,
        inviscid_numerical_flux_func=inviscid_numerical_flux_func,
        quadrature_tag=quadrature_tag, dd=dd, comm_tag=comm_tag, use_esdg=True,
        operator_states_quad=operator_states_quad,
        entropy_conserving_flux_func=entropy_conserving_flux_func,
        limiter_func=limiter_func, entropy_min=entropy_min)


def _esdg_for_ns(dcoll, state, gas_model, boundaries, time, es_volume=False):
    """entropy_stable_euler_operator as ns_operator calls it
    (mirgecom/navierstokes.py:564-570): with ns_operator's operator states,
    whose volume state is the plain projected one unless the caller's
    states were made with entropy_stable=True."""
    return _euler_esdg(dcoll, state, gas_model, boundaries, time, None, None,
                       es_volume=es_volume)


def _euler_quad(dcoll, state, gas_model, boundaries, time, kind):
    from mirgecom_amd._opcommon import quad_boundary_states
    from mirgecom_amd.device import get_device_discretization
    mix = bool(getattr(gas_model.eos, "is_mixture", False))
    q = state.cv.q
    if not q.is_cuda or q.dtype != torch.float64:
        raise TypeError("the state must be a float64 tensor on the GPU")
    dev = get_device_discretization(dcoll, q.device)
    q = q.contiguous()
    temp = state.temperature.contiguous() if mix else None
    qbq, tbq = quad_boundary_states(dcoll, dev, q, state.dim, gas_model,
                                    boundaries, time, temp)
    rhs = torch.empty_like(q)
    niter = gas_model.eos.temperature_niter if mix else 0
    _lib.call("mg_euler_rhs_quad", dev.mesh,
              gas_model.device_params(state.cv.nspecies, with_transport=False),
              dev.quad_ops(), _lib.ptr(q), _lib.ptr(temp), _lib.ptr(qbq),
              _lib.ptr(tbq), niter, kind, _lib.ptr(rhs), _lib.stream_handle())
    return ConservedVars(state.dim, rhs)


def _euler_esdg(dcoll, state, gas_model, boundaries, time, numerical_flux_func,
                entropy_conserving_flux_func, es_volume=True):
    """The flux-differencing RHS: Chandrashekar's two-point flux for the
    ideal gas (``mg_euler_rhs_esdg``), Renac's for the uiuc_7sp mixture
    (``mg_euler_rhs_esdg_mixture``).  ``es_volume=False``: the volume points
    take the plain projected state (what ns_operator hands its inviscid
    operator, mirgecom/navierstokes.py:564-570) instead of the entropy-stable
    projected one."""
    from mirgecom_amd._opcommon import quad_boundary_states
    from mirgecom_amd.device import get_device_discretization
    from mirgecom_amd.inviscid import (
        entropy_conserving_flux_chandrashekar, entropy_conserving_flux_renac,
        entropy_stable_inviscid_facial_flux,
        entropy_stable_inviscid_facial_flux_rusanov)
    mix = bool(getattr(gas_model.eos, "is_mixture", False))
    builtin = entropy_conserving_flux_renac if mix else \
        entropy_conserving_flux_chandrashekar
    if entropy_conserving_flux_func is None:
        warn("No entropy_conserving_flux_func was given for ESDG. Setting EC "
             f"flux to {builtin.__name__}.", stacklevel=3)
    elif entropy_conserving_flux_func is not builtin:
        raise NotImplementedError(f"{builtin.__name__} is the built-in "
                                  "entropy-conserving flux for this gas")
    if numerical_flux_func is None:
        warn("No inviscid_numerical_flux_func was given for ESDG. Automatically "
             "setting facial flux to entropy-stable Rusanov "
             "(entropy_stable_inviscid_facial_flux_rusanov).", stacklevel=3)
        numerical_flux_func = entropy_stable_inviscid_facial_flux_rusanov
    if numerical_flux_func is entropy_stable_inviscid_facial_flux_rusanov:
        dissipation = 1
    elif numerical_flux_func is entropy_stable_inviscid_facial_flux:
        dissipation = 0
    else:
        raise NotImplementedError("ESDG takes entropy_stable_inviscid_facial_flux"
                                  "_rusanov or entropy_stable_inviscid_facial_flux")
    q = state.cv.q
    if not q.is_cuda or q.dtype != torch.float64:
        raise TypeError("the state must be a float64 tensor on the GPU")
    dev = get_device_discretization(dcoll, q.device)
    q = q.contiguous()
    ev = torch.empty_like(q)
    rhs = torch.empty_like(q)
    if mix:
        temp = state.temperature.contiguous()
        qbq, tbq = quad_boundary_states(dcoll, dev, q, state.dim, gas_model,
                                        boundaries, time, temp,
                                        entropy_stable=True)
        _lib.call("mg_euler_rhs_esdg_mixture", dev.mesh,
                  gas_model.device_params(state.cv.nspecies, with_transport=False),
                  dev.quad_ops(), dev.esdg_ops(), _lib.ptr(q), _lib.ptr(temp),
                  _lib.ptr(qbq), _lib.ptr(tbq), gas_model.eos.temperature_niter,
                  dissipation, 1 if es_volume else 0, _lib.ptr(ev), _lib.ptr(rhs),
                  _lib.stream_handle())
        return ConservedVars(state.dim, rhs)
    qbq, _ = quad_boundary_states(dcoll, dev, q, state.dim, gas_model,
                                  boundaries, time)
    _lib.call("mg_euler_rhs_esdg", dev.mesh,
              gas_model.device_params(state.cv.nspecies, with_transport=False),
              dev.quad_ops(), dev.esdg_ops(), _lib.ptr(q), _lib.ptr(qbq),
              dissipation, _lib.ptr(ev), _lib.ptr(rhs), _lib.stream_handle())
    return ConservedVars(state.dim, rhs)
