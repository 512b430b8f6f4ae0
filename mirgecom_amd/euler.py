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
from mirgecom_amd.dof_desc import DD_VOLUME_ALL, DISCR_TAG_BASE
from mirgecom_amd.fluid import ConservedVars
from mirgecom_amd.inviscid import flux_kind, inviscid_facial_flux_rusanov


def euler_operator(dcoll, state, gas_model, boundaries, time=0.0,
                   inviscid_numerical_flux_func=None,
                   quadrature_tag=DISCR_TAG_BASE, dd=DD_VOLUME_ALL,
                   comm_tag=None, use_esdg=False, operator_states_quad=None,
                   entropy_conserving_flux_func=None, limiter_func=None,
                   entropy_min=None):
    r"""RHS of the Euler equations, :math:`\dot{q} = -\nabla\cdot F`.

    Same arguments and errors as the reference (mirgecom/euler.py:313-372);
    ESDG, overintegration and limiters are outside the device path and raise
    ``NotImplementedError``."""
    check_dd(dd)
    check_unsupported(quadrature_tag, use_esdg, operator_states_quad,
                      limiter_func, entropy_min)
    warn_mixture_limiter(state, limiter_func)
    if inviscid_numerical_flux_func is None:
        warn("inviscid_numerical_flux_func unspecified, defaulting to "
             "inviscid_facial_flux_rusanov.", stacklevel=2)
        inviscid_numerical_flux_func = inviscid_facial_flux_rusanov
    kind = flux_kind(inviscid_numerical_flux_func)
    prep = Prepared(dcoll, state, gas_model, boundaries, time)
    rhs = torch.empty_like(prep.q)
    _lib.call("mg_euler_rhs", prep.dev.mesh, prep.gas, _lib.ptr(prep.q),
              _lib.ptr(prep.temp), _lib.ptr(prep.qb), _lib.ptr(prep.tb), kind,
              _lib.ptr(rhs), prep.stream)
    return ConservedVars(state.dim, rhs)
