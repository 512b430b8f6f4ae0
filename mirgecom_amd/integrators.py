"""Time integration (mirgecom/integrators/explicit_rk.py:31-38).

``rk4_step`` has the reference's signature and works on any state with
``+``/``*``.  For tensor-backed ``ConservedVars`` on the device with a scalar
``dt`` the stage updates and the FP64 accumulation run in the fused HIP
kernel ``mg_rk_axpy`` (one pass per stage, no temporaries).  A per-node
``dt`` (``local_dt`` marching: a ``(K, Np)`` tensor that broadcasts over the
fields) takes the generic tensor arithmetic."""

from __future__ import annotations

import numbers

import numpy as np
import torch

from mirgecom_amd import _lib
from mirgecom_amd.fluid import ConservedVars


def _axpy(y, k, a, out, acc=None, w=0.0):
    n = y.numel()
    _lib.check(_lib.load().mg_rk_axpy(
        n, _lib.ptr(y), _lib.ptr(k), float(a), _lib.ptr(out),
        _lib.ptr(acc) if acc is not None else None, float(w),
        _lib.stream_handle()))


def _is_scalar(dt):
    return isinstance(dt, numbers.Real) or (
        isinstance(dt, np.ndarray) and dt.ndim == 0)


def _stage_rhs(rhs, t, state, like):
    """``rhs(t, state).q`` as a contiguous float64 tensor of ``like``'s shape
    on its device (the fused update reads it by raw pointer)."""
    k = rhs(t, state)
    k = k.q if isinstance(k, ConservedVars) else k
    if not isinstance(k, torch.Tensor):
        raise TypeError("the RHS of a device state must be a tensor-backed "
                        "ConservedVars")
    if k.dtype != torch.float64 or k.device != like.device:
        raise TypeError(f"RHS is {k.dtype} on {k.device}; the state is "
                        f"{like.dtype} on {like.device}")
    if tuple(k.shape) != tuple(like.shape):
        raise ValueError(f"RHS shape {tuple(k.shape)} differs from the state "
                         f"shape {tuple(like.shape)}")
    return k.contiguous()


def rk4_step(state, t, dt, rhs):
    """Classical RK4: ``k1..k4 = rhs(t + c dt, state + a dt k)``."""
    if not (isinstance(state, ConservedVars) and state.q.is_cuda
            and state.q.dtype == torch.float64 and _is_scalar(dt)):
        k1 = rhs(t, state)
        k2 = rhs(t + dt / 2, state + dt / 2 * k1)
        k3 = rhs(t + dt / 2, state + dt / 2 * k2)
        k4 = rhs(t + dt, state + dt * k3)
        return state + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    dt = float(dt)
    y = state.q.contiguous()
    acc = y.clone()
    tmp = torch.empty_like(y)
    k = _stage_rhs(rhs, t, state, y)
    _axpy(y, k, dt / 2, tmp, acc, dt / 6)
    k = _stage_rhs(rhs, t + dt / 2, ConservedVars(state.dim, tmp), y)
    tmp2 = torch.empty_like(y)
    _axpy(y, k, dt / 2, tmp2, acc, dt / 3)
    k = _stage_rhs(rhs, t + dt / 2, ConservedVars(state.dim, tmp2), y)
    _axpy(y, k, dt, tmp, acc, dt / 3)
    k = _stage_rhs(rhs, t + dt, ConservedVars(state.dim, tmp), y)
    _axpy(acc, k, dt / 6, acc)
    return ConservedVars(state.dim, acc)
