"""Bound-preserving limiter: call surface of mirgecom/limiter.py:39-125.

``bound_preserving_limiter`` scales each element's nodal values about the
cell average (Zhang & Shu, J. Comput. Phys. 230 (2011), eq. 2.9) so that they
lie inside ``[mmin, mmax]``.  It runs in the HIP kernel ``mg_bound_limiter``
(one thread per element, elements staged through LDS), on every field of a
``(..., K, Np)`` tensor in one launch.
"""

from __future__ import annotations

import numpy as np
import torch

from mirgecom_amd import _lib
from mirgecom_amd.dof_desc import DD_VOLUME_ALL, as_dofdesc

_LIMIT_MAX = 1
_LIMIT_MODIFY_AVERAGE = 2


def _mass_weights(dev):
    """Row sums of the reference mass matrix on the device: the cell average
    of an affine simplex is ``sum_j w_j u_j / sum_j w_j``."""
    w = getattr(dev, "_limiter_weights", None)
    if w is None:
        re = dev.host.refel
        w = torch.as_tensor(np.ascontiguousarray(re.mass.sum(axis=0)),
                            dtype=torch.float64, device=dev.device)
        dev._limiter_weights = w
    return w


def bound_preserving_limiter(dcoll, field, mmin=0.0, mmax=None,
                             modify_average=False, dd=None):
    """mirgecom/limiter.py:39-125.

    ``field`` is a volume field ``(K, Np)`` or a stack of them
    ``(..., K, Np)`` (FP64, on the GPU), or an object array of such tensors
    (the reference's array containers).  Returns the limited field(s); with
    ``modify_average`` the cell average is first clipped into the bounds.
    """
    if dd is not None and as_dofdesc(dd) != DD_VOLUME_ALL:
        raise NotImplementedError("limiting on a volume other than "
                                  "DD_VOLUME_ALL")
    if isinstance(field, np.ndarray) and field.dtype == object:
        out = np.empty(field.shape, dtype=object)
        for i in np.ndindex(field.shape):
            out[i] = bound_preserving_limiter(dcoll, field[i], mmin, mmax,
                                              modify_average, dd)
        return out
    k, npe = dcoll.nelements, dcoll.refel.nnodes
    if not isinstance(field, torch.Tensor) or field.device.type != "cuda":
        raise TypeError("field must be a tensor on the GPU")
    if field.dtype != torch.float64:
        raise TypeError("field must be float64")
    if tuple(field.shape[-2:]) != (k, npe):
        raise ValueError(f"field shape {tuple(field.shape)} does not end in "
                         f"(K, Np) = ({k}, {npe})")
    from mirgecom_amd.device import get_device_discretization
    dev = get_device_discretization(dcoll)
    src = field.contiguous()
    out = torch.empty_like(src)
    nf = src.numel() // (k * npe)
    flags = ((_LIMIT_MAX if mmax is not None else 0)
             | (_LIMIT_MODIFY_AVERAGE if modify_average else 0))
    _lib.call("mg_bound_limiter", npe, k, nf, _lib.ptr(_mass_weights(dev)),
              _lib.ptr(src), _lib.ptr(out), float(mmin),
              float(mmax) if mmax is not None else 0.0, flags,
              _lib.stream_handle())
    return out


__all__ = ["bound_preserving_limiter"]
