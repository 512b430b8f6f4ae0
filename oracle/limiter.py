"""Oracle restatement of the bound-preserving limiter (test infrastructure
only): mirgecom/limiter.py:39-125, the linear scaling of Zhang & Shu
(J. Comput. Phys. 2011, sec. 2.3, eq. 2.9) about the cell average.

Pinned by the positivity and upper-bound checks of test/test_limiter.py:79-131
and by what the scaling fixes: cell averages are unchanged, fields inside the
bounds are unchanged, limited fields lie inside the bounds.
"""

from __future__ import annotations

import numpy as np


def elementwise_integral(disc, field):
    """``sum_ij J M_ij u_j`` per element: ``(..., K, Np)`` -> ``(..., K)``."""
    return np.einsum("k,ij,...kj->...k", disc.J, disc.el.mass, field)


def bound_preserving_limiter(disc, field, mmin=0.0, mmax=None,
                             modify_average=False):
    """mirgecom/limiter.py:39-125 on ``(..., K, Np)`` fields."""
    field = np.asarray(field, dtype=np.float64)
    vols = np.abs(elementwise_integral(disc, np.ones_like(field)))
    avgs = elementwise_integral(disc, field) / vols
    if modify_average:
        avgs = np.where(avgs > mmin, avgs, mmin)
    mmin_i = field.min(axis=-1)
    theta = np.minimum(1.0, np.where(mmin_i < mmin,
                                     np.abs((mmin - avgs) / (mmin_i - avgs + 1e-13)),
                                     1.0))
    if mmax is not None:
        if modify_average:
            avgs = np.where(avgs > mmax, mmax, avgs)
        mmax_i = field.max(axis=-1)
        theta = np.minimum(theta, np.where(mmax_i > mmax,
                                           np.abs((mmax - avgs) / (mmax_i - avgs + 1e-13)),
                                           1.0))
    return theta[..., None] * (field - avgs[..., None]) + avgs[..., None]
