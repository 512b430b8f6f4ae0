"""Pins of the oracle's bound-preserving limiter (oracle/limiter.py):
test/test_limiter.py:79-131 restated, plus the properties the scaling fixes."""

import numpy as np
import pytest

from oracle.limiter import bound_preserving_limiter, elementwise_integral
from tests.cases import Problem


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_positivity(dim, order):
    """test/test_limiter.py:79-104: x - 0.001 on [0, 0.5]^d, 2 cells/axis."""
    pb = Problem(dim, order, (2,) * dim, 0.0, 0.5, (False,) * dim)
    field = pb.x[0] - 0.001
    lim = bound_preserving_limiter(pb.odisc, field, mmin=0.0)
    assert lim.min() > -1e-13


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_upper_bound(dim, order):
    """test/test_limiter.py:106-131: 0.5 + x + 0.01 limited to [0, 1]."""
    pb = Problem(dim, order, (2,) * dim, 0.0, 0.5, (False,) * dim)
    field = 0.5 + pb.x[0] + 0.01
    lim = bound_preserving_limiter(pb.odisc, field, mmin=0.0, mmax=1.0)
    assert lim.max() < 1.0 + 1e-13


def test_averages_kept_and_inside_untouched():
    pb = Problem(2, 3, (3, 3), 0.0, 1.0, (False, False))
    rng = np.random.default_rng(0)
    field = 0.5 + 0.6 * np.sin(7 * pb.x[0]) * np.cos(5 * pb.x[1]) \
        + 0.01 * rng.standard_normal(pb.x[0].shape)
    lim = bound_preserving_limiter(pb.odisc, field, mmin=0.0, mmax=1.0)
    a0 = elementwise_integral(pb.odisc, field)
    a1 = elementwise_integral(pb.odisc, lim)
    assert np.abs(a1 - a0).max() < 1e-15 * np.abs(a0).max() * 10
    inside = (field.min(axis=-1) >= 0.0) & (field.max(axis=-1) <= 1.0)
    assert inside.any() and (~inside).any()
    # theta = 1 there: (u - avg) + avg, equal up to one rounding
    assert np.abs(lim[inside] - field[inside]).max() < 4e-16 * np.abs(field).max()
    avg = a0 / elementwise_integral(pb.odisc, np.ones_like(field))
    ok = (avg > 0.0) & (avg < 1.0)
    assert lim[ok].min() > -1e-12 and lim[ok].max() < 1.0 + 1e-12
