"""Pins of the oracle constant-pressure reactor (oracle/reactor.py): what the
physics fixes (enthalpy and element conservation at constant pressure),
RK4's fourth-order convergence, and the acceptance window the reference's
own reactor test uses (test/test_chemistry.py:258-260: 1800 K < T < 3200 K
after ignition)."""

import numpy as np

from oracle import reactor, thermochem as tc

ELEMENTS = {"C": [2, 0, 1, 1, 0, 0, 0], "H": [4, 0, 0, 0, 2, 2, 0],
            "O": [0, 2, 2, 1, 1, 0, 0], "N": [0, 0, 0, 0, 0, 0, 2]}


def _premixed(phi):
    x = np.zeros((7, len(phi)))
    x[0], x[1], x[6] = phi, 3.0, 3 * 3.76
    y = x * tc.W[:, None]
    return y / y.sum(axis=0)


def _elements(y):
    return {el: np.einsum("k,k...->...", np.array(c, float) / tc.W, y)
            for el, c in ELEMENTS.items()}


def test_conservation_and_ignition_window():
    phi = np.array([0.7, 1.0, 1.3])
    y0 = _premixed(phi)
    t0 = np.full(3, 1600.0)
    p = np.array([0.7, 1.0, 1.0]) * 101325.0
    h0 = tc.mixture_enthalpy_mass(t0, y0)
    t, y, t_ign = reactor.march(t0, p, y0, 1e-8, 1500)
    assert np.all(t_ign > 0) and np.all(t_ign < 1.5e-5)
    # the reference's acceptance window (stoichiometric mixture)
    assert 1800.0 < t[1] < 3200.0
    assert np.all(t > t0 + 1000.0)
    assert np.abs(tc.mixture_enthalpy_mass(t, y) - h0).max() < 1e-8 * np.abs(h0).max()
    e0, e1 = _elements(y0), _elements(y)
    for el in ELEMENTS:
        assert np.abs(e1[el] - e0[el]).max() < 1e-13, el
    assert np.abs(y.sum(axis=0) - 1.0).max() < 1e-13
    # fuel-lean burns its fuel, rich keeps some
    assert y[0, 0] < 1e-4 * y0[0, 0]


def test_rk4_fourth_order():
    """Halving dt cuts the error of the state at a fixed time 16-fold.  The
    march starts from a partly reacted state: at Y = 0 the fractional
    reaction orders (0.5, 0.65, 0.75) make the rates non-smooth."""
    y0 = _premixed(np.array([1.0]))
    t0 = np.array([1500.0])
    p = np.array([101325.0])
    t0, y0, _ = reactor.march(t0, p, y0, 1e-9, 2000)
    tend = 4e-6
    ref = reactor.march(t0, p, y0, tend / 800, 800)
    errs = []
    for n in (25, 50):
        t, y, _ = reactor.march(t0, p, y0, tend / n, n)
        errs.append(np.abs(y - ref[1]).max())
    rate = np.log2(errs[0] / errs[1])
    assert 3.5 < rate < 4.6, rate


def test_substep_matches_march():
    y0 = _premixed(np.array([0.8, 1.3]))
    t0 = np.array([1700.0, 1450.0])
    p = np.array([101325.0, 2e5])
    states = np.concatenate([t0[:, None], p[:, None], y0.T, np.full((2, 1), 1e-8)], 1)
    dy = reactor.substep(states, 1e-8, 4)
    _, y, _ = reactor.march(t0, p, y0, 2.5e-9, 4)
    assert np.abs(dy - (y - y0).T).max() < 1e-16
