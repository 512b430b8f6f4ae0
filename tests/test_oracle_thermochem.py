"""Pins of the oracle's uiuc_7sp thermochemistry by what the mathematics and
the mechanism fix (Cantera/pyrometheus are absent here, so the reference's
own comparisons test/test_eos.py:55-472 and test/test_chemistry.py:51-351
cannot run):

* NASA7 self-consistency: d(h)/dT = cp, d(s)/dT = cp/T, continuity at T_mid
* Newton temperature inversion: round trip T -> e(T) -> T
  (cf. test/test_eos.py:240-244)
* mass conservation sum_k W_k omega_k = 0 and element conservation
* detailed balance of the reversible R2 at its equilibrium composition
* zero-level pinning of negative concentrations
  (mirgecom/thermochemistry.py:87-94, test/test_eos.py:337-354)
* 0-D autoignition reaches 1800 K < T < 3200 K
  (acceptance of test/test_chemistry.py:258-260)
"""

import numpy as np
import pytest

from oracle import thermochem as tc

ELEMENTS = {"C": [2, 0, 1, 1, 0, 0, 0], "H": [4, 0, 0, 0, 2, 2, 0],
            "O": [0, 2, 2, 1, 1, 0, 0], "N": [0, 0, 0, 0, 0, 0, 2]}


def _random_state(n, seed, tlo=900.0, thi=2500.0):
    rng = np.random.default_rng(seed)
    y = rng.uniform(0.01, 1.0, (7, n))
    y /= y.sum(axis=0)
    t = rng.uniform(tlo, thi, n)
    rho = rng.uniform(0.1, 2.0, n)
    return rho, t, y


def test_nasa7_derivatives():
    t = np.linspace(350, 3000, 200)
    t = t[np.abs(t - tc.T_MID) > 1]
    h = tc.h_rt(t) * t                    # h / R
    eps = 1e-4
    dh = (tc.h_rt(t + eps) * (t + eps) - tc.h_rt(t - eps) * (t - eps)) / (2 * eps)
    assert np.abs(dh - tc.cp_r(t)).max() < 1e-6 * np.abs(tc.cp_r(t)).max()
    ds = (tc.s_r(t + eps) - tc.s_r(t - eps)) / (2 * eps)
    assert np.abs(ds - tc.cp_r(t) / t).max() < 1e-8
    del h


def test_nasa7_continuity_at_tmid():
    lo = tc.T_MID - 1e-9
    hi = tc.T_MID + 1e-9
    for f in (tc.cp_r, tc.h_rt, tc.s_r):
        a, b = f(np.array(lo)), f(np.array(hi))
        assert np.abs(a - b).max() < 2e-3 * np.abs(a).max()


@pytest.mark.parametrize("niter,tol", [(5, 1e-6), (10, 1e-9)])
def test_newton_round_trip(niter, tol):
    rho, t, y = _random_state(500, 1)
    e = tc.mixture_energy_mass(t, y)
    seed = t + np.random.default_rng(2).uniform(-100, 100, t.shape)
    t2 = tc.get_temperature(e, seed, y, niter)
    assert np.abs(t2 - t).max() < tol
    # cv is d e / d T
    de = (tc.mixture_energy_mass(t + 1e-3, y)
          - tc.mixture_energy_mass(t - 1e-3, y)) / 2e-3
    assert np.allclose(de, tc.mixture_cv_mass(t, y), rtol=1e-7)
    # cp - cv = R_mix
    assert np.allclose(tc.mixture_cp_mass(t, y) - tc.mixture_cv_mass(t, y),
                       tc.specific_gas_constant(y), rtol=1e-13)


def test_mass_and_element_conservation():
    rho, t, y = _random_state(1000, 3)
    src = tc.species_source_terms(rho, t, y)
    scale = np.abs(src).max(axis=0)
    assert np.abs(src.sum(axis=0)).max() < 1e-12 * scale.max()
    omega = tc.net_production_rates(rho, t, y)
    for el, counts in ELEMENTS.items():
        res = np.einsum("k,k...->...", np.array(counts, float), omega)
        assert np.abs(res).max() < 1e-12 * np.abs(omega).max(), el


def test_r2_detailed_balance():
    """At C_CO2 = K_c C_CO C_O2^0.5 the net rate of the reversible
    CO + 0.5 O2 <=> CO2 vanishes; K_c = exp(-dG/RT) (p0/RT)^dnu."""
    t = np.array([1200.0, 1800.0, 2600.0])
    kc = tc.equilibrium_constants(t)[1]
    g = tc.g_rt(t)
    kc_closed = np.exp(-(g[2] - g[3] - 0.5 * g[1])) * (tc.P_ATM / (tc.R_U * t))**-0.5
    assert np.allclose(kc, kc_closed, rtol=1e-14)
    c = np.zeros((7, 3))
    c[1] = 2.0e-3          # O2
    c[3] = 1.0e-3          # CO
    c[2] = kc * c[3] * np.sqrt(c[1])
    y = c * tc.W[:, None]
    rho = y.sum(axis=0)
    r = tc.rates_of_progress(rho, t, y / rho)
    fwd = tc.fwd_rate_coefficients(t)[1] * c[3] * np.sqrt(c[1])
    assert np.abs(r[1]).max() < 1e-12 * fwd.max()


def test_zero_level_pin():
    rho, t, y = _random_state(10, 4)
    y[0, :5] = -1e-3
    c = tc.concentrations(rho, y)
    assert np.all(c >= 0)
    r = tc.rates_of_progress(rho, t, y)
    assert np.all(np.isfinite(r))
    assert np.all(r[0, :5] == 0.0)     # R1 needs C2H4


def test_autoignition_bounds():
    """Constant-pressure 0-D reactor, stoichiometric ethylene/air
    (O2 + 3.76 N2) at 1200 K and 1 atm as in test/test_chemistry.py:51-110:
    ignites and burns to 1800 K < T < 3200 K within 1 ms."""
    x = np.zeros(7)
    x[0], x[1], x[6] = 1.0, 3.0, 3 * 3.76
    y = x * tc.W / np.sum(x * tc.W)
    t = 1200.0
    p = 101325.0
    h0 = tc.mixture_enthalpy_mass(np.array(t), y)
    dt = 1e-6
    for _ in range(1000):
        rho = tc.density(p, np.array(t), y)
        y = y + dt * tc.species_source_terms(rho, np.array(t), y) / rho
        for _ in range(3):
            t = t + float((h0 - tc.mixture_enthalpy_mass(np.array(t), y))
                          / tc.mixture_cp_mass(np.array(t), y))
    assert 1800.0 < t < 3200.0
    assert y[0] < 1e-3 * x[0]        # fuel consumed


def test_newton_tolerance_mode():
    """Tolerance mode: converged to the tolerance from far seeds, the last
    update within it; a fixed-count solve with as many iterations agrees."""
    rho, t, y = _random_state(300, 11, tlo=1000.0, thi=2600.0)
    e = tc.mixture_energy_mass(t, y)
    seed = np.full_like(t, 300.0)
    tt, it = tc.get_temperature_tol(e, seed, y, 1e-6)
    assert np.all(it > 0)
    assert np.abs(tt - t).max() < 1e-6
    nmax = int(it.max())
    tf = tc.get_temperature(e, seed, y, nmax)
    assert np.abs(tf - tt).max() < 1e-6
    _, it2 = tc.get_temperature_tol(e, seed, y, 1e-6, max_iter=1)
    assert np.all(it2 == -1)
