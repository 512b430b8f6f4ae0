"""The pure-species transport fits of the uiuc_7sp mixture
(mirgecom_amd/transport_fits.py; host-side data of MixtureAveragedTransport,
mirgecom/transport.py:324-502).  Cantera, which makes these fits for the
reference, is absent: the kinetic theory is checked against the tabulated
Lennard-Jones collision integrals, the Chapman-Enskog relation between
self-diffusion and viscosity, and measured gas properties at 300 K; the fits
against the kinetic theory they approximate."""

import math

import numpy as np
import pytest

from mirgecom_amd import mechanisms as mech
from mirgecom_amd import transport_fits as tf


@pytest.mark.parametrize("tstar,om11,om22", [
    # Lennard-Jones 12-6 collision integrals as tabulated in Bird, Stewart &
    # Lightfoot, Transport Phenomena (2nd ed.), table E.2 (Omega_D, Omega_mu)
    (0.3, 2.649, 2.840), (1.0, 1.440, 1.593), (2.0, 1.075, 1.176),
    (5.0, 0.8422, 0.9256), (10.0, 0.7424, 0.8242)])
def test_collision_integrals_vs_table(tstar, om11, om22):
    assert tf.omega11(tstar) == pytest.approx(om11, rel=3e-3)
    assert tf.omega22(tstar) == pytest.approx(om22, rel=3e-3)


def test_dipole_terms():
    """Brokaw's polar terms: +0.19 / +0.2 delta*^2 / T*."""
    assert tf.omega11(2.0, 0.5) - tf.omega11(2.0) == pytest.approx(
        0.19 * 0.25 / 2.0, rel=1e-12)
    assert tf.omega22(2.0, 0.5) - tf.omega22(2.0) == pytest.approx(
        0.2 * 0.25 / 2.0, rel=1e-12)


def test_chapman_enskog_self_diffusion():
    """rho D_kk / mu_k = 6/5 Omega22*/Omega11* (same T*, delta*)."""
    sp = tf.SpeciesParameters()
    t = np.array([300.0, 1000.0, 2500.0])
    mu, lam, pd = tf.kinetic_theory(t, sp)
    for k in range(sp.n):
        tstar = tf.K_B * t / sp.eps[k]
        delta = sp.delta_ij[k, k]
        ratio = sp.mw[k] * pd[k, k] / (tf.R_U * t * mu[k])
        want = 1.2 * tf.omega22(tstar, delta) / tf.omega11(tstar, delta)
        np.testing.assert_allclose(ratio, want, rtol=1e-12)


def test_h2o_reduced_dipole():
    """delta* = mu_d^2 / (2 4 pi eps0 eps sigma^3) for H2O (1.844 D,
    572.4 K, 2.605 A) and zero for every pair with a non-polar partner."""
    sp = tf.SpeciesParameters()
    k = mech.SPECIES.index("H2O")
    mu_d = 1.844 * 3.33564095198152e-30
    want = mu_d ** 2 / (2 * 4 * math.pi * 8.8541878128e-12
                        * 572.4 * 1.380649e-23 * (2.605e-10) ** 3)
    assert sp.delta_ij[k, k] == pytest.approx(want, rel=1e-12)
    others = [j for j in range(sp.n) if j != k]
    assert np.all(sp.delta_ij[k, others] == 0.0)
    assert np.all(sp.delta_ij[np.ix_(others, others)] == 0.0)


# measured values at 300 K and 1 atm (CRC Handbook / Marrero & Mason 1972)
_MU = {"N2": 1.79e-5, "O2": 2.07e-5, "H2": 0.896e-5, "CO2": 1.50e-5,
       "CO": 1.77e-5, "C2H4": 1.03e-5}
_LAM = {"N2": 0.0259, "O2": 0.0263, "H2": 0.183, "CO": 0.0250}
_D = {("O2", "N2"): 0.210e-4, ("H2", "N2"): 0.780e-4, ("CO2", "N2"): 0.165e-4}


def test_against_measured_gas_data():
    mu, lam, pd = tf.kinetic_theory(np.array([300.0]))
    i = mech.SPECIES.index
    for s, v in _MU.items():
        assert mu[i(s), 0] == pytest.approx(v, rel=0.03), s
    for s, v in _LAM.items():
        assert lam[i(s), 0] == pytest.approx(v, rel=0.05), s
    for (a, b), v in _D.items():
        assert pd[i(a), i(b), 0] / 101325.0 == pytest.approx(v, rel=0.06), (a, b)
    np.testing.assert_allclose(pd[:, :, 0], pd[:, :, 0].T, rtol=1e-15)


def test_fits_reproduce_kinetic_theory():
    visc, cond, diff = tf.transport_fits()
    tmin, tmax = mech.phase_temperature_range()
    assert (tmin, tmax) == (300.0, 3500.0)
    t = np.linspace(tmin, tmax, 301)
    lt = np.log(t)
    mu, lam, pd = tf.kinetic_theory(t)
    pv = np.polynomial.polynomial.polyval
    assert np.abs(np.sqrt(t) * pv(lt, visc.T) ** 2 / mu - 1).max() < 2e-3
    assert np.abs(np.sqrt(t) * pv(lt, cond.T) / lam - 1).max() < 1e-2
    for a in range(7):
        for b in range(7):
            assert np.abs(t ** 1.5 * pv(lt, diff[a, b]) / pd[a, b] - 1).max() < 3e-3
            assert np.array_equal(diff[a, b], diff[b, a])


def test_packed_layout():
    """308 doubles: the device struct's size; Wilke factors in place."""
    p = tf.packed_transport()
    assert p.size == 308
    visc, cond, diff = tf.cached_fits()
    assert np.array_equal(p[:35], visc.ravel())
    assert np.array_equal(p[35:70], cond.ravel())
    assert np.array_equal(p[70:75], diff[0, 1])
    assert np.array_equal(p[170:175], diff[5, 6])
    assert np.array_equal(p[175:180], diff[0, 0])
    w = mech.molecular_weights()
    phi_a = p[210:259].reshape(7, 7)
    phi_b = p[259:308].reshape(7, 7)
    assert phi_a[2, 3] == pytest.approx((1 + w[2] / w[3]) ** -0.5 / math.sqrt(8))
    assert phi_b[2, 3] == pytest.approx((w[3] / w[2]) ** 0.25)
    np.testing.assert_allclose(np.diag(phi_a) * (1 + np.diag(phi_b)) ** 2, 1.0,
                               rtol=1e-15)
