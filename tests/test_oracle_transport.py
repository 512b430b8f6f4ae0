"""Pins of the oracle's mixture-averaged transport (oracle/transport.py,
mirgecom/transport.py:324-502 on pyrometheus' mixture rules):
test/test_transport.py:119-159 restated (pure-species states: mixture
viscosity and conductivity equal the species values, the singular
diffusivity for the present species), the trace-species closed form
d_j = D_jk/p, Wilke's symmetry phi_jk = phi_kj (mu_j/mu_k)(W_k/W_j), the
binary-mixture form of the diffusivity, and the Lewis-number branch."""

import numpy as np
import pytest

from mirgecom_amd.transport_fits import transport_fits
from oracle import thermochem as tc
from oracle import transport as otr

FITS = transport_fits()
NSP = 7


def _state(y, t, p=101325.0, dim=2):
    """Conserved state (F, n) at rest for mass fractions y (7, n)."""
    rho = tc.density(p, t, y)
    e = tc.mixture_energy_mass(t, y)
    zero = 0 * t
    return np.stack([rho, rho * e] + [zero] * dim + [rho * yk for yk in y])


@pytest.mark.parametrize("k", range(NSP))
def test_pure_species(k):
    t = np.array([300.0, 800.0, 1500.0, 3000.0])
    y = np.zeros((NSP, t.size))
    y[k] = 1.0
    mu_k = otr.species_viscosities(t, FITS[0])[k]
    lam_k = otr.species_thermal_conductivities(t, FITS[1])[k]
    np.testing.assert_allclose(otr.mixture_viscosity(t, y, FITS[0]), mu_k,
                               rtol=1e-14)
    np.testing.assert_allclose(otr.mixture_thermal_conductivity(t, y, FITS[1]),
                               lam_k, rtol=1e-14)
    tr = otr.MixtureAveragedTransport(FITS, singular_diffusivity=1e-6,
                                      factor=1.0)
    cv = _state(y, t)
    tv = tr.transport_vars(cv, t, tc.PyrometheusMixture(), 2)
    assert np.all(tv.species_diffusivity[k] == 1e-6)
    # every other species is a trace in pure k: d_j = D_jk / p
    bd = otr.species_binary_diffusivities(t, FITS[2])
    for j in range(NSP):
        if j != k:
            np.testing.assert_allclose(tv.species_diffusivity[j],
                                       bd[j, k] / 101325.0, rtol=1e-13)
    np.testing.assert_allclose(tv.bulk_viscosity, 0.6 * mu_k, rtol=1e-14)


def test_wilke_symmetry():
    t = np.array([400.0, 2000.0])
    mu = otr.species_viscosities(t, FITS[0])
    w = tc.W
    phi = otr.wilke_phi(mu, w)
    for k in range(NSP):
        np.testing.assert_allclose(phi[k][k], 1.0, rtol=1e-15)
        for j in range(NSP):
            np.testing.assert_allclose(
                phi[j][k], phi[k][j] * (mu[j] / mu[k]) * (w[k] / w[j]),
                rtol=1e-13)


def test_binary_mixture_diffusivity():
    """Two species a, b only: d_a = D_ab (1 - Y_a) / (p X_b)."""
    a, b = 1, 6   # O2, N2
    t = np.array([500.0, 1200.0])
    y = np.zeros((NSP, 2))
    y[a], y[b] = 0.3, 0.7
    p = np.array([101325.0, 2e5])
    d = otr.species_mass_diffusivities(p, t, y, FITS[2])
    x = otr.mole_fractions(y)
    bd = otr.species_binary_diffusivities(t, FITS[2])
    np.testing.assert_allclose(d[a], bd[a, b] * (1 - y[a]) / (p * x[b]),
                               rtol=1e-13)
    np.testing.assert_allclose(d[b], bd[a, b] * (1 - y[b]) / (p * x[a]),
                               rtol=1e-13)


def test_lewis_and_factor():
    rng = np.random.default_rng(2)
    t = rng.uniform(400, 2500, 16)
    y = rng.uniform(0.01, 1.0, (NSP, 16))
    y /= y.sum(axis=0)
    cv = _state(y, t)
    eos = tc.PyrometheusMixture()
    le = np.linspace(0.8, 1.4, NSP)
    base = otr.MixtureAveragedTransport(FITS).transport_vars(cv, t, eos, 2)
    tr = otr.MixtureAveragedTransport(FITS, factor=2.0, lewis=le, alpha=0.4)
    tv = tr.transport_vars(cv, t, eos, 2)
    np.testing.assert_allclose(tv.viscosity, 2.0 * base.viscosity, rtol=1e-15)
    np.testing.assert_allclose(tv.bulk_viscosity, 0.8 * base.viscosity,
                               rtol=1e-15)
    cp = tc.mixture_cp_mass(t, y)
    # rho d_k Le_k c_p = kappa
    np.testing.assert_allclose(cv[0] * tv.species_diffusivity * le[:, None] * cp,
                               np.broadcast_to(tv.thermal_conductivity, (NSP, 16)),
                               rtol=1e-14)
    f2 = otr.MixtureAveragedTransport(FITS, factor=2.0).transport_vars(cv, t, eos, 2)
    np.testing.assert_allclose(f2.species_diffusivity,
                               2.0 * base.species_diffusivity, rtol=1e-15)


def test_mixture_bounds():
    """Mixture viscosity / conductivity lie between the species extremes."""
    rng = np.random.default_rng(5)
    t = rng.uniform(300, 3500, 64)
    y = rng.uniform(0.0, 1.0, (NSP, 64))
    y /= y.sum(axis=0)
    mu = otr.mixture_viscosity(t, y, FITS[0])
    mus = otr.species_viscosities(t, FITS[0])
    assert np.all(mu <= mus.max(axis=0) * (1 + 1e-12))
    assert np.all(mu >= mus.min(axis=0) * (1 - 1e-12))
    lam = otr.mixture_thermal_conductivity(t, y, FITS[1])
    lams = otr.species_thermal_conductivities(t, FITS[1])
    assert np.all(lam <= lams.max(axis=0)) and np.all(lam >= lams.min(axis=0))
