"""Pins of the oracle's state-based entropy-stable operator
(oracle/esdg.py ``esdg_operator``; mirgecom/euler.py:108-310) and Renac's
two-point flux (``renac``, mirgecom/inviscid.py:521-598):

* with the ideal gas and Chandrashekar's flux the state-based pipeline equals
  the pinned ``esdg_euler_operator`` (periodic and prescribed boundaries):
  projection, hybridized states, face conversions and assembly are pinned;
* Renac's flux is symmetric, its mass / momentum / species parts are
  consistent with the physical flux, and its energy part at equal states is
  ``(c_v(T) T + |u|^2/2) rho u + p u`` (the form eq. 24 of Renac 2021
  reduces to);
* a uniform mixture state gives 0, and the mixture RHS conserves mass,
  momentum, energy and every species on a periodic box (the flux
  differencing of a symmetric two-point flux is conservative).
"""

import numpy as np
import pytest

from oracle import operators as oop, physics as ph, thermochem as tc
from oracle.esdg import esdg_euler_operator, esdg_operator, renac
from tests.cases import Problem, mixture_state, smooth_perturbed_state

MGM = ph.GasModel(tc.PyrometheusMixture())


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 3), (3, 2)])
def test_ideal_gas_equals_pinned_operator(dim, order):
    pb = Problem(dim, order, (3,) * dim, 0.0, 1.0, (True,) * dim,
                 own_matrices=True)
    q = smooth_perturbed_state(pb.x, seed=3, amp=0.1, nspecies=2)
    gm = ph.GasModel(ph.IdealSingleGas(gamma=1.4))
    for diss in (True, False):
        a = esdg_euler_operator(pb.odisc, pb.box, q, 1.4, {}, dissipation=diss)
        b = esdg_operator(pb.odisc, pb.box, q, gm, {}, dissipation=diss)
        assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("kind", ["prescribed", "dummy"])
def test_ideal_gas_boundaries_equal_pinned_operator(kind):
    pb = Problem(2, 2, (3, 3), 0.0, 1.0, (False, False), own_matrices=True)
    q = smooth_perturbed_state(pb.x, seed=5, amp=0.1)
    gm = ph.GasModel(ph.IdealSingleGas(gamma=1.4))

    def bfunc(xb, sm, t):
        return ph.make_fluid_state(smooth_perturbed_state(xb, seed=5, amp=0.1),
                                   gm, 2)

    bnd = {"all": (oop.PrescribedFluidBoundary(bfunc) if kind == "prescribed"
                   else oop.DummyBoundary())}
    a = esdg_euler_operator(pb.odisc, pb.box, q, 1.4, bnd)
    b = esdg_operator(pb.odisc, pb.box, q, gm, bnd)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()


def _mix_states(n=40, seed=2):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 2 * np.pi, (3, n))
    q, t = mixture_state(x, seed=seed)
    return ph.make_fluid_state(q, MGM, 3, temperature_seed=t + 1.0)


def test_renac_symmetry_and_consistency():
    sl = _mix_states(seed=2)
    sr = _mix_states(seed=3)
    eos = MGM.eos
    f_lr = renac(sl, sr, eos, 3)
    f_rl = renac(sr, sl, eos, 3)
    np.testing.assert_allclose(f_lr, f_rl, rtol=1e-12, atol=1e-12 * np.abs(f_lr).max())
    f = renac(sl, sl, eos, 3)
    phys = ph.inviscid_flux(sl)
    rows = [0] + list(range(2, phys.shape[0]))
    np.testing.assert_allclose(f[rows], phys[rows], rtol=1e-11,
                               atol=1e-11 * np.abs(phys).max())
    u = sl.velocity
    cvm = tc.mixture_cv_mass(sl.temperature, sl.species_mass_fractions)
    e_flux = ((cvm * sl.temperature + 0.5 * np.sum(u * u, axis=0)) * sl.mass * u
              + sl.pressure * u)
    np.testing.assert_allclose(f[1], e_flux, rtol=1e-11)


def test_mixture_uniform_zero():
    pb = Problem(2, 2, (3, 3), 0.0, 1.0, (True, True), own_matrices=True)
    y = np.array([0.05, 0.2, 0.03, 0.01, 0.02, 0.001, 0.0])
    y[-1] = 1 - y.sum()
    t0, p0 = 1500.0, 101325.0
    rho = tc.density(np.array([p0]), np.array([t0]), y[:, None])[0]
    e = tc.mixture_energy_mass(np.array([t0]), y[:, None])[0]
    u = np.array([3.0, -2.0])
    one = np.ones(pb.x.shape[1:])
    q = np.stack([rho * one, rho * (e + 0.5 * u @ u) * one]
                 + [rho * u[d] * one for d in range(2)]
                 + [rho * y[k] * one for k in range(7)])
    st = ph.make_fluid_state(q, MGM, 2, temperature_seed=t0 * one)
    h = 1.0 / 3
    # round-off scale: the penalty alpha (q+ - q-) / h with alpha = |u| + c
    # and the fluxes (q u, p) / h
    scale = (st.wavespeed.max() * np.abs(q).max() + st.pressure.max()) / h
    for diss in (True, False):
        r = esdg_operator(pb.odisc, pb.box, q, MGM, {}, temperature=t0 * one,
                          dissipation=diss)
        assert np.abs(r).max() < 1e-12 * scale


def test_mixture_conservation():
    pb = Problem(2, 2, (3, 3), 0.0, 2 * np.pi, (True, True), own_matrices=True)
    q, t = mixture_state(pb.x, seed=41)
    r = esdg_operator(pb.odisc, pb.box, q, MGM, {}, temperature=t)
    # integral of each RHS row: sum_k J_k 1^T M r_k
    od = pb.odisc
    mass = od.el.mass
    tot = np.einsum("ij,fkj,k->f", mass, r, od.J)
    scale = np.einsum("ij,fkj,k->f", mass, np.abs(r), od.J)
    assert np.all(np.abs(tot) < 1e-11 * scale)
    assert np.abs(r).max() > 1e-3 * np.abs(q).max()
