"""Pins of the oracle's overintegrated Navier-Stokes operator
(oracle/overintegration.py ``grad_operators_quad`` / ``ns_operator_quad``;
ns_operator with DISCR_TAG_QUAD, mirgecom/navierstokes.py:350-575):

* uniform states give 0 (test/test_navierstokes.py:81-235 with ``quad``);
* the gradient of the conserved state equals the nodal oracle's for any
  state (both integrate the degree <= 2p weak-form integrands exactly), and
  so does the temperature gradient where T is a polynomial of degree <= p;
* for a state whose inviscid and viscous fluxes are of degree <= p, the
  whole RHS equals the nodal oracle's (prescribed exact boundary state);
* the exact shear flow of test/test_navierstokes.py:615-712 gives a RHS that
  converges at order >= p - 1/2.
"""

import numpy as np
import pytest

from oracle import operators as oop, physics as ph
from oracle.overintegration import (
    QuadStates, grad_operators_quad, ns_operator_quad)
from tests.cases import Problem, smooth_perturbed_state


def _gas(gamma=1.4, gas_const=1.0, mu=0.05, kappa=0.07, mub=0.01, d=None):
    return ph.GasModel(ph.IdealSingleGas(gamma=gamma, gas_const=gas_const),
                       ph.SimpleTransport(viscosity=mu, thermal_conductivity=kappa,
                                          bulk_viscosity=mub,
                                          species_diffusivity=d))


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 3), (3, 2)])
@pytest.mark.parametrize("nspecies", [0, 3])
def test_uniform_zero(dim, order, nspecies):
    n = (3,) * dim
    pb = Problem(dim, order, n, 0.0, 1.0, (True,) * dim, own_matrices=True)
    gm = _gas(d=np.linspace(0.01, 0.03, nspecies) if nspecies else None)
    from oracle import initializers as oini
    y = [0.1 * (k + 1) for k in range(nspecies)] if nspecies else None
    q = oini.uniform(pb.x, rho=1.2, p=2.0, velocity=[0.3, -0.2, 0.1][:dim],
                     mass_fractions=y)
    r = ns_operator_quad(pb.odisc, pb.box, q, gm, {})
    assert np.abs(r).max() < 1e-11


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 3), (3, 2)])
def test_grad_cv_equals_nodal(dim, order):
    pb = Problem(dim, order, (3,) * dim, 0.0, 2 * np.pi, (True,) * dim,
                 own_matrices=True)
    q = smooth_perturbed_state(pb.x, seed=3, amp=0.1)
    gm = _gas()
    qs = QuadStates(pb.odisc, pb.box, q, gm, {}, 0.0, 2 * order)
    gcv, _ = grad_operators_quad(qs)
    ops = oop.OperatorStates(pb.odisc, pb.box, ph.make_fluid_state(q, gm, dim),
                             gm, {}, 0.0)
    ref = oop.grad_cv_operator(ops)
    assert np.abs(gcv - ref).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("dim,order", [(2, 2), (3, 2)])
def test_grad_t_equals_nodal_for_polynomial_t(dim, order):
    """rho and u constant, p of degree p: T = p / (rho R) is then the same
    polynomial at the points as at the nodes."""
    pb = Problem(dim, order, (3,) * dim, 0.0, 1.0, (False,) * dim,
                 own_matrices=True)
    gm = _gas()

    def state(x):
        rho = 1.3 + 0 * x[0]
        v = np.array([0.3, -0.2, 0.1])[:dim]
        p = 2.0 + 0.3 * x[0] ** order - 0.2 * x[-1] * x[0] ** (order - 1)
        e = p / 0.4 + 0.5 * rho * float(v @ v)
        return np.stack([rho, e] + [rho * vi for vi in v])

    q = state(pb.x)

    def bfunc(xb, sm, t):
        return ph.make_fluid_state(state(xb), gm, dim)

    bnd = {"all": oop.PrescribedFluidBoundary(bfunc)}
    qs = QuadStates(pb.odisc, pb.box, q, gm, bnd, 0.0, 2 * order)
    gcv, gt = grad_operators_quad(qs)
    ops = oop.OperatorStates(pb.odisc, pb.box, ph.make_fluid_state(q, gm, dim),
                             gm, bnd, 0.0)
    ref_t = oop.grad_t_operator(ops)
    ref_cv = oop.grad_cv_operator(ops)
    assert np.abs(gt - ref_t).max() < 1e-10 * np.abs(ref_t).max()
    assert np.abs(gcv - ref_cv).max() < 1e-10 * np.abs(ref_cv).max()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", ["rusanov", "hll"])
def test_equals_nodal_for_degree_p_fluxes(dim, flux):
    """p = 3, rho = 1, linear velocity and pressure: the inviscid fluxes are
    of degree 3, the viscous ones of degree 1, T is linear; both operators
    integrate exactly, so the RHS agree (and are not small)."""
    order = 3
    n = (3, 2) if dim == 2 else (2, 2, 1)
    pb = Problem(dim, order, n, 0.0, 1.0, (False,) * dim, own_matrices=True)
    gm = _gas(mu=0.3, kappa=0.5, mub=0.1)
    nf = {"rusanov": ph.inviscid_facial_flux_rusanov,
          "hll": ph.inviscid_facial_flux_hll}[flux]

    def state(x):
        rho = 1.0 + 0 * x[0]
        v = [0.2 + 0.1 * x[i] - 0.05 * x[(i + 1) % dim] for i in range(dim)]
        p = 2.0 + 0.3 * x[0] - 0.2 * x[-1]
        e = p / 0.4 + 0.5 * sum(vi * vi for vi in v)
        return np.stack([rho, e] + v)

    q = state(pb.x)

    def bfunc(xb, sm, t):
        return ph.make_fluid_state(state(xb), gm, dim)

    bnd = {"all": oop.PrescribedFluidBoundary(bfunc)}
    rq = ns_operator_quad(pb.odisc, pb.box, q, gm, bnd,
                          inviscid_numerical_flux_func=nf)
    rn = oop.ns_operator(pb.odisc, pb.box, ph.make_fluid_state(q, gm, dim), gm,
                         bnd, inviscid_numerical_flux_func=nf)
    scale = np.abs(rn).max()
    assert scale > 1e-2
    assert np.abs(rq - rn).max() < 1e-10 * scale
    # the viscous part alone (inviscid terms off) as well
    rq = ns_operator_quad(pb.odisc, pb.box, q, gm, bnd, inviscid_terms_on=False)
    rn = oop.ns_operator(pb.odisc, pb.box, ph.make_fluid_state(q, gm, dim), gm,
                         bnd, inviscid_terms_on=False)
    # (its RHS, 1e-2, is small against the fluxes, O(1): absolute bound)
    assert np.abs(rn).max() > 1e-3
    assert np.abs(rq - rn).max() < 1e-10


def _shear_flow(x, flow_dir, trans_dir, mu, gamma):
    """mirgecom/initializers.py ShearFlow (Hesthaven & Warburton 2008,
    sec. 7.5.3): rho = 1, u_flow = x_trans^2, p = 2 mu x_flow + 10."""
    dim = x.shape[0]
    rho = 1.0 + 0 * x[0]
    v = [0 * x[0] for _ in range(dim)]
    v[flow_dir] = x[trans_dir] ** 2
    p = 2 * mu * x[flow_dir] + 10
    e = p / (gamma - 1) + 0.5 * rho * sum(vi * vi for vi in v)
    return np.stack([rho, e] + [rho * vi for vi in v])


@pytest.mark.parametrize("order", [2, 3])
def test_shear_flow_eoc(order):
    """test/test_navierstokes.py:615-712 (2-D, flow along x, quad=True): the
    energy RHS of the exact solution converges at order >= p - 1/2."""
    mu, gamma = 0.01, 1.5
    gm = ph.GasModel(ph.IdealSingleGas(gamma=gamma, gas_const=287.0),
                     ph.SimpleTransport(viscosity=mu, thermal_conductivity=0.0))

    def bfunc(xb, sm, t):
        return ph.make_fluid_state(_shear_flow(xb, 0, 1, mu, gamma), gm, 2)

    hs, errs = [], []
    for n in (4, 8, 16):
        pb = Problem(2, order, (n, n), 0.0, 1.0, (False, False),
                     own_matrices=True)
        q = _shear_flow(pb.x, 0, 1, mu, gamma)
        r = ns_operator_quad(pb.odisc, pb.box, q, gm,
                             {"all": oop.PrescribedFluidBoundary(bfunc)})
        hs.append(1.0 / n)
        errs.append(np.sqrt(np.mean(r[1] ** 2)))
    eoc = np.polyfit(np.log(hs), np.log(errs), 1)[0]
    assert eoc >= order - 0.5 or max(errs) < 1e-8


def test_walls_raise():
    pb = Problem(2, 1, (2, 2), 0.0, 1.0, (False, False), own_matrices=True)
    gm = _gas()
    q = smooth_perturbed_state(pb.x, seed=1)

    class Wall:
        kind = "mengaldo"

        def state_plus(self, sm, n, gas_model):
            return sm

    with pytest.raises(NotImplementedError):
        ns_operator_quad(pb.odisc, pb.box, q, gm, {"all": Wall()})
