"""The oracle's DG operators against the reference's own invariant and
convergence tests (restated; the reference cannot be imported here):

* uniform flow => RHS = 0: test/test_euler.py:80-232 (1e-9),
  test/test_navierstokes.py:81-235
* steady isentropic vortex RHS -> 0 with EOC >= p - 0.7:
  test/test_euler.py:405-480 (same 32/48/64 meshes of [-5,5]^2)
* advected vortex RK4 march vs the exact solution, EOC >= p - 0.5:
  test/test_euler.py:816-887
* DG gradient exact for polynomials of degree <= p:
  test/test_operators.py:237-416
* RK4 order on y' = -exp(-t): test/test_time_integrators.py:47-88
"""

import numpy as np
import pytest

from oracle import initializers as oini
from oracle import operators as oop
from oracle import physics as ph
from tests.cases import Problem


def _eoc(hs, errs):
    return np.polyfit(np.log(hs), np.log(errs), 1)[0]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("nspecies", [0, 3])
def test_uniform_rhs_zero(dim, order, nspecies):
    pb = Problem(dim, order, (3,) * dim, 0.0, 1.0, (False,) * dim)
    y = [0.2, 0.3, 0.5][:nspecies] if nspecies else None
    q = oini.uniform(pb.x, rho=1.2, p=2.0, velocity=[0.3, -0.2, 0.1][:dim],
                     mass_fractions=y)
    gm = ph.GasModel(ph.IdealSingleGas())
    st = ph.make_fluid_state(q, gm, dim)
    bnd = {"all": oop.DummyBoundary()}
    for flux in (ph.inviscid_facial_flux_rusanov, ph.inviscid_facial_flux_hll):
        r = oop.euler_operator(pb.odisc, pb.box, st, gm, bnd,
                               numerical_flux_func=flux)
        assert np.abs(r).max() < 1e-9
    gmv = ph.GasModel(ph.IdealSingleGas(),
                      ph.SimpleTransport(viscosity=0.1, thermal_conductivity=0.2,
                                         species_diffusivity=[0.01] * nspecies))
    r = oop.ns_operator(pb.odisc, pb.box, ph.make_fluid_state(q, gmv, dim), gmv,
                        bnd)
    assert np.abs(r).max() < 2e-9


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("flux", ["rusanov", "hll"])
def test_vortex_rhs_eoc(order, flux):
    fn = {"rusanov": ph.inviscid_facial_flux_rusanov,
          "hll": ph.inviscid_facial_flux_hll}[flux]
    gm = ph.GasModel(ph.IdealSingleGas())

    def bfunc(xb, sm, t):
        return ph.make_fluid_state(oini.vortex2d(xb, t, beta=5), gm, 2)

    hs, errs = [], []
    for nel in (32, 48, 64):
        pb = Problem(2, order, (nel, nel), -5.0, 5.0, (False, False),
                     own_matrices=True)
        q = oini.vortex2d(pb.x, beta=5)
        r = oop.euler_operator(pb.odisc, pb.box, ph.make_fluid_state(q, gm, 2), gm,
                               {"all": oop.PrescribedFluidBoundary(bfunc)},
                               numerical_flux_func=fn)
        hs.append(10.0 / nel)
        errs.append(np.abs(r).max())
    assert _eoc(hs, errs) >= order - 0.7 or max(errs) < 1e-9


@pytest.mark.parametrize("order", [2, 3])
def test_vortex_march_eoc(order):
    gm = ph.GasModel(ph.IdealSingleGas())
    kw = dict(beta=5, center=(0, 0), velocity=(1, 1))

    def bfunc(xb, sm, t):
        return ph.make_fluid_state(oini.vortex2d(xb, t, **kw), gm, 2)

    hs, errs = [], []
    for nel in (16, 32, 64):
        pb = Problem(2, order, (nel, nel), -5.0, 5.0, (False, False),
                     own_matrices=True)
        bnd = {"all": oop.PrescribedFluidBoundary(bfunc)}

        def rhs(t, q):
            return oop.euler_operator(pb.odisc, pb.box,
                                      ph.make_fluid_state(q, gm, 2), gm, bnd,
                                      time=t)

        q = oini.vortex2d(pb.x, 0.0, **kw)
        dt = 1e-4
        for step in range(10):
            q = oop.rk4_step(q, step * dt, dt, rhs)
        exact = oini.vortex2d(pb.x, 10 * dt, **kw)
        hs.append(10.0 / nel)
        errs.append(np.abs(q - exact).max())
    assert _eoc(hs, errs) >= order - 0.5 or max(errs) < 1e-11


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_gradient_polynomial_exact(dim, order):
    """Central-flux DG gradient of a continuous polynomial of degree p is
    exact (traces agree, so u* = u)."""
    pb = Problem(dim, order, (3,) * dim, 0.0, 1.0, (False,) * dim,
                 own_matrices=True)
    x = pb.x
    u = x[0]**order + 0.5 * x[-1]**max(order - 1, 0) * x[0] + 1.0
    du = [order * x[0]**(order - 1) + 0.5 * x[-1]**max(order - 1, 0)]
    du += [0 * x[0]] * (dim - 1)
    if order > 1:
        du[-1] = du[-1] + 0.5 * (order - 1) * x[-1]**(order - 2) * x[0]
    od = pb.odisc
    um = od.trace(u)
    n = np.moveaxis(od.normal, -1, 0)[:, :, :, None]
    g = oop.grad_operator(od, u, um[None] * n)
    for i in range(dim):
        assert np.abs(g[i] - du[i]).max() < 1e-10


def test_rk4_order():
    """y' = -exp(-t), y(0) = 1 => y = exp(-t); RK4 error ~ dt^4."""
    errs = []
    dts = [0.1, 0.05, 0.025]
    for dt in dts:
        y = np.array(1.0)
        t = 0.0
        while t < 1.0 - 1e-12:
            y = oop.rk4_step(y, t, dt, lambda tt, yy: -np.exp(-tt))
            t += dt
        errs.append(abs(y - np.exp(-1.0)))
    assert _eoc(dts, errs) > 3.9
