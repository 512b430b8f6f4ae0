"""GPU parity: the HIP right-hand side against the FP64 CPU oracle on the same
mesh and seeded states.  Tolerance: <= 1e-12 relative per field (FP64).
Connectivity is the same int32 map in both (checked bit-exact on the CPU in
tests/test_connectivity.py)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import Problem, rel_err, smooth_perturbed_state  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


def _flux_pair(name):
    from mirgecom_amd import inviscid as gi
    from oracle import physics as ph
    return {
        "rusanov": (gi.inviscid_facial_flux_rusanov,
                    ph.inviscid_facial_flux_rusanov),
        "hll": (gi.inviscid_facial_flux_hll, ph.inviscid_facial_flux_hll),
        "central": (gi.inviscid_facial_flux_central,
                    ph.inviscid_facial_flux_central),
    }[name]


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("flux", ["rusanov", "hll", "central"])
def test_euler_vortex_prescribed(order, flux):
    """2-D isentropic vortex with exact prescribed boundaries (the
    [0,10]x[-5,5] vortex workload on a coarser mesh)."""
    from mirgecom_amd.boundary import PrescribedFluidBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.initializers import Vortex2D
    from oracle import initializers as oini, operators as oop, physics as ph

    pb = Problem(2, order, (6, 6), (0.0, -5.0), (10.0, 5.0), (False, False))
    t = 0.37
    vortex = Vortex2D(beta=5, center=(5, 0), velocity=(1, 0))
    gm = GasModel(eos=IdealSingleGas())

    def bstate(dcoll, dd_bdry, gas_model, state_minus, time=0.0, **kw):
        x = dcoll.nodes(dd_bdry)
        return make_fluid_state(vortex(x, time=time, eos=gas_model.eos),
                                gas_model)

    q0 = oini.vortex2d(pb.x, time=0.1, beta=5, center=(5, 0), velocity=(1, 0))
    state = make_fluid_state(ConservedVars(2, _dev(q0)), gm)
    gflux, oflux = _flux_pair(flux)
    rhs = euler_operator(pb.dcoll, state, gm,
                         {"all": PrescribedFluidBoundary(boundary_state_func=bstate)},
                         time=t, inviscid_numerical_flux_func=gflux)
    ogm = ph.GasModel(ph.IdealSingleGas())
    ost = ph.make_fluid_state(q0, ogm, 2)

    def obstate(xb, sm, time):
        return ph.make_fluid_state(
            oini.vortex2d(xb, time, beta=5, center=(5, 0), velocity=(1, 0)), ogm, 2)

    ref = oop.euler_operator(pb.odisc, pb.box, ost, ogm,
                             {"all": oop.PrescribedFluidBoundary(obstate)},
                             time=t, numerical_flux_func=oflux)
    err = rel_err(_host(rhs.q), ref)
    print("euler vortex", order, flux, err)
    assert err.max() < TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_euler_3d_periodic(order):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from oracle import operators as oop, physics as ph

    pb = Problem(3, order, (3, 3, 3), 0.0, 2 * np.pi, (True, True, True))
    q0 = smooth_perturbed_state(pb.x, seed=3)
    gm = GasModel(eos=IdealSingleGas())
    state = make_fluid_state(ConservedVars(3, _dev(q0)), gm)
    rhs = euler_operator(pb.dcoll, state, gm, {},
                         inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
    ogm = ph.GasModel(ph.IdealSingleGas())
    ref = oop.euler_operator(pb.odisc, pb.box, ph.make_fluid_state(q0, ogm, 3),
                             ogm, {})
    err = rel_err(_host(rhs.q), ref)
    print("euler 3d", order, err)
    assert err.max() < TOL


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_ns_periodic(dim, order):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    from oracle import operators as oop, physics as ph

    n = (4,) * dim if dim == 2 else (3,) * dim
    pb = Problem(dim, order, n, 0.0, 2 * np.pi, (True,) * dim)
    q0 = smooth_perturbed_state(pb.x, seed=7)
    mu, kappa, mub = 0.05, 0.07, 0.01
    gm = GasModel(eos=IdealSingleGas(),
                  transport=SimpleTransport(viscosity=mu, thermal_conductivity=kappa,
                                            bulk_viscosity=mub))
    state = make_fluid_state(ConservedVars(dim, _dev(q0)), gm)
    rhs, gcv, gt = ns_operator(pb.dcoll, gm, state, {}, return_gradients=True)
    ogm = ph.GasModel(ph.IdealSingleGas(),
                      ph.SimpleTransport(viscosity=mu, thermal_conductivity=kappa,
                                         bulk_viscosity=mub))
    ref, rgcv, rgt = oop.ns_operator(pb.odisc, pb.box,
                                     ph.make_fluid_state(q0, ogm, dim), ogm, {},
                                     return_gradients=True)
    eg = rel_err(_host(gcv.q).reshape(-1, *rgcv.shape[2:]),
                 rgcv.reshape(-1, *rgcv.shape[2:]))
    et = rel_err(_host(gt), rgt)
    er = rel_err(_host(rhs.q), ref)
    print("ns", dim, order, eg.max(), et.max(), er)
    assert eg.max() < TOL and et.max() < TOL and er.max() < TOL


@pytest.mark.parametrize("transport", ["simple", "powerlaw"])
def test_ns_species_dummy_boundary(transport):
    """Passive species, DummyBoundary on a non-periodic 3-D box."""
    from mirgecom_amd.boundary import DummyBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import PowerLawTransport, SimpleTransport
    from oracle import operators as oop, physics as ph

    pb = Problem(3, 2, (3, 3, 2), 0.0, (1.0, 0.8, 0.6), (False, True, False))
    q0 = smooth_perturbed_state(pb.x, seed=11, nspecies=7)
    d = np.linspace(0.01, 0.07, 7)
    if transport == "simple":
        gt_ = SimpleTransport(viscosity=0.03, thermal_conductivity=0.05,
                              species_diffusivity=d)
        ot_ = ph.SimpleTransport(viscosity=0.03, thermal_conductivity=0.05,
                                 species_diffusivity=d)
    else:
        gt_ = PowerLawTransport(beta=0.01, n=0.7, species_diffusivity=d)
        ot_ = ph.PowerLawTransport(beta=0.01, n=0.7, species_diffusivity=d)
    gm = GasModel(eos=IdealSingleGas(gas_const=1.0), transport=gt_)
    state = make_fluid_state(ConservedVars(3, _dev(q0)), gm)
    with pytest.warns(UserWarning):
        bnd = {"all": DummyBoundary()}
    rhs = ns_operator(pb.dcoll, gm, state, bnd)
    ogm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0), ot_)
    ref = oop.ns_operator(pb.odisc, pb.box, ph.make_fluid_state(q0, ogm, 3), ogm,
                          {"all": oop.DummyBoundary()})
    err = rel_err(_host(rhs.q), ref)
    print("ns species", transport, err)
    assert err.max() < TOL


def test_ns_paths_agree_and_prescribed_viscous_boundary():
    """Navier-Stokes with a prescribed (non-trivial) boundary state: the
    fused path (nodal net fluxes), the path with returned gradients and the
    path from explicitly given gradients all match the oracle."""
    from mirgecom_amd.boundary import PrescribedFluidBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.initializers import Vortex2D
    from mirgecom_amd.inviscid import inviscid_facial_flux_hll
    from mirgecom_amd.navierstokes import (
        grad_cv_operator, grad_t_operator, ns_operator)
    from mirgecom_amd.transport import SimpleTransport
    from oracle import initializers as oini, operators as oop, physics as ph

    pb = Problem(2, 3, (6, 5), (0.0, -5.0), (10.0, 5.0), (False, False))
    vortex = Vortex2D(beta=5, center=(5, 0), velocity=(1, 0))
    tr = dict(viscosity=0.02, thermal_conductivity=0.03, bulk_viscosity=0.01)
    gm = GasModel(eos=IdealSingleGas(), transport=SimpleTransport(**tr))

    def bstate(dcoll, dd_bdry, gas_model, state_minus, time=0.0, **kw):
        # exterior = exact vortex at a shifted time, so q_ext != q-
        return make_fluid_state(vortex(dcoll.nodes(dd_bdry), time=time + 0.3),
                                gas_model)

    q0 = oini.vortex2d(pb.x, time=0.1, beta=5, center=(5, 0), velocity=(1, 0))
    st = make_fluid_state(ConservedVars(2, _dev(q0)), gm)
    bnd = {"all": PrescribedFluidBoundary(boundary_state_func=bstate)}
    ogm = ph.GasModel(ph.IdealSingleGas(), ph.SimpleTransport(**tr))

    def obstate(xb, sm, time):
        return ph.make_fluid_state(
            oini.vortex2d(xb, time + 0.3, beta=5, center=(5, 0), velocity=(1, 0)),
            ogm, 2)

    for gflux, oflux in [(None, None), (inviscid_facial_flux_hll,
                                        ph.inviscid_facial_flux_hll)]:
        kw = {} if gflux is None else {"inviscid_numerical_flux_func": gflux}
        okw = {} if oflux is None else {"inviscid_numerical_flux_func": oflux}
        ref, rgcv, rgt = oop.ns_operator(
            pb.odisc, pb.box, ph.make_fluid_state(q0, ogm, 2), ogm,
            {"all": oop.PrescribedFluidBoundary(obstate)}, time=0.2,
            return_gradients=True, **okw)
        r1 = ns_operator(pb.dcoll, gm, st, bnd, time=0.2, **kw)
        r2, gcv, gt = ns_operator(pb.dcoll, gm, st, bnd, time=0.2,
                                  return_gradients=True, **kw)
        gcv3 = grad_cv_operator(pb.dcoll, gm, bnd, st, time=0.2)
        gt3 = grad_t_operator(pb.dcoll, gm, bnd, st, time=0.2)
        r3 = ns_operator(pb.dcoll, gm, st, bnd, time=0.2, grad_cv=gcv3,
                         grad_t=gt3, **kw)
        for r in (r1, r2, r3):
            err = rel_err(_host(r.q), ref)
            print("ns prescribed", gflux, err)
            assert err.max() < TOL
        eg = rel_err(_host(gcv.q).reshape(-1, *rgcv.shape[2:]),
                     rgcv.reshape(-1, *rgcv.shape[2:]))
        assert eg.max() < TOL and rel_err(_host(gt), rgt).max() < TOL
        assert np.array_equal(_host(gcv.q), _host(gcv3.q))


def test_ns_own_matrices_noise_floor():
    """Against the oracle built entirely from its own element matrices, p=3:
    the remaining difference is the roundoff of constructing the matrices two
    ways (about 2e-12), bounded at 5e-12."""
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    from oracle import operators as oop, physics as ph

    pb = Problem(3, 3, (3, 3, 3), 0.0, 2 * np.pi, (True,) * 3, own_matrices=True)
    q0 = smooth_perturbed_state(pb.x, seed=7)
    gm = GasModel(eos=IdealSingleGas(),
                  transport=SimpleTransport(viscosity=0.05, thermal_conductivity=0.07))
    rhs = ns_operator(pb.dcoll, gm, make_fluid_state(ConservedVars(3, _dev(q0)), gm),
                      {})
    ogm = ph.GasModel(ph.IdealSingleGas(),
                      ph.SimpleTransport(viscosity=0.05, thermal_conductivity=0.07))
    ref = oop.ns_operator(pb.odisc, pb.box, ph.make_fluid_state(q0, ogm, 3), ogm, {})
    err = rel_err(_host(rhs.q), ref)
    print("ns own matrices", err)
    assert err.max() < 5e-12


def test_uniform_state_zero_rhs():
    """Uniform flow => RHS = 0 (test/test_euler.py:80-232,
    test/test_navierstokes.py:81-235), on the GPU."""
    from mirgecom_amd.boundary import DummyBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    from oracle import initializers as oini

    for dim in (2, 3):
        for order in (1, 2, 3):
            pb = Problem(dim, order, (3,) * dim, 0.0, 1.0, (False,) * dim)
            q0 = oini.uniform(pb.x, rho=1.2, p=2.0,
                              velocity=[0.3, -0.2, 0.1][:dim])
            gm = GasModel(eos=IdealSingleGas(),
                          transport=SimpleTransport(viscosity=0.1,
                                                    thermal_conductivity=0.2))
            st = make_fluid_state(ConservedVars(dim, _dev(q0)), gm)
            with pytest.warns(UserWarning):
                bnd = {"all": DummyBoundary()}
            rhs = ns_operator(pb.dcoll, gm, st, bnd)
            assert np.abs(_host(rhs.q)).max() < 1e-9


def test_rk4_vortex_march():
    """Ten RK4 steps of the 2-D vortex (the vortex workload on a coarser mesh):
    fused device RK4 vs oracle RK4."""
    from mirgecom_amd.boundary import PrescribedFluidBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.initializers import Vortex2D
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.integrators import rk4_step
    from oracle import initializers as oini, operators as oop, physics as ph

    pb = Problem(2, 3, (8, 8), (0.0, -5.0), (10.0, 5.0), (False, False))
    vortex = Vortex2D(beta=5, center=(5, 0), velocity=(1, 0))
    gm = GasModel(eos=IdealSingleGas())

    def bstate(dcoll, dd_bdry, gas_model, state_minus, time=0.0, **kw):
        return make_fluid_state(vortex(dcoll.nodes(dd_bdry), time=time), gas_model)

    bnd = {"all": PrescribedFluidBoundary(boundary_state_func=bstate)}

    def rhs(t, cv):
        return euler_operator(pb.dcoll, make_fluid_state(cv, gm), gm, bnd, time=t,
                              inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)

    ogm = ph.GasModel(ph.IdealSingleGas())

    def obstate(xb, sm, time):
        return ph.make_fluid_state(
            oini.vortex2d(xb, time, beta=5, center=(5, 0), velocity=(1, 0)), ogm, 2)

    obnd = {"all": oop.PrescribedFluidBoundary(obstate)}

    def orhs(t, q):
        return oop.euler_operator(pb.odisc, pb.box, ph.make_fluid_state(q, ogm, 2),
                                  ogm, obnd, time=t)

    q0 = oini.vortex2d(pb.x, 0.0, beta=5, center=(5, 0), velocity=(1, 0))
    cv = ConservedVars(2, _dev(q0))
    q = q0.copy()
    dt = 0.01
    for step in range(10):
        cv = rk4_step(cv, step * dt, dt, rhs)
        q = oop.rk4_step(q, step * dt, dt, orhs)
    err = rel_err(_host(cv.q), q)
    print("rk4", err)
    assert err.max() < TOL
