"""GPU parity of the Mengaldo boundary family (mirgecom/boundary.py:
391-1857) against the oracle (oracle/boundaries.py): Navier-Stokes and Euler
right-hand sides on non-periodic boxes with every wall type, through the
fused path, the path from explicit gradients and (looser) the mixed path.
Tolerance 1e-12 relative per field for FP64."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import (  # noqa: E402
    Problem, mixture_state, rel_err, smooth_perturbed_state)

pytestmark = pytest.mark.gpu

TOL = 1e-12
KINDS = ["adiabatic_slip", "isothermal_slip", "isothermal", "adiabatic_noslip",
         "farfield", "outflow"]


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


def _pair(kind, dim, nsp=0, tw=1.1, mix=False):
    from mirgecom_amd import boundary as gb
    from oracle import boundaries as ob
    if mix:
        tw = 1500.0
        y = np.array([0.05, 0.2, 0.02, 0.01, 0.03, 0.002, 0.688])
        ff = dict(free_stream_pressure=101325.0,
                  free_stream_velocity=np.linspace(1.0, 3.0, dim),
                  free_stream_temperature=1400.0, free_stream_mass_fractions=y)
        p_out = 101000.0
    else:
        y = np.full(nsp, 1.0 / max(nsp, 1)) if nsp else None
        ff = dict(free_stream_pressure=1.05, free_stream_velocity=
                  np.linspace(0.2, 0.4, dim), free_stream_temperature=1.2,
                  free_stream_mass_fractions=y)
        p_out = 0.97
    if kind == "adiabatic_slip":
        return gb.AdiabaticSlipBoundary(), ob.AdiabaticSlipBoundary()
    if kind == "isothermal_slip":
        return (gb.IsothermalSlipWallBoundary(tw),
                ob.IsothermalSlipWallBoundary(tw))
    if kind == "isothermal":
        return gb.IsothermalWallBoundary(tw), ob.IsothermalWallBoundary(tw)
    if kind == "adiabatic_noslip":
        return gb.AdiabaticNoslipWallBoundary(), ob.AdiabaticNoslipWallBoundary()
    if kind == "farfield":
        return gb.FarfieldBoundary(**ff), ob.FarfieldBoundary(**ff)
    return gb.PressureOutflowBoundary(p_out), ob.PressureOutflowBoundary(p_out)


def _ideal(dim, order, nsp, kind):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.transport import SimpleTransport
    from oracle import physics as ph

    n = (4, 3) if dim == 2 else (3, 2, 2)
    pb = Problem(dim, order, n, 0.0, (1.0, 0.8, 0.6)[:dim], (False,) * dim)
    q0 = smooth_perturbed_state(pb.x, seed=17, nspecies=nsp)
    d = np.linspace(0.01, 0.03, nsp) if nsp else None  # passive species
    gm = GasModel(eos=IdealSingleGas(gas_const=1.0),
                  transport=SimpleTransport(viscosity=0.02, bulk_viscosity=0.01,
                                            thermal_conductivity=0.03,
                                            species_diffusivity=d))
    ogm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0),
                      ph.SimpleTransport(viscosity=0.02, bulk_viscosity=0.01,
                                         thermal_conductivity=0.03,
                                         species_diffusivity=d))
    gbc, obc = _pair(kind, dim, nsp)
    st = make_fluid_state(ConservedVars(dim, _dev(q0)), gm)
    ost = ph.make_fluid_state(q0, ogm, dim)
    return pb, gm, ogm, st, ost, {"all": gbc}, {"all": obc}


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dim,order,nsp", [(2, 3, 0), (3, 2, 7)])
def test_ns_walls(kind, dim, order, nsp):
    from mirgecom_amd.inviscid import inviscid_facial_flux_hll
    from mirgecom_amd.navierstokes import ns_operator
    from oracle import operators as oop, physics as ph

    pb, gm, ogm, st, ost, gbnd, obnd = _ideal(dim, order, nsp, kind)
    ref = oop.ns_operator(pb.odisc, pb.box, ost, ogm, obnd)
    rhs, gcv, gt = ns_operator(pb.dcoll, gm, st, gbnd, return_gradients=True)
    err = rel_err(_host(rhs.q), ref)
    print("ns walls fused", kind, dim, err)
    assert err.max() < TOL
    # gradients: CV_bc / T_bc boundary values
    ops = oop.OperatorStates(pb.odisc, pb.box, ost, ogm, obnd, 0.0)
    gref = oop.grad_cv_operator(ops)
    # rho = 1 +- 0.05: grad rho is ~20x smaller than the values whose
    # differences form it, so its round-off floor is ~20x the RHS's
    assert rel_err(_host(gcv.q).reshape(gref.shape[0], -1),
                   gref.reshape(gref.shape[0], -1)).max() < 20 * TOL
    # from explicit gradients (mg_ns_rhs), with HLL
    rhs2 = ns_operator(pb.dcoll, gm, st, gbnd, grad_cv=gcv, grad_t=gt,
                       inviscid_numerical_flux_func=inviscid_facial_flux_hll)
    ref2 = oop.ns_operator(pb.odisc, pb.box, ost, ogm, obnd,
                           inviscid_numerical_flux_func=ph.inviscid_facial_flux_hll)
    err2 = rel_err(_host(rhs2.q), ref2)
    print("ns walls from gradients, hll", kind, dim, err2)
    assert err2.max() < TOL
    # mixed precision
    rhs3 = ns_operator(pb.dcoll, gm, st, gbnd, precision="mixed")
    err3 = rel_err(_host(rhs3.q), ref)
    print("ns walls mixed", kind, dim, err3)
    assert err3.max() < 1e-3


@pytest.mark.parametrize("kind", ["adiabatic_slip", "farfield", "outflow",
                                  "isothermal"])
def test_euler_walls(kind):
    from mirgecom_amd.euler import euler_operator
    from oracle import operators as oop

    pb, gm, ogm, st, ost, gbnd, obnd = _ideal(2, 3, 0, kind)
    rhs = euler_operator(pb.dcoll, st, gm, gbnd)
    ref = oop.euler_operator(pb.odisc, pb.box, ost, ogm, obnd)
    err = rel_err(_host(rhs.q), ref)
    print("euler walls", kind, err)
    assert err.max() < TOL


@pytest.mark.parametrize("kind", ["adiabatic_slip", "isothermal", "farfield",
                                  "outflow"])
def test_ns_walls_mixture(kind):
    """uiuc_7sp mixture: Newton temperatures of state_plus / state_bc seeded
    with the interior temperature; impermeable species gradients."""
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    from oracle import operators as oop, physics as ph, thermochem as tc

    pb = Problem(3, 2, (3, 2, 2), 0.0, (0.01, 0.008, 0.006), (False,) * 3)
    q, t = mixture_state(pb.x * 300.0, seed=23)
    d = np.linspace(1e-5, 3e-5, 7)
    gm = GasModel(eos=PyrometheusMixture(),
                  transport=SimpleTransport(viscosity=2e-5,
                                            thermal_conductivity=0.05,
                                            species_diffusivity=d))
    ogm = ph.GasModel(tc.PyrometheusMixture(),
                      ph.SimpleTransport(viscosity=2e-5,
                                         thermal_conductivity=0.05,
                                         species_diffusivity=d))
    gbc, obc = _pair(kind, 3, mix=True)
    st = make_fluid_state(ConservedVars(3, _dev(q)), gm,
                          temperature_seed=_dev(t + 2.0))
    ost = ph.make_fluid_state(q, ogm, 3, temperature_seed=t + 2.0)
    with pytest.warns(UserWarning):
        rhs = ns_operator(pb.dcoll, gm, st, {"all": gbc})
    ref = oop.ns_operator(pb.odisc, pb.box, ost, ogm, {"all": obc})
    err = rel_err(_host(rhs.q), ref)
    print("ns walls mixture", kind, err)
    assert err.max() < 1e-11
    # mixed precision (FP32 gradients, node-parallel FP32 net fluxes with the
    # p_ref shift, shifted exterior fluxes)
    with pytest.warns(UserWarning):
        rhs3 = ns_operator(pb.dcoll, gm, st, {"all": gbc}, precision="mixed")
    err3 = rel_err(_host(rhs3.q), ref)
    print("ns walls mixture mixed", kind, err3)
    assert err3.max() < 1e-3
