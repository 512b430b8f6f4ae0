"""Mixed-precision Euler RHS (mg_euler_rhs_mixed: nodal fluxes relative to
each element's first node, FP32 contraction; physics, face fluxes, state and
RHS in FP64) against the FP64 path (itself checked against the oracle to
1e-12): per-field RHS error for the ideal gas and the uiuc_7sp mixture, with
periodic and prescribed boundaries, and the state after RK4 steps of the
reacting box at dt = 1e-8 s, where the target is <= 1e-6 relative per
field."""

import warnings

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import (  # noqa: E402
    Problem, mixture_state, rel_err, rel_err_cv, smooth_perturbed_state)

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


def _both(dcoll, st, gm, bnd, flux):
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.inviscid import (
        inviscid_facial_flux_hll, inviscid_facial_flux_rusanov)
    fn = {"rusanov": inviscid_facial_flux_rusanov, "hll": inviscid_facial_flux_hll}[flux]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r64 = euler_operator(dcoll, st, gm, bnd, inviscid_numerical_flux_func=fn)
        r32 = euler_operator(dcoll, st, gm, bnd, inviscid_numerical_flux_func=fn,
                             precision="mixed")
    return _host(r32.q), _host(r64.q)


@pytest.mark.parametrize("dim,order", [(3, 3), (3, 1), (2, 4)])
@pytest.mark.parametrize("flux", ["rusanov", "hll"])
def test_ideal_gas(dim, order, flux):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    pb = Problem(dim, order, (4,) * dim, 0.0, 2 * np.pi, (True,) * dim)
    q = smooth_perturbed_state(pb.x, seed=11)
    gm = GasModel(eos=IdealSingleGas())
    st = make_fluid_state(ConservedVars(dim, _dev(q)), gm)
    a, b = _both(pb.dcoll, st, gm, {}, flux)
    err = rel_err_cv(a, b, dim)
    print(f"ideal gas dim={dim} p={order} {flux}: mixed vs FP64 {err}", flush=True)
    assert err.max() < 1e-5


def test_prescribed_boundary():
    from mirgecom_amd.boundary import PrescribedFluidBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    pb = Problem(2, 3, (5, 5), -1.0, 1.0, (False, False))
    q = smooth_perturbed_state(pb.x, seed=12)
    gm = GasModel(eos=IdealSingleGas())
    st = make_fluid_state(ConservedVars(2, _dev(q)), gm)

    def bnd_state(dcoll, dd_bdry, gas_model, state_minus, **kw):
        cv = state_minus.cv
        return make_fluid_state(ConservedVars(2, cv.q * 1.01), gas_model)

    bnd = {"all": PrescribedFluidBoundary(boundary_state_func=bnd_state)}
    a, b = _both(pb.dcoll, st, gm, bnd, "rusanov")
    err = rel_err_cv(a, b, 2)
    print("prescribed boundary: mixed vs FP64", err, flush=True)
    assert err.max() < 1e-5


@pytest.mark.parametrize("flux", ["rusanov", "hll"])
def test_mixture(flux):
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    pb = Problem(3, 2, (4, 4, 4), 0.0, 2 * np.pi, (True,) * 3)
    q, t = mixture_state(pb.x, seed=31)
    gm = GasModel(eos=PyrometheusMixture())
    st = make_fluid_state(ConservedVars(3, _dev(q)), gm, temperature_seed=_dev(t))
    a, b = _both(pb.dcoll, st, gm, {}, flux)
    err = rel_err(a, b)
    print(f"mixture {flux}: mixed vs FP64 {err}", flush=True)
    assert err.max() < 1e-5


def test_reacting_box_march():
    """20 RK4 steps (Newton T from the previous stage, exact sources) of a
    6^3 x 6 reacting box at dt = 1e-8 s: mixed vs FP64 state within 1e-6
    per field (momentum over the largest momentum component, species over
    the largest species row)."""
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.workloads import reacting_box
    wl = reacting_box(n=6, order=2)
    gm = wl.gas_model
    finals = {}
    for prec in ("float64", "mixed"):
        tbox = [wl.temperature.clone()]

        def rhs(t, cv, prec=prec, tbox=tbox):
            st = make_fluid_state(cv, gm, temperature_seed=tbox[0])
            tbox[0] = st.temperature
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                r = euler_operator(wl.dcoll, st, gm, {}, precision=prec,
                                   inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
            return gm.eos.add_species_source_terms(cv, st.temperature, r)

        cv = wl.cv
        for k in range(20):
            cv = rk4_step(cv, k * wl.dt, wl.dt, rhs)
        finals[prec] = _host(cv.q)
    a, b = finals["mixed"], finals["float64"]
    err = rel_err_cv(a, b, 3)
    sp = slice(5, 12)
    err[sp] = np.abs(a[sp] - b[sp]).reshape(7, -1).max(axis=1) / np.abs(b[sp]).max()
    print("reacting box march: mixed vs FP64", err, flush=True)
    assert np.all(np.isfinite(a))
    assert err.max() < 1e-6


@pytest.mark.parametrize("dim,order,n", [(3, 1, (4, 4, 4)), (3, 2, (4, 4, 4)),
                                         (3, 3, (3, 3, 4)), (2, 2, (7, 7)),
                                         (2, 4, (7, 7))])
def test_mixture_column_groups(dim, order, n):
    """The mixed-precision mixture kernels take two contraction columns per
    lane where the LDS allows (3-D p <= 2: 10 elements per RHS workgroup, 9
    per gradient workgroup; 2-D: 11 and 10), one otherwise (3-D p = 3).  The
    element counts here (384, 216, 98) are no multiple of those, so the last
    workgroup is partial.  Euler and Navier-Stokes RHS against FP64 within
    2e-5 (Euler; at p = 4 the FP32 contraction's rounding reaches 1.2e-5 on
    a momentum component, bit for bit as with one column per lane) and 1e-3
    (NS: FP32 gradients and nodal fluxes, the bound of the mixed mixture NS
    tests) per field, with the Rusanov flux, and with HLL for Euler."""
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    pb = Problem(dim, order, n, 0.0, 2 * np.pi, (True,) * dim)
    q, t = mixture_state(pb.x, seed=41 + order)
    gm = GasModel(eos=PyrometheusMixture())
    st = make_fluid_state(ConservedVars(dim, _dev(q)), gm, temperature_seed=_dev(t))
    for flux in ("rusanov", "hll"):
        a, b = _both(pb.dcoll, st, gm, {}, flux)
        err = rel_err_cv(a, b, dim)
        print(f"column groups dim={dim} p={order} K={pb.dcoll.mesh.nelements} "
              f"Euler {flux}: {err}", flush=True)
        assert err.max() < 2e-5
    gmn = GasModel(eos=PyrometheusMixture(), transport=SimpleTransport(
        viscosity=5e-5, thermal_conductivity=0.1, species_diffusivity=np.full(7, 2.3e-4)))
    stn = make_fluid_state(ConservedVars(dim, _dev(q)), gmn, temperature_seed=_dev(t))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r64 = ns_operator(pb.dcoll, gmn, stn, {})
        r32 = ns_operator(pb.dcoll, gmn, stn, {}, precision="mixed")
    err = rel_err(_host(r32.q), _host(r64.q))
    print(f"column groups dim={dim} p={order} NS: {err}", flush=True)
    assert err.max() < 1e-3
