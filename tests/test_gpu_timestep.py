"""GPU node-local timestep (mg_local_timestep) against the oracle, the
reference's own check test/test_viscous.py:499-555, and get_sim_timestep's
three modes (mirgecom/simutil.py:224-306).  Tolerance 1e-13 relative."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import Problem, mixture_state, smooth_perturbed_state  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("transport", [None, "simple", "powerlaw"])
def test_timestep_ideal_species(transport):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.inviscid import get_inviscid_cfl, get_inviscid_timestep
    from mirgecom_amd.transport import PowerLawTransport, SimpleTransport
    from mirgecom_amd.viscous import get_viscous_timestep
    from oracle import operators as oop, physics as ph

    pb = Problem(3, 2, (3, 3, 2), 0.0, (1.0, 0.8, 0.6), (False, True, False))
    q0 = smooth_perturbed_state(pb.x, seed=21, nspecies=7)
    d = np.linspace(0.01, 0.07, 7)
    gt_ = ot_ = None
    if transport == "simple":
        gt_ = SimpleTransport(viscosity=0.03, thermal_conductivity=0.05,
                              species_diffusivity=d)
        ot_ = ph.SimpleTransport(viscosity=0.03, thermal_conductivity=0.05,
                                 species_diffusivity=d)
    elif transport == "powerlaw":
        gt_ = PowerLawTransport(beta=0.01, n=0.7, species_diffusivity=d)
        ot_ = ph.PowerLawTransport(beta=0.01, n=0.7, species_diffusivity=d)
    gm = GasModel(eos=IdealSingleGas(gas_const=1.0), transport=gt_)
    st = make_fluid_state(ConservedVars(3, _dev(q0)), gm)
    ogm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0), ot_)
    ost = ph.make_fluid_state(q0, ogm, 3)
    dt_v = _host(get_viscous_timestep(pb.dcoll, st))
    dt_i = _host(get_inviscid_timestep(pb.dcoll, st))
    ref_v = oop.get_viscous_timestep(pb.odisc, ost)
    ref_i = oop.get_inviscid_timestep(pb.odisc, ost)
    print("timestep", transport, _rel(dt_v, ref_v), _rel(dt_i, ref_i))
    assert _rel(dt_v, ref_v) < 1e-13
    assert _rel(dt_i, ref_i) < 1e-13
    if transport is not None:
        assert np.all(dt_v < dt_i)
    cfl = _host(get_inviscid_cfl(pb.dcoll, st, 1e-3))
    assert _rel(cfl, 1e-3 / ref_i) < 1e-13


def test_timestep_mixture():
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.transport import SimpleTransport
    from mirgecom_amd.viscous import get_viscous_timestep
    from oracle import operators as oop, physics as ph, thermochem as tc

    pb = Problem(3, 2, (3, 3, 3), 0.0, 2 * np.pi, (True,) * 3)
    q, t = mixture_state(pb.x, seed=31)
    d = np.linspace(1e-5, 3e-5, 7)
    gm = GasModel(eos=PyrometheusMixture(),
                  transport=SimpleTransport(viscosity=2e-5,
                                            thermal_conductivity=0.05,
                                            species_diffusivity=d))
    st = make_fluid_state(ConservedVars(3, _dev(q)), gm,
                          temperature_seed=_dev(t + 3.0))
    ogm = ph.GasModel(tc.PyrometheusMixture(),
                      ph.SimpleTransport(viscosity=2e-5,
                                         thermal_conductivity=0.05,
                                         species_diffusivity=d))
    ost = ph.make_fluid_state(q, ogm, 3, temperature_seed=t + 3.0)
    err = _rel(_host(get_viscous_timestep(pb.dcoll, st)),
               oop.get_viscous_timestep(pb.odisc, ost))
    print("mixture timestep", err)
    assert err < 1e-13


@pytest.mark.parametrize("vel", [0, 1])
@pytest.mark.parametrize("mu", [-1, 0, 1, 2])
def test_viscous_timestep_reference_case(mu, vel):
    """test/test_viscous.py:499-555 at order 1: uniform state with c = 1,
    mu a multiple of the smallest length scale (no transport if negative);
    dt = h / (|v| + c + mu / h)."""
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import make_conserved
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.transport import SimpleTransport
    from mirgecom_amd.viscous import get_viscous_timestep

    dim = 3
    pb = Problem(dim, 1, (4,) * dim, 1.0, 2.0, (False,) * dim)
    K, npe = pb.dcoll.nelements, pb.dcoll.refel.nnodes
    ones = torch.ones((K, npe), dtype=torch.float64, device="cuda")
    chlen = _dev(np.repeat(pb.dcoll.lengthscales[:, None], npe, axis=1))
    mu = mu * float(chlen.min())
    tv = None
    if mu >= 0:
        tv = SimpleTransport(viscosity=mu)
    else:
        mu = 0.0
    gm = GasModel(eos=IdealSingleGas(), transport=tv)
    cv = make_conserved(dim, mass=ones, energy=ones / (1.4 * 0.4),
                        momentum=torch.stack([ones * vel] * dim))
    st = make_fluid_state(cv, gm)
    dt = get_viscous_timestep(pb.dcoll, st)
    expected = chlen / (st.wavespeed + mu / chlen)
    assert float(((expected - dt) / expected).abs().max()) < 1e-15


def test_sim_timestep_modes():
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.simutil import get_sim_timestep
    from mirgecom_amd.transport import SimpleTransport
    from mirgecom_amd.viscous import get_viscous_timestep

    pb = Problem(2, 3, (4, 4), 0.0, 1.0, (True, True))
    q0 = smooth_perturbed_state(pb.x, seed=41)
    gm = GasModel(eos=IdealSingleGas(),
                  transport=SimpleTransport(viscosity=0.01))
    st = make_fluid_state(ConservedVars(2, _dev(q0)), gm)
    vdt = _host(get_viscous_timestep(pb.dcoll, st))
    dt = get_sim_timestep(pb.dcoll, st, t=0.0, dt=1.0, cfl=0.5, t_final=10.0,
                          constant_cfl=True)
    assert dt == pytest.approx(0.5 * vdt.min(), rel=1e-15)
    dt = get_sim_timestep(pb.dcoll, st, t=9.99999, dt=1.0, cfl=0.5,
                          t_final=10.0, constant_cfl=True)
    assert dt == pytest.approx(1e-5)
    ldt = _host(get_sim_timestep(pb.dcoll, st, t=0.0, dt=1.0, cfl=0.5,
                                 local_dt=True))
    assert ldt.shape == vdt.shape
    assert np.all(ldt == 0.5 * vdt.min(axis=1, keepdims=True))


def test_local_dt_rk4_device_state():
    """advance_state(local_dt=True) with rk4_step on a device ConservedVars:
    the per-node dt (K, Np) takes the tensor path and matches the same march
    done in NumPy on the host copy."""
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.simutil import get_sim_timestep
    from mirgecom_amd.steppers import advance_state

    pb = Problem(2, 2, (4, 4), 0.0, 1.0, (True, True))
    q0 = smooth_perturbed_state(pb.x, seed=43)
    gm = GasModel(eos=IdealSingleGas())

    def rhs(t, cv):
        return euler_operator(pb.dcoll, make_fluid_state(cv, gm), gm, {},
                              time=t)

    st = make_fluid_state(ConservedVars(2, _dev(q0)), gm)
    ldt = get_sim_timestep(pb.dcoll, st, t=0.0, dt=1.0, cfl=0.3,
                           local_dt=True)
    assert tuple(ldt.shape) == tuple(q0.shape[1:])
    _, _, out = advance_state(rhs, rk4_step, st.cv, t_final=0.0, dt=ldt,
                              local_dt=True, max_steps=3)
    # the same march with host arithmetic between the device RHS calls
    y = q0.copy()
    d = _host(ldt)
    for _ in range(3):
        def k(yy, t):
            return _host(rhs(t, ConservedVars(2, _dev(yy))).q)
        k1 = k(y, 0.0)
        k2 = k(y + d / 2 * k1, 0.0)
        k3 = k(y + d / 2 * k2, 0.0)
        k4 = k(y + d * k3, 0.0)
        y = y + d / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert _rel(_host(out.q), y) < 1e-13


def test_rk4_rejects_malformed_rhs():
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.integrators import rk4_step

    y = ConservedVars(2, torch.ones((4, 3, 6), dtype=torch.float64,
                                    device="cuda"))
    with pytest.raises(TypeError):
        rk4_step(y, 0.0, 0.1, lambda t, s: ConservedVars(2, s.q.float()))
    with pytest.raises(ValueError):
        rk4_step(y, 0.0, 0.1, lambda t, s: ConservedVars(2, s.q[:, :2]))
    # a non-contiguous RHS of the right shape is accepted (copied)
    out = rk4_step(y, 0.0, 0.1,
                   lambda t, s: ConservedVars(2, (-s.q.transpose(1, 2))
                                              .contiguous().transpose(1, 2)))
    assert torch.allclose(out.q, torch.full_like(y.q, float(np.exp(-0.1))),
                          rtol=1e-6)
