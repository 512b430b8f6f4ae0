"""operator_states_quad on the GPU (mirgecom/gas_model.py:728-904 as used by
examples/combozzle.py): the operators given the states of
make_operator_fluid_states return exactly what they return without them;
the states' volume part on the quadrature equals the oracle's projected
state; the entropy-stable mixture operator takes the volume state the given
states carry (plain, or entropy-stable projected) as the reference does
(mirgecom/euler.py:159-171)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import Problem, mixture_state, rel_err, smooth_perturbed_state  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("quad", [False, True])
def test_same_rhs_with_states(quad):
    from mirgecom_amd.boundary import DummyBoundary
    from mirgecom_amd.dof_desc import DISCR_TAG_BASE, DISCR_TAG_QUAD
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state, make_operator_fluid_states
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.navierstokes import grad_cv_operator, ns_operator
    from mirgecom_amd.transport import SimpleTransport
    tag = DISCR_TAG_QUAD if quad else DISCR_TAG_BASE
    pb = Problem(3, 2, (3, 3, 2), 0.0, 1.0, (True, True, False))
    q = smooth_perturbed_state(pb.x, seed=8, amp=0.1)
    gm = GasModel(eos=IdealSingleGas(gas_const=1.0),
                  transport=SimpleTransport(viscosity=0.05, thermal_conductivity=0.07))
    with pytest.warns(UserWarning):
        bnd = {"all": DummyBoundary()}
    st = make_fluid_state(ConservedVars(3, _dev(q)), gm)
    ops = make_operator_fluid_states(pb.dcoll, st, gm, bnd, quadrature_tag=tag)
    kw = dict(quadrature_tag=tag, inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
    a = ns_operator(pb.dcoll, gm, st, bnd, **kw)
    b = ns_operator(pb.dcoll, gm, st, bnd, operator_states_quad=ops, **kw)
    assert torch.equal(a.q, b.q)
    a = euler_operator(pb.dcoll, st, gm, bnd, **kw)
    b = euler_operator(pb.dcoll, st, gm, bnd, operator_states_quad=ops, **kw)
    assert torch.equal(a.q, b.q)
    g1 = grad_cv_operator(pb.dcoll, gm, bnd, st, quadrature_tag=tag)
    g2 = grad_cv_operator(pb.dcoll, gm, bnd, st, quadrature_tag=tag,
                          operator_states_quad=ops)
    assert torch.equal(g1.q, g2.q)
    if quad:
        from oracle import operators as oop, physics as ph
        from oracle.overintegration import QuadStates
        ogm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0))
        qs = QuadStates(pb.odisc, pb.box, q, ogm, {"all": oop.DummyBoundary()}, 0.0, 4)
        vol = ops.vol_state_quad
        got = _host(vol.cv.q).reshape(q.shape[0], -1)
        assert rel_err(got, qs.vol.cv).max() < 1e-12
        assert rel_err(_host(vol.temperature).reshape(1, -1),
                       qs.vol.temperature[None]).max() < 1e-12


@pytest.mark.parametrize("entropy_stable", [False, True])
def test_mixture_esdg_takes_the_given_volume_state(entropy_stable):
    from mirgecom_amd.dof_desc import DISCR_TAG_QUAD
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state, make_operator_fluid_states
    from oracle import physics as ph, thermochem as tc
    from oracle.esdg import esdg_operator
    pb = Problem(2, 2, (4, 3), 0.0, 2 * np.pi, (True, True))
    q, t = mixture_state(pb.x, seed=70)
    gm = GasModel(eos=PyrometheusMixture())
    st = make_fluid_state(ConservedVars(2, _dev(q)), gm, temperature_seed=_dev(t + 2.0))
    ops = make_operator_fluid_states(pb.dcoll, st, gm, {}, quadrature_tag=DISCR_TAG_QUAD,
                                     entropy_stable=entropy_stable)
    with pytest.warns(UserWarning):
        got = _host(euler_operator(pb.dcoll, st, gm, {}, use_esdg=True,
                                   quadrature_tag=DISCR_TAG_QUAD,
                                   operator_states_quad=ops).q)
    ref = esdg_operator(pb.odisc, pb.box, q, ph.GasModel(tc.PyrometheusMixture()), {},
                        temperature=_host(st.temperature), es_volume=entropy_stable)
    err = rel_err(got, ref)
    print("err", err)
    assert err.max() < 1e-10
