"""GPU parity of the entropy-stable flux-differencing Euler operator
(euler_operator with use_esdg=True, quadrature_tag=DISCR_TAG_QUAD;
mg_euler_rhs_esdg) against the oracle (oracle/esdg.py): periodic boxes at
p = 1..4 in 2-D and 3-D with the entropy-stable Rusanov and the
entropy-conserving face flux, passive species, prescribed and dummy
boundaries; and the defining property on the GPU result itself: discrete
entropy conservation without face dissipation.  3-D p = 3 runs the paired
pair loop (each two-point flux once for both rows).  Tolerance 1e-11 relative per
field (1e-10 at p = 4, the oracle's monomial-basis round-off)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import Problem, rel_err, smooth_perturbed_state  # noqa: E402

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


def _run(pb, q, dim, dissipation=True, bnd=None, nsp=0):
    from mirgecom_amd.dof_desc import DISCR_TAG_QUAD
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.inviscid import (
        entropy_stable_inviscid_facial_flux,
        entropy_stable_inviscid_facial_flux_rusanov)
    gm = GasModel(eos=IdealSingleGas(gamma=GAMMA, gas_const=1.0))
    st = make_fluid_state(ConservedVars(dim, _dev(q)), gm)
    fn = (entropy_stable_inviscid_facial_flux_rusanov if dissipation
          else entropy_stable_inviscid_facial_flux)
    return _host(euler_operator(pb.dcoll, st, gm, bnd or {}, use_esdg=True,
                                quadrature_tag=DISCR_TAG_QUAD,
                                inviscid_numerical_flux_func=fn).q)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1),
                                       (3, 2), (3, 3)])
@pytest.mark.parametrize("dissipation", [True, False])
def test_periodic(dim, order, dissipation):
    from oracle.esdg import esdg_euler_operator
    n = (4, 3) if dim == 2 else (3, 3, 3)
    pb = Problem(dim, order, n, 0.0, (1.0, 0.8, 0.9)[:dim], (True,) * dim)
    q = smooth_perturbed_state(pb.x, seed=4 + order, amp=0.15)
    print(f"esdg periodic dim={dim} p={order} dissipation={dissipation}",
          flush=True)
    got = _run(pb, q, dim, dissipation)
    ref = esdg_euler_operator(pb.odisc, pb.box, q, GAMMA, {},
                              dissipation=dissipation)
    err = rel_err(got, ref)
    print("err", err)
    assert err.max() < (1e-10 if order >= 4 else 1e-11)
    if not dissipation:
        from tests.test_oracle_esdg import _entropy_rate
        rate, scale = _entropy_rate(pb, q, got)
        assert abs(rate) < 1e-12 * scale


@pytest.mark.parametrize("order", [2, 3])
def test_passive_species(order):
    from oracle.esdg import esdg_euler_operator
    pb = Problem(3, order, (3, 3, 3), 0.0, 1.0, (True,) * 3)
    q = smooth_perturbed_state(pb.x, seed=9, nspecies=7, amp=0.1)
    print(f"esdg passive species p={order}", flush=True)
    got = _run(pb, q, 3)
    ref = esdg_euler_operator(pb.odisc, pb.box, q, GAMMA, {})
    assert rel_err(got, ref).max() < 1e-11


@pytest.mark.parametrize("kind", ["prescribed", "dummy"])
def test_boundaries(kind):
    from mirgecom_amd.boundary import DummyBoundary, PrescribedFluidBoundary
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.initializers import Vortex2D
    from oracle import initializers as oini, operators as oop, physics as ph
    from oracle.esdg import esdg_euler_operator
    pb = Problem(2, 3, (6, 5), -5.0, 5.0, (False, False))
    vortex = Vortex2D(center=[0, 0], velocity=[0, 0])
    ogm = ph.GasModel(ph.IdealSingleGas(gamma=GAMMA, gas_const=1.0))
    q = oini.vortex2d(pb.x, beta=5)

    def gstate(dcoll, dd_bdry, gas_model, state_minus, time=0.0, **kw):
        return make_fluid_state(vortex(dcoll.nodes(dd_bdry), time=time,
                                       eos=gas_model.eos), gas_model)

    def ostate(xb, sm, t):
        return ph.make_fluid_state(oini.vortex2d(xb, t, beta=5), ogm, 2)

    if kind == "prescribed":
        gb = {"all": PrescribedFluidBoundary(boundary_state_func=gstate)}
        ob = {"all": oop.PrescribedFluidBoundary(ostate)}
    else:
        gb, ob = {"all": DummyBoundary()}, {"all": oop.DummyBoundary()}
    print(f"esdg boundaries kind={kind}", flush=True)
    got = _run(pb, q, 2, bnd=gb)
    ref = esdg_euler_operator(pb.odisc, pb.box, q, GAMMA, ob)
    err = rel_err(got, ref)
    print("err", err)
    assert err.max() < 1e-11


def test_3d_p4_unsupported():
    """3-D p = 4 has no parity coverage: the library refuses it."""
    from mirgecom_amd._lib import MgError
    pb = Problem(3, 4, (3, 3, 3), 0.0, 2 * np.pi, (True,) * 3)
    q = smooth_perturbed_state(pb.x, seed=5)
    with pytest.raises(MgError, match="3-D p = 4"):
        _run(pb, q, 3)
