"""GPU parity of the bound-preserving limiter (mirgecom/limiter.py:39-125,
kernel ``mg_bound_limiter``) against the oracle (oracle/limiter.py), the
reference's own limiter tests (test/test_limiter.py:44-131) restated, and the
``limiter_func`` hook of ``make_fluid_state`` (mirgecom/gas_model.py:373-383).

Tolerance: 1e-13 relative to the field's magnitude (FP64; the cell average
is summed in another order than the oracle's)."""

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import Problem  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                           device="cuda")


def _host(t):
    return t.detach().cpu().numpy()


def _rough_field(pb, seed, nf=1):
    """Smooth field with element-scale noise: some elements inside [0, 1],
    some poking out of either bound, some with averages outside."""
    rng = np.random.default_rng(seed)
    x = pb.x
    out = []
    for _ in range(nf):
        k = rng.uniform(2, 6, size=pb.dim)
        v = 0.5 + 0.7 * np.prod([np.sin(k[i] * x[i] + i) for i in range(pb.dim)],
                                axis=0)
        out.append(v + 0.05 * rng.standard_normal(x[0].shape))
    return np.stack(out)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 3), (2, 4), (3, 1), (3, 2),
                                       (3, 3), (3, 4)])
@pytest.mark.parametrize("mmax,modify", [(None, False), (1.0, False),
                                         (1.0, True), (None, True)])
def test_parity(dim, order, mmax, modify):
    from mirgecom_amd.limiter import bound_preserving_limiter
    from oracle.limiter import bound_preserving_limiter as ref_lim
    n = (7, 5) if dim == 2 else (4, 3, 3)   # K = 70 / 216: ragged blocks
    pb = Problem(dim, order, n, 0.0, 1.0, (False,) * dim)
    field = _rough_field(pb, seed=dim * 10 + order, nf=3)
    print(f"limiter parity dim={dim} order={order} mmax={mmax} "
          f"modify={modify}", flush=True)
    got = _host(bound_preserving_limiter(pb.dcoll, _dev(field), mmin=0.0,
                                         mmax=mmax, modify_average=modify))
    want = ref_lim(pb.odisc, field, mmin=0.0, mmax=mmax,
                   modify_average=modify)
    changed = np.abs(want - field).max(axis=-1) > 1e-12
    assert changed.any() and (~changed).any()
    assert np.abs(got - want).max() < TOL * np.abs(want).max()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_positivity_preserving(dim, order):
    """test/test_limiter.py:86-111."""
    from mirgecom_amd.limiter import bound_preserving_limiter
    pb = Problem(dim, order, (2,) * dim, 0.0, 0.5, (False,) * dim)
    field = pb.x[0] - 0.001
    lim = _host(bound_preserving_limiter(pb.dcoll, _dev(field), mmin=0.0))
    assert lim.min() > -1e-13


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_bound_preserving(dim, order):
    """test/test_limiter.py:114-136."""
    from mirgecom_amd.limiter import bound_preserving_limiter
    pb = Problem(dim, order, (2,) * dim, 0.0, 0.5, (False,) * dim)
    field = 0.5 + pb.x[0] + 0.01
    lim = _host(bound_preserving_limiter(pb.dcoll, _dev(field), mmin=0.0,
                                         mmax=1.0))
    assert lim.max() < 1.0 + 1e-13


def test_object_array_and_in_place_abi():
    """Object arrays of fields are limited entry by entry; the C ABI also
    limits in place (``out == field``) with the same result."""
    from mirgecom_amd import _lib
    from mirgecom_amd.limiter import bound_preserving_limiter, _mass_weights
    from mirgecom_amd.device import get_device_discretization
    pb = Problem(3, 3, (3, 3, 3), 0.0, 1.0, (False,) * 3)
    field = _dev(_rough_field(pb, seed=5, nf=2))
    stacked = _host(bound_preserving_limiter(pb.dcoll, field, mmin=0.0,
                                             mmax=1.0))
    obj = np.empty(2, dtype=object)
    obj[0], obj[1] = field[0], field[1]
    lim = bound_preserving_limiter(pb.dcoll, obj, mmin=0.0, mmax=1.0)
    for i in range(2):
        assert np.array_equal(_host(lim[i]), stacked[i])
    buf = field.clone()
    w = _mass_weights(get_device_discretization(pb.dcoll))
    k, npe = pb.dcoll.nelements, pb.dcoll.refel.nnodes
    _lib.call("mg_bound_limiter", npe, k, 2, _lib.ptr(w), _lib.ptr(buf),
              _lib.ptr(buf), 0.0, 1.0, 1, _lib.stream_handle())
    assert np.array_equal(_host(buf), stacked)
    # zero elements / zero fields are no-ops
    _lib.call("mg_bound_limiter", npe, 0, 2, _lib.ptr(w), _lib.ptr(buf),
              _lib.ptr(buf), 0.0, 1.0, 1, _lib.stream_handle())
    torch.cuda.synchronize()


def test_fluid_api():
    """test/test_limiter.py:44-83: a density limiter as ``limiter_func``.

    The reference's assertion there is vacuous (its own comment says so): a
    uniform density 1 - eps below mmin = 1 has theta = 1 and stays as it is.
    Restated as what the limiter does: unchanged without ``modify_average``,
    lifted to exactly 1 with it."""
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import make_conserved
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.limiter import bound_preserving_limiter
    dim, eps = 2, 1e-3
    pb = Problem(dim, 1, (4, 4), -0.5, 0.5, (False, False))
    ones = torch.ones((pb.dcoll.nelements, 3), dtype=torch.float64,
                      device="cuda")
    eos = IdealSingleGas(gas_const=1.0)
    gas_model = GasModel(eos=eos)
    mom = torch.zeros((dim,) + ones.shape, dtype=torch.float64, device="cuda")
    cv = make_conserved(dim, mass=ones - eps, energy=2.5 * ones, momentum=mom)

    for modify, expect in ((False, 1.0 - eps), (True, 1.0)):
        seen = {}

        def _limit_fluid_cv(cv, temperature_seed, gas_model, dd=None):
            seen["dd"] = dd
            density_lim = bound_preserving_limiter(pb.dcoll, cv.mass, mmin=1.0,
                                                   modify_average=modify)
            return make_conserved(dim=dim, mass=density_lim, energy=cv.energy,
                                  momentum=density_lim * cv.velocity)

        state = make_fluid_state(cv=cv, gas_model=gas_model,
                                 limiter_func=_limit_fluid_cv)
        mass = _host(state.mass_density)
        assert np.abs(mass - expect).max() < 1e-15
        assert "dd" in seen
        p = _host(state.pressure)
        assert np.abs(p - 0.4 * 2.5).max() < 1e-14


def test_species_limiter_func_mixture():
    """The species limiter of examples/gas-in-box.py:463-497 as
    ``limiter_func`` on a uiuc_7sp mixture: mass fractions limited into
    [0, 1] with ``modify_average``, renormalised, density and energy rebuilt
    at the unlimited pressure and temperature, which the function returns
    with the CV as ``[cv, p, T]``."""
    from mirgecom_amd.fluid import make_conserved
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.limiter import bound_preserving_limiter
    from mirgecom_amd.eos import PyrometheusMixture
    dim, nsp = 2, 7
    pb = Problem(dim, 3, (4, 4), 0.0, 1.0, (False, False))
    eos = PyrometheusMixture(temperature_guess=1500.0)
    gas_model = GasModel(eos=eos)
    rng = np.random.default_rng(3)
    shape = (pb.dcoll.nelements, pb.dcoll.refel.nnodes)
    y = np.abs(rng.uniform(0.02, 0.3, size=(nsp,) + shape))
    y /= y.sum(axis=0)
    y[1] -= 0.03   # undershoots below zero in places
    y[6] += 0.03
    t = _dev(1500.0 + 50.0 * pb.x[0])
    yd = _dev(y)
    rho = eos.get_density(_dev(np.full(shape, 101325.0)), t, yd)
    vel = _dev(np.stack([0.5 + 0 * pb.x[0], 0.2 + 0 * pb.x[1]]))
    energy = rho * (eos.get_internal_energy(t, yd)
                    + 0.5 * (vel * vel).sum(dim=0))
    cv = make_conserved(dim, mass=rho, energy=energy, momentum=rho * vel,
                        species_mass=rho * yd)
    assert y.min() < 0

    def _limiter(cv, temperature_seed, gas_model, dd=None):
        temperature = gas_model.eos.temperature(
            cv=cv, temperature_seed=temperature_seed)
        pressure = gas_model.eos.pressure(cv=cv, temperature=temperature)
        spec = bound_preserving_limiter(pb.dcoll, cv.species_mass_fractions,
                                        mmin=0.0, mmax=1.0,
                                        modify_average=True, dd=dd)
        spec = spec / spec.sum(dim=0)
        mass = gas_model.eos.get_density(pressure=pressure,
                                         temperature=temperature,
                                         species_mass_fractions=spec)
        energy = mass * (gas_model.eos.get_internal_energy(
            temperature, species_mass_fractions=spec)
            + 0.5 * (cv.velocity * cv.velocity).sum(dim=0))
        cv_lim = make_conserved(dim=dim, mass=mass, energy=energy,
                                momentum=mass * cv.velocity,
                                species_mass=mass * spec)
        out = np.empty(3, dtype=object)
        out[0], out[1], out[2] = cv_lim, pressure, temperature
        return out

    state = make_fluid_state(cv=cv, gas_model=gas_model,
                             temperature_seed=torch.full_like(t, 1400.0),
                             limiter_func=_limiter)
    yl = _host(state.cv.species_mass_fractions)
    assert yl.min() > -1e-13 and yl.max() < 1.0 + 1e-13
    assert np.abs(yl.sum(axis=0) - 1.0).max() < 1e-13
    assert np.abs(_host(state.temperature) - _host(t)).max() < 1e-7
    assert np.abs(_host(state.pressure) - 101325.0).max() < 1e-6
