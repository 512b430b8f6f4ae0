"""Pins of the oracle's Mengaldo boundary family (oracle/boundaries.py):

* the closed-form boundary states of test/test_bc.py:136-972 (far field,
  pressure outflow sub- and supersonic, isothermal and adiabatic no-slip
  walls, slip wall), restated on boundary nodes;
* the normal-axes utility (test/test_bc.py:77-130), and the frame-free
  slip velocity gradient G - (G n) n^T + n (n^T G n) n^T against the
  reference's rotated construction (where that is defined, n_z != 0);
* operator invariants: a uniform gas at rest in a closed box has a zero
  Navier-Stokes RHS for every wall type, a far field equal to the interior
  and an outflow at the interior pressure give zero RHS, and slip / no-slip
  walls conserve mass (slip walls also energy, for Euler).
"""

import numpy as np
import pytest

from oracle import boundaries as ob, operators as oop, physics as ph
from tests.cases import Problem, smooth_perturbed_state


def _uniform_boundary_state(dim, n, vel, gm, rho=1.0, p=1.0):
    one = np.ones(n)
    v = np.asarray(vel, dtype=np.float64)[:, None] * one
    e = p / 0.4 + 0.5 * rho * np.sum(v * v, axis=0)
    cv = np.concatenate([rho * one[None], e[None], rho * v])
    return ph.make_fluid_state(cv, gm, dim)


def _normals(dim, n, seed):
    rng = np.random.default_rng(seed)
    nv = rng.normal(size=(dim, n))
    if dim == 3:
        nv[2] = np.where(np.abs(nv[2]) < 0.1, 0.5, nv[2])
    return nv / np.sqrt(np.sum(nv * nv, axis=0))


@pytest.mark.parametrize("dim", [2, 3])
def test_normal_axes(dim):
    n = _normals(dim, 50, 1)
    axes = ob.normal_axes(n)
    assert len(axes) == dim
    assert np.allclose(axes[0], n, atol=1e-15)
    for i in range(dim):
        assert np.allclose(np.sum(axes[i] ** 2, axis=0), 1.0, atol=1e-14)
        for j in range(i + 1, dim):
            assert np.abs(np.sum(axes[i] * axes[j], axis=0)).max() < 1e-14
    rot = ob.rotation_matrix(n)
    for i in range(dim):
        assert np.array_equal(rot[i], axes[i])


@pytest.mark.parametrize("dim", [2, 3])
def test_slip_velocity_gradient_projector_form(dim):
    rng = np.random.default_rng(2)
    npts = 40
    n = _normals(dim, npts, 3)
    gm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0))
    sm = _uniform_boundary_state(dim, npts, rng.normal(size=dim), gm)
    gcv = rng.normal(size=(2 + dim, dim, npts))
    gv = ph.velocity_gradient(sm.cv, gcv, dim)
    got = ob.slip_grad_velocity_bc(sm, gcv, n)
    ref = ob.slip_grad_velocity_bc_rotated(sm, gcv, n)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(gv).max()
    # the tangential velocity has no normal derivative any more
    for ax in ob.normal_axes(n)[1:]:
        tgn = np.einsum("i...,ij...,j...->...", ax, got, n)
        assert np.abs(tgn).max() < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_farfield_state(dim):
    """test/test_bc.py:136-301"""
    gm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0),
                     ph.SimpleTransport(viscosity=5.0, thermal_conductivity=3.0))
    ff_t, ff_p = 2.0, 0.5
    ff_v = np.ones(dim)
    bnd = ob.FarfieldBoundary(free_stream_pressure=ff_p,
                              free_stream_velocity=ff_v,
                              free_stream_temperature=ff_t)
    rho = ff_p / ff_t
    for vdir in range(dim):
        for parity in (1.0, -1.0):
            vel = np.zeros(dim)
            vel[vdir] = parity
            n = _normals(dim, 12, vdir)
            sm = _uniform_boundary_state(dim, 12, vel, gm)
            sp = bnd.state_plus(sm, n, gm)
            assert np.allclose(sp.cv[0], rho, rtol=1e-15)
            assert np.allclose(sp.cv[2:], rho * ff_v[:, None], rtol=1e-15)
            assert np.allclose(sp.cv[1], ff_p / 0.4 + 0.5 * rho * dim,
                               rtol=1e-15)
            assert np.all(bnd.temperature_bc(sm) == ff_t)
            assert np.allclose(bnd.state_bc(sm, n, gm).cv, sp.cv, rtol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_outflow_state(dim):
    """test/test_bc.py:304-426: subsonic -> p+ = 2 p_b - p-, supersonic
    (|v.n n| > c) -> p+ = p-; rho and momentum from the interior."""
    gm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0))
    p_b = 1.0 / 1.01
    bnd = ob.PressureOutflowBoundary(boundary_pressure=p_b)
    for vdir in range(dim):
        for parity in (1.0, -1.0):
            n = np.zeros((dim, 6))
            n[vdir] = parity
            vel = np.zeros(dim)
            vel[vdir] = parity
            sm = _uniform_boundary_state(dim, 6, vel, gm)
            sp = bnd.state_plus(sm, n, gm)
            p_plus = 2 * p_b - 1.0
            assert np.allclose(sp.cv[0], 1.0)
            assert np.allclose(sp.cv[2:], vel[:, None])
            assert np.allclose(sp.cv[1], p_plus / 0.4 + 0.5, rtol=1e-15)
            assert np.allclose(sp.pressure, p_plus, rtol=1e-14)
            assert np.allclose(sp.temperature, p_plus, rtol=1e-14)
            # state_bc imposes p_b itself
            sb = bnd.state_bc(sm, n, gm)
            assert np.allclose(sb.pressure, p_b, rtol=1e-14)
            # supersonic: interior pressure kept
            sm2 = _uniform_boundary_state(dim, 6, 1.5 * vel, gm)
            sp2 = bnd.state_plus(sm2, n, gm)
            assert np.allclose(sp2.cv[1], 1.0 / 0.4 + 0.5 * 2.25, rtol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["isothermal", "adiabatic_noslip",
                                  "adiabatic_slip", "isothermal_slip"])
def test_wall_states(dim, kind):
    """test/test_bc.py:429-972: no-slip walls reflect the momentum in
    state_plus and zero it in state_bc; slip walls reflect / drop its normal
    part; isothermal walls impose T_w in T_bc and in the state_bc energy."""
    gm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0),
                     ph.SimpleTransport(viscosity=5.0, thermal_conductivity=3.0))
    bnd = {"isothermal": ob.IsothermalWallBoundary(wall_temperature=2.0),
           "adiabatic_noslip": ob.AdiabaticNoslipWallBoundary(),
           "adiabatic_slip": ob.AdiabaticSlipBoundary(),
           "isothermal_slip": ob.IsothermalSlipWallBoundary(2.0)}[kind]
    rng = np.random.default_rng(5)
    n = _normals(dim, 20, 7)
    sm = _uniform_boundary_state(dim, 20, rng.normal(size=dim), gm)
    sp = bnd.state_plus(sm, n, gm)
    sb = bnd.state_bc(sm, n, gm)
    mom = sm.momentum
    mn = np.sum(mom * n, axis=0)
    slip = "slip" in kind and "noslip" not in kind
    if slip:
        assert np.allclose(sp.momentum, mom - 2 * mn * n, atol=1e-15)
        assert np.abs(np.sum(sb.momentum * n, axis=0)).max() < 1e-15
        assert np.abs(np.sum((sm.momentum + sp.momentum) * n, axis=0)).max() \
            < 1e-15
    else:
        assert np.array_equal(sp.momentum, -mom)
        assert np.all(sb.momentum == 0)
    assert np.array_equal(sp.cv[0], sm.cv[0])
    if kind.startswith("isothermal"):
        assert np.all(bnd.temperature_bc(sm) == 2.0)
        assert np.allclose(sb.temperature, 2.0, rtol=1e-14)
    else:
        assert np.array_equal(bnd.temperature_bc(sm), sm.temperature)
        assert np.allclose(sb.temperature, sm.temperature, rtol=1e-14)


@pytest.mark.parametrize("flux", ["rusanov", "hll"])
@pytest.mark.parametrize("dim", [2, 3])
def test_slipwall_flux(dim, flux):
    """test/test_bc.py:1054-1175: the averaged normal momentum and the mass
    and energy numerical fluxes through a slip wall vanish."""
    gm = ph.GasModel(ph.IdealSingleGas())
    bnd = ob.AdiabaticSlipBoundary()
    func = {"rusanov": ph.inviscid_facial_flux_rusanov,
            "hll": ph.inviscid_facial_flux_hll}[flux]
    rng = np.random.default_rng(9)
    n = _normals(dim, 30, 11)
    sm = _uniform_boundary_state(dim, 30, rng.normal(size=dim), gm)
    sp = bnd.state_plus(sm, n, gm)
    avg = 0.5 * (sm.momentum + sp.momentum)
    assert np.abs(np.sum(avg * n, axis=0)).max() < 1e-15
    f = func(sm, sp, gm, n)
    assert np.abs(f[0]).max() < 1e-14
    assert np.abs(f[1]).max() < 1e-13


def _closed_box(dim, order, n=3):
    return Problem(dim, order, (n,) * dim, 0.0, (1.0, 0.8, 0.6)[:dim],
                   (False,) * dim)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["isothermal", "adiabatic_noslip",
                                  "adiabatic_slip", "isothermal_slip",
                                  "farfield", "outflow"])
def test_uniform_rest_state_zero_rhs(dim, kind):
    """Uniform gas at rest (rho = 1, p = 1, R = 1: T = 1): walls at T_w = T,
    a far field equal to it and an outflow at its pressure leave the NS RHS
    zero (cf. test/test_navierstokes.py:81-235)."""
    pb = _closed_box(dim, 2)
    gm = ph.GasModel(ph.IdealSingleGas(gas_const=1.0),
                     ph.SimpleTransport(viscosity=0.1, thermal_conductivity=0.2))
    bnd = {"isothermal": ob.IsothermalWallBoundary(1.0),
           "adiabatic_noslip": ob.AdiabaticNoslipWallBoundary(),
           "adiabatic_slip": ob.AdiabaticSlipBoundary(),
           "isothermal_slip": ob.IsothermalSlipWallBoundary(1.0),
           "farfield": ob.FarfieldBoundary(1.0, np.zeros(dim), 1.0),
           "outflow": ob.PressureOutflowBoundary(1.0)}[kind]
    x = pb.x
    cv = np.concatenate([np.ones((1,) + x[0].shape),
                         np.full((1,) + x[0].shape, 1.0 / 0.4),
                         np.zeros((dim,) + x[0].shape)])
    st = ph.make_fluid_state(cv, gm, dim)
    rhs = oop.ns_operator(pb.odisc, pb.box, st, gm, {"all": bnd})
    # round-off of O(p / h) = O(10) terms (test/test_euler.py uses 1e-9)
    assert np.abs(rhs).max() < 1e-10


@pytest.mark.parametrize("kind", ["adiabatic_slip", "adiabatic_noslip",
                                  "isothermal"])
def test_wall_conservation(kind):
    """Walls carry no mass flux: sum M rho_t = 0 for Navier-Stokes; a slip
    wall carries no energy flux either, so Euler conserves energy."""
    pb = _closed_box(2, 3, n=4)
    gm = ph.GasModel(ph.IdealSingleGas(),
                     ph.SimpleTransport(viscosity=0.05, thermal_conductivity=0.1))
    bnd = {"isothermal": ob.IsothermalWallBoundary(1.2 / 287.1),
           "adiabatic_noslip": ob.AdiabaticNoslipWallBoundary(),
           "adiabatic_slip": ob.AdiabaticSlipBoundary()}[kind]
    q = smooth_perturbed_state(pb.x, seed=3)
    st = ph.make_fluid_state(q, gm, 2)
    mass = pb.odisc.el.mass
    jac = pb.odisc.J

    def integral(f):
        return float(np.einsum("k,ij,kj->", jac, mass, f))

    rhs = oop.ns_operator(pb.odisc, pb.box, st, gm, {"all": bnd})
    scale = np.abs(rhs[0]).max()
    assert abs(integral(rhs[0])) < 1e-12 * scale
    if kind == "adiabatic_slip":
        gme = ph.GasModel(ph.IdealSingleGas())
        rhs = oop.euler_operator(pb.odisc, pb.box, ph.make_fluid_state(q, gme, 2),
                                 gme, {"all": bnd})
        for f in (0, 1):
            assert abs(integral(rhs[f])) < 1e-12 * np.abs(rhs[f]).max()
