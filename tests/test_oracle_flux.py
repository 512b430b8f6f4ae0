"""Known-answer tests of the oracle's pointwise physics.

Restated from the reference's own tests: Rusanov (test/test_flux.py:57-201),
HLL over its wave configurations (test/test_flux.py:204-407), the inviscid
flux (test/test_inviscid.py:64-142), the viscous stress tensor for a linear
velocity field (test/test_viscous.py:63-118), the species diffusive flux
(test/test_viscous.py:281-355) and the ideal-gas EOS on the isentropic
vortex, p = rho^gamma (test/test_eos.py:528-624)."""

import numpy as np
import pytest

from oracle import initializers as oini
from oracle import physics as ph

GAMMA = 1.4


def _pair_states(dim, nspecies, vel_mag):
    p0, rho0 = 15.0, 2.0
    p1, rho1 = 20.0, 4.0
    vel0 = np.array([3.0 * vel_mag + (i + 1) for i in range(dim)])
    vel1 = np.array([1.0 * vel_mag + (i + 1) for i in range(dim)])
    y0 = np.array([float(i + 1) for i in range(nspecies)])
    y1 = np.array([2.0 * (i + 1) for i in range(nspecies)])
    gm = ph.GasModel(ph.IdealSingleGas(gamma=GAMMA))

    def cv(rho, vel, p, y):
        e = p / (GAMMA - 1) + 0.5 * rho * vel @ vel
        return np.concatenate([[rho, e], rho * vel, rho * y])[:, None]

    s0 = ph.make_fluid_state(cv(rho0, vel0, p0, y0), gm, dim)
    s1 = ph.make_fluid_state(cv(rho1, vel1, p1, y1), gm, dim)
    return gm, (p0, rho0, vel0, y0), (p1, rho1, vel1, y1), s0, s1


def _exact_normal_flux(p, rho, vel, y, normal):
    vn = vel @ normal
    e = p / (GAMMA - 1) + 0.5 * rho * vel @ vel
    mom = np.array([rho * vel[i] * vn + p * normal[i] for i in range(len(vel))])
    return np.concatenate([[rho * vn, (e + p) * vn], mom, rho * y * vn])


@pytest.mark.parametrize("nspecies", [0, 1, 10])
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("norm_dir", [1, -1])
@pytest.mark.parametrize("vel_mag", [0, 1, -1])
def test_rusanov_kat(nspecies, dim, norm_dir, vel_mag):
    gm, a, b, s0, s1 = _pair_states(dim, nspecies, vel_mag)
    normal = norm_dir * np.ones(dim) / np.sqrt(dim)
    got = ph.inviscid_facial_flux_rusanov(s0, s1, gm, normal[:, None])[:, 0]
    c0 = np.sqrt(GAMMA * a[0] / a[1])
    c1 = np.sqrt(GAMMA * b[0] / b[1])
    lam = max(np.linalg.norm(a[2]) + c0, np.linalg.norm(b[2]) + c1)
    q0, q1 = s0.cv[:, 0], s1.cv[:, 0]
    exact = 0.5 * (_exact_normal_flux(*a, normal) + _exact_normal_flux(*b, normal)
                   - lam * (q1 - q0))
    assert np.abs(got - exact).max() < 1e-12 * max(1, np.abs(exact).max())


@pytest.mark.parametrize("vel_mag", [0, 1, 2, -1, -4])
@pytest.mark.parametrize("nspecies", [0, 1, 10])
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("norm_dir", [1, -1])
def test_hll_kat(vel_mag, nspecies, dim, norm_dir):
    gm, a, b, s0, s1 = _pair_states(dim, nspecies, vel_mag)
    normal = norm_dir * np.ones(dim) / np.sqrt(dim)
    got = ph.inviscid_facial_flux_hll(s0, s1, gm, normal[:, None])[:, 0]
    p0, rho0, vel0, _ = a
    p1, rho1, vel1, _ = b
    c0 = np.sqrt(GAMMA * p0 / rho0)
    c1 = np.sqrt(GAMMA * p1 / rho1)
    u0, u1 = vel0 @ normal, vel1 @ normal
    p_star = 0.5 * (p0 + p1) + 0.125 * (u0 - u1) * (rho0 + rho1) * (c0 + c1)
    q0f = np.sqrt(1 + (GAMMA + 1) / (2 * GAMMA) * (p_star / p0 - 1)) \
        if p_star > p0 else 1.0
    q1f = np.sqrt(1 + (GAMMA + 1) / (2 * GAMMA) * (p_star / p1 - 1)) \
        if p_star > p1 else 1.0
    s_m = u0 - c0 * q0f
    s_p = u1 + c1 * q1f
    f0 = _exact_normal_flux(*a, normal)
    f1 = _exact_normal_flux(*b, normal)
    q0, q1 = s0.cv[:, 0], s1.cv[:, 0]
    if s_p <= 0:
        exact = f1
    elif s_m >= 0:
        exact = f0
    else:
        exact = (s_p * f0 - s_m * f1 + s_p * s_m * (q1 - q0)) / (s_p - s_m)
    assert np.abs(got - exact).max() < 1e-12 * max(1, np.abs(exact).max())


def test_hll_covers_all_branches():
    """The parameter set above exercises the three HLL branches."""
    seen = set()
    for vel_mag in [0, 1, 2, -1, -4]:
        for norm_dir in [1, -1]:
            gm, a, b, s0, s1 = _pair_states(3, 0, vel_mag)
            n = norm_dir * np.ones(3) / np.sqrt(3)
            c0 = np.sqrt(GAMMA * a[0] / a[1])
            c1 = np.sqrt(GAMMA * b[0] / b[1])
            u0, u1 = a[2] @ n, b[2] @ n
            seen.add("left" if u0 - 2 * c0 >= 0 else
                     "right" if u1 + 2 * c1 <= 0 else "star")
    assert len(seen) >= 2


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_inviscid_flux_identity(dim):
    """F(q) = <rho v, (E+p) v, rho v x v + p I, rho Y v> on random states."""
    rng = np.random.default_rng(dim)
    n = 50
    rho = rng.uniform(0.5, 2, n)
    vel = rng.uniform(-1, 1, (dim, n))
    p = rng.uniform(0.5, 2, n)
    y = rng.uniform(0, 1, (3, n))
    e = p / (GAMMA - 1) + 0.5 * rho * np.sum(vel * vel, axis=0)
    cv = np.concatenate([rho[None], e[None], rho * vel, rho * y])
    st = ph.make_fluid_state(cv, ph.GasModel(ph.IdealSingleGas()), dim)
    f = ph.inviscid_flux(st)
    assert np.allclose(f[0], rho * vel, rtol=1e-14)
    assert np.allclose(f[1], (e + p) * vel, rtol=1e-14)
    for i in range(dim):
        for j in range(dim):
            assert np.allclose(f[2 + i, j], rho * vel[i] * vel[j] + p * (i == j),
                               rtol=1e-14, atol=1e-14)
    assert np.allclose(f[2 + dim:], rho * y[:, None] * vel[None], rtol=1e-14)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_viscous_stress_tensor_linear_velocity(dim):
    """v = A x with constant A: grad v = A, tau = mu (A + A^T) + (mu_B -
    2 mu/3) tr(A) I exactly (test/test_viscous.py:63-118)."""
    rng = np.random.default_rng(7)
    amat = rng.uniform(-1, 1, (dim, dim))
    n = 20
    x = rng.uniform(0, 1, (dim, n))
    rho = 1.3 + 0 * x[0]
    vel = amat @ x
    cv = np.concatenate([rho[None], (3.0 + 0 * rho)[None], rho * vel])
    # gradient of the conserved variables: d(rho v_i)/dx_j = rho A_ij
    grad = np.zeros((cv.shape[0], dim, n))
    grad[2:2 + dim] = rho * amat[:, :, None]
    mu, mu_b = 0.7, 0.2
    gm = ph.GasModel(ph.IdealSingleGas(),
                     ph.SimpleTransport(viscosity=mu, bulk_viscosity=mu_b))
    st = ph.make_fluid_state(cv, gm, dim)
    tau = ph.viscous_stress_tensor(st, grad)
    exact = mu * (amat + amat.T) + (mu_b - 2 * mu / 3) * np.trace(amat) * np.eye(dim)
    assert np.abs(tau - exact[:, :, None]).max() < 1e-12


def test_diffusive_flux_sums_to_zero_for_mixtures():
    """J_alpha = -rho (d_a grad Y_a - Y_a sum_b d_b grad Y_b); for a mixture
    state sum_a J_a = 0 when sum Y = 1 (test/test_viscous.py:281-355)."""
    from oracle import thermochem as tc
    rng = np.random.default_rng(3)
    dim, n = 3, 30
    rho = rng.uniform(0.5, 1.5, n)
    y = rng.uniform(0.01, 1, (7, n))
    y /= y.sum(axis=0)
    gy = rng.uniform(-1, 1, (7, dim, n))
    gy -= gy.mean(axis=0)          # sum_k grad Y_k = 0
    t = rng.uniform(800, 2000, n)
    e = tc.mixture_energy_mass(t, y)
    cv = np.concatenate([rho[None], (rho * e)[None], np.zeros((dim, n)), rho * y])
    grad = np.zeros((cv.shape[0], dim, n))
    grad[2 + dim:] = rho * gy      # grad rho = 0
    d = np.linspace(1e-5, 7e-5, 7)
    gm = ph.GasModel(tc.PyrometheusMixture(),
                     ph.SimpleTransport(species_diffusivity=d))
    st = ph.make_fluid_state(cv, gm, dim, temperature_seed=t)
    j = ph.diffusive_flux(st, grad)
    exact = -rho * (d[:, None, None] * gy
                    - y[:, None] * np.sum(d[:, None, None] * gy, axis=0)[None])
    assert np.abs(j - exact).max() < 1e-14
    assert np.abs(j.sum(axis=0)).max() < 1e-16 * 100 * np.abs(j).max() + 1e-20


def test_ideal_gas_vortex_isentropic():
    """On the isentropic vortex p = rho^gamma (test/test_eos.py:528-624)."""
    x = np.random.default_rng(0).uniform(-5, 5, (2, 200))
    cv = oini.vortex2d(x, beta=5)
    eos = ph.IdealSingleGas()
    p = eos.pressure(cv, 2)
    assert np.abs(p - cv[0]**GAMMA).max() < 1e-15
    t = eos.temperature(cv, 2)
    assert np.allclose(p, cv[0] * 287.1 * t, rtol=1e-15)
    c = eos.sound_speed(cv, 2)
    assert np.allclose(c, np.sqrt(GAMMA * p / cv[0]), rtol=1e-15)
