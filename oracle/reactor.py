"""Oracle constant-pressure reactor (NumPy FP64; test infrastructure only).

Restates the adiabatic constant-pressure 0-D reactor that the reference
drives through Cantera's ``IdealGasConstPressureReactor`` in
test/test_chemistry.py:51-110 (mixture ``cantera.Solution(uiuc_7sp)``,
``ReactorNet.advance``), with the chemistry of ``oracle.thermochem`` (the
uiuc_7sp rates of mirgecom/eos.py:858-875):

* dY_k/dt = W_k omega_k(rho, T, Y) / rho, rho = p / (R(Y) T), p constant;
* the mixture enthalpy h(T, Y) = h0 is conserved (adiabatic, constant p),
  T recovered from it by Newton steps T <- T + (h0 - h(T, Y)) / cp(T, Y).

Cantera integrates (T, Y) with CVODE; this restatement (and the GPU reactor
of mirgecom_amd/csrc/mg_chem.hip that it checks) uses classical RK4 on Y
with T from the enthalpy after every stage (4 Newton steps seeded with the
previous stage's T).  Pinned by enthalpy and element conservation, RK4's
fourth-order convergence in dt and the acceptance window of
test/test_chemistry.py:258-260 (tests/test_oracle_reactor.py); the
comparison with Cantera itself is parity unpinned (Cantera absent).
"""

from __future__ import annotations

import numpy as np

from oracle import thermochem as tc

NEWTON = 4


def temperature_from_enthalpy(h0, t, y, niter=NEWTON):
    for _ in range(niter):
        t = t + (h0 - tc.mixture_enthalpy_mass(t, y)) / tc.mixture_cp_mass(t, y)
    return t


def rates(p, t, y):
    """dY/dt at constant pressure."""
    rho = tc.density(p, t, y)
    return tc.species_source_terms(rho, t, y) / rho


def rk4_step(p, h0, dt, t, y):
    """One classical RK4 step of the reactor; returns ``(t, y)``."""
    k1 = rates(p, t, y)
    y1 = y + 0.5 * dt * k1
    t1 = temperature_from_enthalpy(h0, t, y1)
    k2 = rates(p, t1, y1)
    y2 = y + 0.5 * dt * k2
    t2 = temperature_from_enthalpy(h0, t1, y2)
    k3 = rates(p, t2, y2)
    y3 = y + dt * k3
    t3 = temperature_from_enthalpy(h0, t2, y3)
    k4 = rates(p, t3, y3)
    y = y + dt / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    return temperature_from_enthalpy(h0, t3, y), y


def march(t0, p, y0, dt, nsteps, t_rise=400.0):
    """Reactors (arrays over the last axis) marched ``nsteps``; returns
    ``(t, y, t_ign)`` with t_ign the first time T > t0 + t_rise (-1 if
    never)."""
    t = np.array(t0, dtype=np.float64)
    y = np.array(y0, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    h0 = tc.mixture_enthalpy_mass(t, y)
    t_ign = np.full(t.shape, -1.0)
    for s in range(nsteps):
        t, y = rk4_step(p, h0, dt, t, y)
        new = (t_ign < 0) & (t > np.asarray(t0) + t_rise)
        t_ign[new] = (s + 1) * dt
    return t, y, t_ign


def substep(states, dt, nsub):
    """Change of Y over dt from states ``[n, 10]`` = (T, p, Y, dt)."""
    t = states[:, 0].copy()
    p = states[:, 1]
    y = states[:, 2:9].T.copy()
    y0 = y.copy()
    h0 = tc.mixture_enthalpy_mass(t, y)
    for _ in range(nsub):
        t, y = rk4_step(p, h0, dt / nsub, t, y)
    return (y - y0).T
