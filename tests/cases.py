"""Problem builders shared by the tests: the same mesh and seeded state for
the GPU path (mirgecom_amd) and the CPU oracle."""

from __future__ import annotations

import numpy as np

from mirgecom_amd.discretization import (
    create_discretization_collection, generate_box_mesh)
from oracle.disc import OracleDisc


class Problem:
    """``own_matrices=False`` (default) runs the oracle with the product's
    reference-element matrices (checked against the oracle's own to 1e-13 in
    tests/test_refelem.py), so that a comparison measures the operator
    assembly and physics rather than matrix-construction roundoff."""

    def __init__(self, dim, order, n, a, b, periodic, own_matrices=False):
        self.mesh = generate_box_mesh(dim, a, b, n, periodic=periodic)
        self.dcoll = create_discretization_collection(self.mesh, order)
        self.box = (self.mesh.a, self.mesh.b, self.mesh.periodic)
        self.odisc = OracleDisc(self.mesh.vertices, self.mesh.elements,
                                self.dcoll.refel.nodes, order, self.box)
        if not own_matrices:
            re = self.dcoll.refel
            self.odisc.use_element_matrices(re.mass, re.diff, re.face_mass)
        self.dim = dim
        self.order = order

    @property
    def x(self):
        """Oracle node coordinates ``(dim, K, Np)``."""
        return np.moveaxis(self.odisc.x, -1, 0)


def rel_err(gpu, ref):
    """Per-field ``max|gpu - ref| / max|ref|`` (absolute where ``ref == 0``)."""
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    out = []
    for f in range(ref.shape[0]):
        scale = np.abs(ref[f]).max()
        err = np.abs(gpu[f] - ref[f]).max()
        out.append(err / scale if scale > 0 else err)
    return np.array(out)


def rel_err_cv(gpu, ref, dim):
    """Per-field relative error with the momentum components normalised by
    the largest momentum component (a vector field; a component that starts
    at zero, like w of the Taylor-Green vortex, would otherwise divide by a
    vanishing norm)."""
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    out = rel_err(gpu, ref)
    mom = slice(2, 2 + dim)
    scale = np.abs(ref[mom]).max()
    out[mom] = np.abs(gpu[mom] - ref[mom]).reshape(dim, -1).max(axis=1) / scale
    return out


def smooth_perturbed_state(x, seed, nspecies=0, amp=0.05, gamma=1.4):
    """A smooth, seeded, non-trivial ideal-gas state (conserved ``(F, ...)``)."""
    rng = np.random.default_rng(seed)
    dim = x.shape[0]

    def field(base, a):
        k = rng.integers(1, 3, size=dim)
        ph = rng.uniform(0, 2 * np.pi, size=dim)
        val = np.ones_like(x[0])
        for i in range(dim):
            val = val * np.cos(k[i] * x[i] + ph[i])
        return base + a * val

    rho = field(1.0, amp)
    vel = [field(0.3 * (i + 1) / dim, amp) for i in range(dim)]
    p = field(1.0, amp)
    e = p / (gamma - 1) + 0.5 * rho * sum(v * v for v in vel)
    rows = [rho, e] + [rho * v for v in vel]
    if nspecies:
        ys = [field(0.5 / nspecies, 0.2 / nspecies) for _ in range(nspecies)]
        rows += [rho * y for y in ys]
    return np.stack(rows)


def mixture_state(x, seed, velocity_amp=5.0, t_mean=1600.0, t_amp=300.0):
    """A seeded uiuc_7sp state (conserved ``(F, ...)``) and its temperature:
    stoichiometric-ish ethylene/air with perturbed mixture fraction and some
    products, T = t_mean + t_amp * smooth field, p = 1 atm."""
    from oracle import thermochem as tc
    rng = np.random.default_rng(seed)
    dim = x.shape[0]

    def smooth():
        k = rng.integers(1, 3, size=dim)
        ph = rng.uniform(0, 2 * np.pi, size=dim)
        val = np.ones_like(x[0])
        for i in range(dim):
            val = val * np.cos(k[i] * x[i] + ph[i])
        return val

    t = t_mean + t_amp * smooth()
    # C2H4, O2, CO2, CO, H2O, H2, N2
    y = np.stack([0.06 * (1 + 0.2 * smooth()), 0.2 * (1 + 0.1 * smooth()),
                  0.02 * (1 + 0.5 * smooth()), 0.01 * (1 + 0.5 * smooth()),
                  0.03 * (1 + 0.5 * smooth()), 0.002 * (1 + 0.5 * smooth()),
                  0 * t])
    y[-1] = 1.0 - y[:-1].sum(axis=0)
    p = 101325.0 * (1 + 0.05 * smooth())
    rho = tc.density(p, t, y)
    vel = [velocity_amp * smooth() for _ in range(dim)]
    e = tc.mixture_energy_mass(t, y)
    energy = rho * (e + 0.5 * sum(v * v for v in vel))
    q = np.stack([rho, energy] + [rho * v for v in vel] + [rho * yk for yk in y])
    return q, t
