"""Synthetic workloads of the benchmark configurations (mesh, gas model,
initial state, boundaries), built on the device.

* ``vortex_2d``: 2-D Euler isentropic vortex on [0,10]x[-5,5], 32x32 squares
  split into 2 triangles (2,048 triangles), p = 3, ideal gas gamma = 1.4,
  vortex beta = 5 centred at (5, 0) advected at u = (1, 0), exact prescribed
  boundaries (the setting of the reference's examples/vortex.py), RK4 at
  dt = 0.005 to t = 1.  dt = 0.01 is beyond RK4's stability limit for this
  p = 3 mesh (local CFL 0.6 by the h_e / (|v| + c) estimate): the FP64
  oracle diverges within 20 steps there, so the march uses dt = 0.005, for
  which the L-inf error against the exact vortex at t = 1 is 1.8e-2 (the
  same as at dt = 0.002: the spatial error dominates).
* ``taylor_green``: 3-D compressible Navier-Stokes Taylor-Green vortex on the
  periodic cube [0, 2pi]^3, n^3 cells x 6 tets, p = 3, Re = 1600, Ma = 0.1,
  Pr = 0.71, constant viscosity.
* ``reacting_box``: 3-D periodic box of n^3 cells x 6 tets, p = 2, uiuc_7sp
  mixture; T = 1600 K + 500 K s(x) with s a normalised sum of 6 seeded
  Fourier modes (wavenumbers 1-4), p = 101325 Pa, u = 0, stoichiometric
  ethylene/air with the mixture fraction perturbed by +-20 % through 4 more
  modes.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from mirgecom_amd import mechanisms as mech
from mirgecom_amd.discretization import (
    create_discretization_collection, generate_box_mesh)


@dataclass
class Workload:
    name: str
    dcoll: object
    gas_model: object
    cv: object
    boundaries: dict
    temperature: object = None      # mixture temperature (seed for the next)
    dt: float = 0.0


def vortex_2d(n=32, order=3, device="cuda"):
    from mirgecom_amd.boundary import PrescribedFluidBoundary
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.initializers import Vortex2D

    mesh = generate_box_mesh(2, (0.0, -5.0), (10.0, 5.0), (n, n))
    dcoll = create_discretization_collection(mesh, order)
    vortex = Vortex2D(beta=5, center=(5, 0), velocity=(1, 0))
    gm = GasModel(eos=IdealSingleGas(gamma=1.4))

    def bstate(dcoll, dd_bdry, gas_model, state_minus, time=0.0, **kw):
        return make_fluid_state(vortex(dcoll.nodes(dd_bdry), time=time,
                                       eos=gas_model.eos), gas_model)

    cv = vortex(dcoll.nodes(), time=0.0, eos=gm.eos)
    return Workload("vortex_2d", dcoll, gm, cv,
                    {"all": PrescribedFluidBoundary(boundary_state_func=bstate)},
                    dt=0.005), vortex


def taylor_green(n=24, order=3, reynolds=1600.0, mach=0.1, prandtl=0.71):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.gas_model import GasModel
    from mirgecom_amd.initializers import TaylorGreenVortex
    from mirgecom_amd.transport import SimpleTransport

    mesh = generate_box_mesh(3, 0.0, 2 * math.pi, (n,) * 3, periodic=(True,) * 3)
    dcoll = create_discretization_collection(mesh, order)
    tgv = TaylorGreenVortex(mach=mach)
    mu, kappa = tgv.transport_for(reynolds=reynolds, prandtl=prandtl)
    gm = GasModel(eos=IdealSingleGas(gamma=1.4, gas_const=287.1),
                  transport=SimpleTransport(viscosity=mu,
                                            thermal_conductivity=kappa))
    cv = tgv(dcoll.nodes(), eos=gm.eos)
    h = 2 * math.pi / n
    c0 = math.sqrt(1.4 * tgv.p0(1.4) / tgv.rho0)
    dt = 0.3 * h / (order**2 * (tgv.v0 + c0))      # CFL 0.3
    return Workload("taylor_green", dcoll, gm, cv, {}, dt=dt)


def _fourier_field(x, rng, nmodes, kmax=4):
    """Normalised sum of ``nmodes`` random Fourier modes with integer
    wavevectors in [1, kmax]^3 and uniform phases; max |s| = 1."""
    s = torch.zeros_like(x[0])
    for _ in range(nmodes):
        k = rng.integers(1, kmax + 1, size=3)
        ph = float(rng.uniform(0, 2 * np.pi))
        amp = float(rng.uniform(0.5, 1.0))
        s = s + amp * torch.cos(k[0] * x[0] + k[1] * x[1] + k[2] * x[2] + ph)
    return s / s.abs().max()


def stoichiometric_mixture_fraction():
    """Z_st of ethylene / air (O2 + 3.76 N2 by mole)."""
    w = mech.molecular_weights()
    sp = {s: i for i, s in enumerate(mech.SPECIES)}
    fuel = w[sp["C2H4"]]
    oxid = 3.0 * (w[sp["O2"]] + 3.76 * w[sp["N2"]])
    return fuel / (fuel + oxid)


def reacting_box(n=32, order=2, seed=2024, dt=1e-8):
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import make_conserved
    from mirgecom_amd.gas_model import GasModel

    mesh = generate_box_mesh(3, 0.0, 2 * math.pi, (n,) * 3, periodic=(True,) * 3)
    dcoll = create_discretization_collection(mesh, order)
    x = dcoll.nodes()
    rng = np.random.default_rng(seed)
    t = 1600.0 + 500.0 * _fourier_field(x, rng, 6)
    zst = stoichiometric_mixture_fraction()
    z = zst * (1.0 + 0.2 * _fourier_field(x, rng, 4))
    w = mech.molecular_weights()
    sp = {s: i for i, s in enumerate(mech.SPECIES)}
    yo2_air = w[sp["O2"]] / (w[sp["O2"]] + 3.76 * w[sp["N2"]])
    y = torch.zeros((7,) + tuple(t.shape), dtype=t.dtype, device=t.device)
    y[sp["C2H4"]] = z
    y[sp["O2"]] = (1 - z) * yo2_air
    y[sp["N2"]] = (1 - z) * (1 - yo2_air)
    eos = PyrometheusMixture(temperature_niter=5)
    p = 101325.0
    rho = eos.get_density(torch.full_like(t, p), t, y)
    energy = rho * eos.get_internal_energy(t, y)
    cv = make_conserved(3, mass=rho, energy=energy,
                        momentum=torch.zeros((3,) + tuple(t.shape),
                                             dtype=t.dtype, device=t.device),
                        species_mass=rho * y)
    return Workload("reacting_box", dcoll, GasModel(eos=eos), cv, {},
                    temperature=t, dt=dt)
