"""The uiuc_7sp ethylene-air mechanism as numeric tables.

Transcribed from mirgecom/mechanisms/uiuc_7sp.yaml:1-154 (species
C2H4, O2, CO2, CO, H2O, H2, N2 in the phase's order, NASA7 two-range
polynomials, three reactions in cm/mol/cal units) and converted to the SI
kmol/m^3 units in which Cantera/pyrometheus evaluate it:

* molecular weights from the element masses H 1.008, C 12.011, N 14.007,
  O 15.999 kg/kmol;
* gas constant 8314.46261815324 J/(kmol K), standard pressure 101325 Pa;
* Arrhenius pre-factors scaled by ``(1e-3)**(order - 1)`` (cm^3/mol ->
  m^3/kmol) for reaction orders 1.15, 1.5, 1.25; activation energies
  ``Ea/R = Ea[cal/mol] * 4184 / R``.

Reactions (uiuc_7sp.yaml:146-154):
  R1  C2H4 + O2 => 2 CO + 2 H2     A=1e10, b=0,   Ea=36000, orders C2H4^0.5 O2^0.65
  R2  CO + 0.5 O2 <=> CO2          A=1.1e7, b=0.7, Ea=11000 (reversible)
  R3  H2 + 0.5 O2 => H2O           A=7e8,  b=0,   Ea=12000, orders H2^0.75 O2^0.5
"""

from __future__ import annotations

import math

import numpy as np

SPECIES = ("C2H4", "O2", "CO2", "CO", "H2O", "H2", "N2")
NSPECIES = 7

_ATOMIC = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999}
_COMPOSITION = {
    "C2H4": {"C": 2, "H": 4}, "O2": {"O": 2}, "CO2": {"C": 1, "O": 2},
    "CO": {"C": 1, "O": 1}, "H2O": {"H": 2, "O": 1}, "H2": {"H": 2},
    "N2": {"N": 2},
}

# NASA7 [low range, high range], T_mid = 1000 K for all species
_NASA7 = {
    "O2": ([3.78245636, -2.99673416e-03, 9.84730201e-06, -9.68129509e-09,
            3.24372837e-12, -1063.94356, 3.65767573],
           [3.28253784, 1.48308754e-03, -7.57966669e-07, 2.09470555e-10,
            -2.16717794e-14, -1088.45772, 5.45323129]),
    "H2O": ([4.19864056, -2.0364341e-03, 6.52040211e-06, -5.48797062e-09,
             1.77197817e-12, -3.02937267e+04, -0.849032208],
            [3.03399249, 2.17691804e-03, -1.64072518e-07, -9.7041987e-11,
             1.68200992e-14, -3.00042971e+04, 4.9667701]),
    "CO": ([3.57953347, -6.1035368e-04, 1.01681433e-06, 9.07005884e-10,
            -9.04424499e-13, -1.4344086e+04, 3.50840928],
           [2.71518561, 2.06252743e-03, -9.98825771e-07, 2.30053008e-10,
            -2.03647716e-14, -1.41518724e+04, 7.81868772]),
    "CO2": ([2.35677352, 8.98459677e-03, -7.12356269e-06, 2.45919022e-09,
             -1.43699548e-13, -4.83719697e+04, 9.90105222],
            [3.85746029, 4.41437026e-03, -2.21481404e-06, 5.23490188e-10,
             -4.72084164e-14, -4.8759166e+04, 2.27163806]),
    "C2H4": ([3.95920148, -7.57052247e-03, 5.70990292e-05, -6.91588753e-08,
              2.69884373e-11, 5089.77593, 4.09733096],
             [2.03611116, 0.0146454151, -6.71077915e-06, 1.47222923e-09,
              -1.25706061e-13, 4939.88614, 10.3053693]),
    "N2": ([3.298677, 1.4082404e-03, -3.963222e-06, 5.641515e-09,
            -2.444854e-12, -1020.8999, 3.950372],
           [2.92664, 1.4879768e-03, -5.68476e-07, 1.0097038e-10,
            -6.753351e-15, -922.7977, 5.980528]),
    "H2": ([2.34433112, 7.98052075e-03, -1.9478151e-05, 2.01572094e-08,
            -7.37611761e-12, -917.935173, 0.683010238],
           [3.3372792, -4.94024731e-05, 4.99456778e-07, -1.79566394e-10,
            2.00255376e-14, -950.158922, -3.20502331]),
}

GAS_CONSTANT = 8314.46261815324
ONE_ATM = 101325.0
T_MID = 1000.0

# (A [cm, mol, s], b, Ea [cal/mol], total forward order)
_REACTIONS = [(1.0e10, 0.0, 36000.0, 1.15),
              (1.1e7, 0.7, 11000.0, 1.5),
              (7.0e8, 0.0, 12000.0, 1.25)]


# Lennard-Jones transport parameters of each species (the yaml's ``transport``
# blocks): geometry, diameter [Angstrom], well depth eps/k_B [K], dipole moment
# [Debye], polarizability [Angstrom^3], rotational relaxation number at 298 K.
_TRANSPORT = {
    "C2H4": ("nonlinear", 3.971, 280.8, 0.0, 0.0, 1.5),
    "O2": ("linear", 3.458, 107.4, 0.0, 1.6, 3.8),
    "CO2": ("linear", 3.763, 244.0, 0.0, 2.65, 2.1),
    "CO": ("linear", 3.65, 98.1, 0.0, 1.95, 1.8),
    "H2O": ("nonlinear", 2.605, 572.4, 1.844, 0.0, 4.0),
    "H2": ("linear", 2.92, 38.0, 0.0, 0.79, 280.0),
    "N2": ("linear", 3.621, 97.53, 0.0, 1.76, 4.0),
}

# NASA7 validity of each species ([T_min, T_max]); the phase is valid on the
# intersection, which is where the transport fits are made.
_T_RANGE = {"C2H4": (200.0, 3500.0), "O2": (200.0, 3500.0),
            "CO2": (200.0, 3500.0), "CO": (200.0, 3500.0),
            "H2O": (200.0, 3500.0), "H2": (300.0, 5000.0),
            "N2": (300.0, 5000.0)}


def transport_parameters():
    """Per-species ``(geometry, sigma [A], eps/k_B [K], dipole [D],
    polarizability [A^3], z_rot(298))`` in species order."""
    return [_TRANSPORT[s] for s in SPECIES]


def phase_temperature_range():
    """``(T_min, T_max)`` where every species' thermo is valid."""
    return (max(_T_RANGE[s][0] for s in SPECIES),
            min(_T_RANGE[s][1] for s in SPECIES))


def molecular_weights():
    return np.array([sum(_ATOMIC[el] * n for el, n in _COMPOSITION[s].items())
                     for s in SPECIES])


def nasa7_tables():
    lo = np.array([_NASA7[s][0] for s in SPECIES])
    hi = np.array([_NASA7[s][1] for s in SPECIES])
    return lo, hi


def arrhenius_si():
    """``(log A_SI, b, Ea/R)`` per reaction."""
    log_a = np.array([math.log(a) + (order - 1) * math.log(1e-3)
                      for a, _, _, order in _REACTIONS])
    b = np.array([r[1] for r in _REACTIONS])
    ea_r = np.array([r[2] * 4184.0 / GAS_CONSTANT for r in _REACTIONS])
    return log_a, b, ea_r


def packed_mechanism():
    """Constants packed in the order of the device-side mechanism struct."""
    lo, hi = nasa7_tables()
    wt = molecular_weights()
    log_a, b, ea_r = arrhenius_si()
    return np.concatenate([
        lo.ravel(), hi.ravel(), np.full(NSPECIES, T_MID), wt, 1.0 / wt,
        [GAS_CONSTANT, ONE_ATM], log_a, b, ea_r]).astype(np.float64)
