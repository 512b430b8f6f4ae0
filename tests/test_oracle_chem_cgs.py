"""The oracle's uiuc_7sp rates against an independent evaluation in the
units of the mechanism file (mirgecom/mechanisms/uiuc_7sp.yaml:5: cm, mol,
cal/mol): concentrations in mol/cm^3, k = A T^b exp(-Ea / (R_cal T)) with
R_cal = 8.31446261815324 / 4.184 cal/(mol K), K_c of R2 with p0 = 1013250 dyn/cm^2
and R = 8.31446261815324e7 erg/(mol K), rates in mol/(cm^3 s) converted to
kmol/(m^3 s) only at the end (x 1e3).  An error in the SI path's order-
dependent A conversion (A 1e3^(1 - order)), its activation-energy units or
its K_c units would fail here; the rate constants are typed in from the
mechanism file's reaction list (uiuc_7sp.yaml:146-154)."""

import numpy as np

from oracle import thermochem as tc

R_CAL = 8.31446261815324 / 4.184     # cal / (mol K)
R_ERG = 8.31446261815324e7           # erg / (mol K)
P0_CGS = 1013250.0                   # dyn / cm^2
# reactions: A [cm, mol, s], b, Ea [cal/mol], forward orders
R1 = (1.0e10, 0.0, 36000.0, {"C2H4": 0.5, "O2": 0.65})
R2 = (1.1e7, 0.7, 11000.0, {"CO": 1.0, "O2": 0.5})
R3 = (7.0e8, 0.0, 12000.0, {"H2": 0.75, "O2": 0.5})
IDX = {s: i for i, s in enumerate(tc.SPECIES)}
W_G = tc.W                           # g/mol (= kg/kmol)


def _cgs_rates(rho_si, t, y):
    rho = rho_si * 1e-3                          # g/cm^3
    c = rho * y / W_G[:, None]                   # mol/cm^3
    c = np.maximum(c, 0.0)
    out = []
    for a, b, ea, orders in (R1, R2, R3):
        k = a * t**b * np.exp(-ea / (R_CAL * t))
        r = k * np.prod([c[IDX[s]] ** o for s, o in orders.items()], axis=0)
        out.append(r)
    # R2 reverse: K_c = exp(-dG/RT) (p0 / (R T))^dnu, dnu = 1 - 1.5
    g = tc.g_rt(t)
    dg = g[IDX["CO2"]] - g[IDX["CO"]] - 0.5 * g[IDX["O2"]]
    kc = np.exp(-dg) * (P0_CGS / (R_ERG * t)) ** (-0.5)
    a, b, ea, _ = R2
    kf2 = a * t**b * np.exp(-ea / (R_CAL * t))
    out[1] = out[1] - kf2 / kc * c[IDX["CO2"]]
    return np.array(out) * 1e3                   # kmol / (m^3 s)


def test_rates_match_cgs_evaluation():
    rng = np.random.default_rng(7)
    n = 400
    y = rng.uniform(0.01, 1.0, (7, n))
    y /= y.sum(axis=0)
    t = rng.uniform(900.0, 2800.0, n)
    rho = rng.uniform(0.05, 3.0, n)
    got = tc.rates_of_progress(rho, t, y)
    ref = _cgs_rates(rho, t, y)
    for r in range(3):
        scale = np.abs(ref[r]).max()
        assert np.abs(got[r] - ref[r]).max() <= 1e-11 * scale, r


def test_species_sources_match_cgs_evaluation():
    rng = np.random.default_rng(8)
    y = rng.uniform(0.0, 0.3, (7, 50))
    y[6] = 1 - y[:6].sum(axis=0)
    y = np.abs(y) / np.abs(y).sum(axis=0)
    t = rng.uniform(1000.0, 2500.0, 50)
    rho = rng.uniform(0.1, 1.5, 50)
    r = _cgs_rates(rho, t, y)
    nu = np.zeros((7, 3))
    for s, v in (("C2H4", -1), ("O2", -1), ("CO", 2), ("H2", 2)):
        nu[IDX[s], 0] = v
    for s, v in (("CO", -1), ("O2", -0.5), ("CO2", 1)):
        nu[IDX[s], 1] = v
    for s, v in (("H2", -1), ("O2", -0.5), ("H2O", 1)):
        nu[IDX[s], 2] = v
    ref = W_G[:, None] * (nu @ r)                 # kg / (m^3 s)
    got = tc.species_source_terms(rho, t, y)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
