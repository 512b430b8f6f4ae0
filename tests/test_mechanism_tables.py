"""The transcribed uiuc_7sp tables (mirgecom_amd/mechanisms.py) against the
reference's mechanism file mirgecom/mechanisms/uiuc_7sp.yaml (read as data
with yaml.safe_load; skipped where the reference tree is not mounted)."""

import os

import numpy as np
import pytest

from mirgecom_amd import mechanisms as mech

YAML = "/root/reference/mirgecom/mechanisms/uiuc_7sp.yaml"


@pytest.mark.skipif(not os.path.exists(YAML), reason="reference not mounted")
def test_tables_match_yaml():
    import yaml
    with open(YAML) as f:
        doc = yaml.safe_load(f)
    phase = doc["phases"][0]
    assert tuple(phase["species"]) == mech.SPECIES
    assert doc["units"] == {"length": "cm", "quantity": "mol",
                            "activation-energy": "cal/mol"}
    by_name = {s["name"]: s for s in doc["species"]}
    lo, hi = mech.nasa7_tables()
    for k, name in enumerate(mech.SPECIES):
        sp = by_name[name]
        assert sp["composition"] == mech._COMPOSITION[name]
        th = sp["thermo"]
        assert th["model"] == "NASA7"
        assert th["temperature-ranges"][1] == mech.T_MID
        assert np.array_equal(np.array(th["data"][0]), lo[k])
        assert np.array_equal(np.array(th["data"][1]), hi[k])
        tr = sp["transport"]
        geo, sig, eps, dip, pol, zrot = mech._TRANSPORT[name]
        assert tr["model"] == "gas" and tr["geometry"] == geo
        assert (tr["diameter"], tr["well-depth"]) == (sig, eps)
        assert tr.get("dipole", 0.0) == dip
        assert tr.get("polarizability", 0.0) == pol
        assert tr.get("rotational-relaxation", 0.0) == zrot
        assert tuple(th["temperature-ranges"][::2]) == mech._T_RANGE[name]
    assert phase["transport"] == "mixture-averaged"
    rxns = doc["reactions"]
    assert len(rxns) == 3
    for r, (a, b, ea, order) in zip(rxns, mech._REACTIONS):
        rc = r["rate-constant"]
        assert (rc["A"], rc["b"], rc["Ea"]) == (a, b, ea)
        if "orders" in r:
            assert sum(r["orders"].values()) == pytest.approx(order)
    assert "<=>" in rxns[1]["equation"]
    assert "=>" in rxns[0]["equation"] and "<=>" not in rxns[0]["equation"]


def test_molecular_weights():
    w = mech.molecular_weights()
    assert w[mech.SPECIES.index("O2")] == pytest.approx(31.998)
    assert w[mech.SPECIES.index("C2H4")] == pytest.approx(28.054)
    assert w[mech.SPECIES.index("N2")] == pytest.approx(28.014)


def test_arrhenius_units():
    log_a, b, ea_r = mech.arrhenius_si()
    # R1 order 1.15: A_SI = 1e10 * (1e-3)^0.15
    assert np.exp(log_a[0]) == pytest.approx(1e10 * 1e-3**0.15, rel=1e-14)
    assert np.exp(log_a[1]) == pytest.approx(1.1e7 * 1e-3**0.5, rel=1e-14)
    assert ea_r[0] == pytest.approx(36000 * 4.184 * 1000 / 8314.46261815324)
