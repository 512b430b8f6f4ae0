"""Batched ensemble (mg_ensemble_prepare / mg_ensemble_ns_rhs) against the
same cases evaluated one by one through the single-case entry points, and
through the Python operators (make_fluid_state + ns_operator(precision=
"mixed") + the exact species sources): bit-equal right-hand sides for every
launch shape, bit-equal graph-replayed RK4 steps, and each case's mixed RHS
within the mixed-precision tolerance of its FP64 RHS (the FP64 path is the
one checked against the oracle)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

SPECS = None


def _specs():
    from mirgecom_amd.ensemble import CaseSpec
    return [CaseSpec(3, 1, 800.0, 1), CaseSpec(4, 2, 2500.0, 2),
            CaseSpec(3, 3, 4000.0, 3), CaseSpec(3, 4, 1200.0, 4),
            CaseSpec(5, 2, 600.0, 5)]


@pytest.fixture(scope="module")
def ens():
    from mirgecom_amd.ensemble import ReactingEnsemble
    print("ensemble of 5 cases (p = 1..4), exact chemistry", flush=True)
    return ReactingEnsemble(_specs())


TUNINGS = [{"fused": 1, "persistent": 0, "wg_per_cu": [4] * 5},
           {"fused": 0, "persistent": 0, "wg_per_cu": [4] * 5},
           {"fused": 1, "persistent": 1, "wg_per_cu": [2, 2, 2, 2, 2]},
           {"fused": 0, "persistent": 1, "wg_per_cu": [1, 8, 4, 2, 1]}]


@pytest.mark.parametrize("tuning", range(len(TUNINGS)))
def test_batched_equals_sequential(ens, tuning):
    print("launch shape", TUNINGS[tuning], flush=True)
    t0 = ens.temperature.clone()
    ens.prepare(TUNINGS[tuning])
    ens.rhs_sequential("y")
    torch.cuda.synchronize()
    seq_k, seq_t = ens.k.clone(), ens.temperature.clone()
    ens.temperature.copy_(t0)
    ens.k.fill_(float("nan"))
    ens.rhs("y")
    torch.cuda.synchronize()
    assert torch.isfinite(ens.k).all()
    assert torch.equal(ens.k, seq_k)
    assert torch.equal(ens.temperature, seq_t)
    ens.temperature.copy_(t0)


def test_case_equals_python_operators(ens):
    """Case by case: make_fluid_state (Newton from the seed) + ns_operator
    (mixed) + get_species_source_terms equals the ensemble's RHS bit for
    bit; and the mixed RHS is within 1e-3 (per field, relative) of the FP64
    one."""
    import warnings

    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator
    t0 = ens.temperature.clone()
    ens.prepare(TUNINGS[0])
    ens.rhs("y")
    torch.cuda.synchronize()
    for i in range(len(ens.specs)):
        gm = ens.gas_models[i]
        cv = ConservedVars(3, ens.case_view(ens.y, i).clone())
        seed = t0[int(ens.node_off[i]):int(ens.node_off[i + 1])].view_as(ens.case_t(i))
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            st = make_fluid_state(cv, gm, temperature_seed=seed.clone())
            r = ns_operator(ens.dcolls[i], gm, st, {}, precision="mixed")
            r64 = ns_operator(ens.dcolls[i], gm, st, {})
        r = gm.eos.add_species_source_terms(cv, st.temperature, r)
        r64 = gm.eos.add_species_source_terms(cv, st.temperature, r64)
        got = ens.case_view(ens.k, i)
        assert torch.equal(got, r.q), i
        a, b = got.reshape(12, -1), r64.q.reshape(12, -1)
        scale = b.abs().max(1).values
        # fields whose RHS vanishes (species absent everywhere): absolute
        scale = torch.where(scale > 0, scale, torch.ones_like(scale))
        err = ((a - b).abs().max(1).values / scale).cpu().numpy()
        print(f"case {i} p={ens.specs[i].order}: mixed vs FP64 rel err {err.max():.2e}",
              flush=True)
        assert err.max() < 1e-3
    ens.temperature.copy_(t0)


def test_graph_step_equals_eager(ens):
    y0, t0 = ens.y.clone(), ens.temperature.clone()
    ens.prepare(TUNINGS[0])
    ens.step_eager(sequential=True)
    ens.step_eager(sequential=True)
    torch.cuda.synchronize()
    want_y, want_t = ens.y.clone(), ens.temperature.clone()
    ens.y.copy_(y0)
    ens.temperature.copy_(t0)
    ens.capture()
    for rep in range(2):
        ens.y.copy_(y0)
        ens.temperature.copy_(t0)
        ens.step()
        torch.cuda.synchronize()
        _ = (ens.y * 2).sum()      # unrelated eager work between replays
        ens.step()
        torch.cuda.synchronize()
        assert torch.equal(ens.y, want_y), rep
        assert torch.equal(ens.temperature, want_t), rep
    ens.y.copy_(y0)
    ens.temperature.copy_(t0)


def test_autotune_picks_a_measured_shape(ens):
    from mirgecom_amd.ensemble import autotune
    table = autotune(ens, reps=2)
    ms = [r["ms"] for r in table["trials"]]
    assert table["best_ms"] == min(ms)
    assert len(ms) >= 6


def test_surrogate_ensemble_equals_sequential():
    """With the surrogate (untrained weights, gate bounds open so that the
    network path runs) the batched and the sequential sources agree bit for
    bit as well."""
    from mirgecom_amd.ensemble import ReactingEnsemble
    from mirgecom_amd.surrogate import ChemistrySurrogate
    sg = ChemistrySurrogate(seed=3)
    nm = sg.norm
    for i in range(10):
        nm.in_mean[i], nm.in_scale[i] = 0.0, 1.0
        nm.lo[i], nm.hi[i] = -1e300, 1e300
    nm.in_mean[0], nm.in_scale[0] = 7.3, 0.2
    nm.in_mean[1], nm.in_scale[1] = 11.5, 0.1
    nm.in_mean[9], nm.in_scale[9] = 1e-8, 1e-8
    for k in range(7):
        nm.out_mean[k], nm.out_scale[k] = 0.0, 1e-9
    nm.check_tol, nm.sample_period, nm.sample_offset = 1e300, 100, 3
    e = ReactingEnsemble(_specs()[:3], surrogate=sg)
    t0 = e.temperature.clone()
    e.rhs_sequential("y")
    torch.cuda.synchronize()
    want = e.k.clone()
    e.temperature.copy_(t0)
    e.rhs("y")
    torch.cuda.synchronize()
    assert torch.equal(e.k, want)
