"""The torch entropy-conserving / entropy-stable flux functions of the
call surface (mirgecom/inviscid.py:451-711) against the oracle's
restatement (oracle/esdg.py), on CPU tensors."""

import numpy as np
import torch

from tests.cases import smooth_perturbed_state


class _Pair:
    def __init__(self, a, b):
        self.int, self.ext = a, b


def _states(dim, nsp):
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (dim, 30))
    ql = smooth_perturbed_state(x, seed=11, nspecies=nsp, amp=0.2)
    qr = smooth_perturbed_state(x, seed=12, nspecies=nsp, amp=0.2)
    gm = GasModel(eos=IdealSingleGas(gamma=1.4, gas_const=1.0))
    sl = make_fluid_state(ConservedVars(dim, torch.as_tensor(ql)), gm)
    sr = make_fluid_state(ConservedVars(dim, torch.as_tensor(qr)), gm)
    return gm, ql, qr, sl, sr


def test_chandrashekar_matches_oracle():
    from mirgecom_amd.inviscid import entropy_conserving_flux_chandrashekar
    from oracle.esdg import chandrashekar
    for dim, nsp in ((2, 0), (3, 2)):
        gm, ql, qr, sl, sr = _states(dim, nsp)
        got = entropy_conserving_flux_chandrashekar(gm, sl, sr).q.numpy()
        np.testing.assert_allclose(got, chandrashekar(ql, qr, 1.4, dim),
                                   rtol=1e-13, atol=1e-14)


def test_entropy_stable_facial_fluxes():
    from mirgecom_amd.inviscid import (
        entropy_stable_inviscid_facial_flux,
        entropy_stable_inviscid_facial_flux_rusanov)
    from oracle.esdg import es_facial_flux
    dim = 3
    gm, ql, qr, sl, sr = _states(dim, 1)
    rng = np.random.default_rng(8)
    n = rng.standard_normal((dim, 30))
    n /= np.linalg.norm(n, axis=0)
    nt = torch.as_tensor(n)
    pair = _Pair(sl, sr)
    np.testing.assert_allclose(
        entropy_stable_inviscid_facial_flux(pair, gm, nt).q.numpy(),
        es_facial_flux(ql, qr, n, 1.4, dim, dissipation=False),
        rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(
        entropy_stable_inviscid_facial_flux_rusanov(pair, gm, nt).q.numpy(),
        es_facial_flux(ql, qr, n, 1.4, dim, dissipation=True),
        rtol=1e-13, atol=1e-14)


def _mix_states():
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import FluidState, GasModel, _LazyDV
    from oracle import physics as ph, thermochem as tc
    from tests.cases import mixture_state
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 2 * np.pi, (3, 30))
    ql, tl = mixture_state(x, seed=4)
    qr, tr = mixture_state(x, seed=5)
    gm = GasModel(eos=PyrometheusMixture())
    ogm = ph.GasModel(tc.PyrometheusMixture())
    osl = ph.make_fluid_state(ql, ogm, 3, temperature_seed=tl + 1.0)
    osr = ph.make_fluid_state(qr, ogm, 3, temperature_seed=tr + 1.0)
    # the host states take the oracle's temperatures (the device Newton
    # solve is not available on the CPU)

    def host(q, t):
        cv = ConservedVars(3, torch.as_tensor(q))
        return FluidState(cv, _LazyDV(cv, gm.eos, torch.as_tensor(t)), gm)

    return gm, ogm, host(ql, osl.temperature), host(qr, osr.temperature), osl, osr


def test_renac_matches_oracle():
    from mirgecom_amd.inviscid import entropy_conserving_flux_renac
    from oracle.esdg import renac
    gm, ogm, sl, sr, osl, osr = _mix_states()
    got = entropy_conserving_flux_renac(gm, sl, sr).q.numpy()
    want = renac(osl, osr, ogm.eos, 3)
    np.testing.assert_allclose(got, want, rtol=1e-12,
                               atol=1e-13 * np.abs(want).max())


def test_mixture_facial_flux_defaults_to_renac():
    from mirgecom_amd.inviscid import entropy_stable_inviscid_facial_flux_rusanov
    from oracle.esdg import renac
    gm, ogm, sl, sr, osl, osr = _mix_states()
    rng = np.random.default_rng(10)
    n = rng.standard_normal((3, 30))
    n /= np.linalg.norm(n, axis=0)
    got = entropy_stable_inviscid_facial_flux_rusanov(_Pair(sl, sr), gm,
                                                      torch.as_tensor(n)).q.numpy()
    alpha = np.maximum(osl.wavespeed, osr.wavespeed)
    want = (np.einsum("fd...,d...->f...", renac(osl, osr, ogm.eos, 3), n)
            - 0.5 * alpha * (osr.cv - osl.cv))
    np.testing.assert_allclose(got, want, rtol=1e-11,
                               atol=1e-12 * np.abs(want).max())


def test_entropy_variable_conversions():
    """gas_model.conservative_to_entropy_vars / entropy_to_conservative_vars
    (mirgecom/gas_model.py:1026-1100) against the oracle, ideal gas and
    mixture (the state's own pressure)."""
    from mirgecom_amd.gas_model import (
        conservative_to_entropy_vars, entropy_to_conservative_vars)
    from oracle.esdg import conserved_vars, entropy_vars_state
    gm, ql, qr, sl, sr = _states(3, 2)
    ev = conservative_to_entropy_vars(1.4, sl)
    np.testing.assert_allclose(ev.q.numpy(),
                               entropy_vars_state(ql, sl.pressure.numpy(), 1.4, 3),
                               rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(entropy_to_conservative_vars(1.4, ev).q.numpy(), ql,
                               rtol=1e-12, atol=1e-14)
    gm, ogm, sl, sr, osl, osr = _mix_states()
    g = gm.eos.gamma(sl.cv, sl.temperature)
    ev = conservative_to_entropy_vars(g, sl)
    og = ogm.eos.gamma(osl.cv, osl.temperature)
    want = entropy_vars_state(osl.cv, osl.pressure, og, 3)
    from tests.cases import rel_err
    assert rel_err(ev.q.numpy(), want).max() < 1e-11
    back = entropy_to_conservative_vars(g, ev).q.numpy()
    assert rel_err(back, conserved_vars(want, og, 3)).max() < 1e-11
