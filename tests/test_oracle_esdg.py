"""Pins of the oracle's entropy-stable Euler operator (oracle/esdg.py;
mirgecom/euler.py:108-310 with use_esdg): the properties that define the
scheme -- discrete entropy conservation with the entropy-conserving face
flux, entropy dissipation with the entropy-stable Rusanov flux, free-stream
preservation, global conservation -- and the pieces: entropy-variable round
trip (gas_model.py:1026-1100), the two-point flux's consistency and symmetry
(inviscid.py:451-518), the logarithmic mean's series branch."""

import numpy as np
import pytest

from oracle import physics as ph
from oracle.esdg import (
    _HybridOps, chandrashekar, conserved_vars, entropy_vars,
    esdg_euler_operator, ln_mean)
from tests.cases import Problem, smooth_perturbed_state

GAMMA = 1.4


def test_entropy_vars_round_trip():
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (3, 50))
    q = smooth_perturbed_state(x, seed=1, nspecies=2, amp=0.2)
    back = conserved_vars(entropy_vars(q, GAMMA, 3), GAMMA, 3)
    np.testing.assert_allclose(back, q, rtol=1e-13, atol=1e-14)


def test_two_point_flux():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (2, 40))
    ql = smooth_perturbed_state(x, seed=2, nspecies=1, amp=0.2)
    qr = smooth_perturbed_state(x, seed=3, nspecies=1, amp=0.2)
    # consistency: F_S(q, q) = F(q)
    gm = ph.GasModel(ph.IdealSingleGas(gamma=GAMMA))
    f = ph.inviscid_flux(ph.make_fluid_state(ql, gm, 2))
    np.testing.assert_allclose(chandrashekar(ql, ql, GAMMA, 2), f, rtol=1e-12,
                               atol=1e-14)
    # symmetry
    np.testing.assert_allclose(chandrashekar(ql, qr, GAMMA, 2),
                               chandrashekar(qr, ql, GAMMA, 2), rtol=1e-13)
    # log mean: series branch near x = y, exact branch elsewhere
    a = np.array([1.0, 1.0, 2.0])
    b = np.array([1.0 + 1e-6, 1.3, 5.0])
    want = np.where(np.abs(b - a) > 1e-3, (b - a) / np.log(b / a), (a + b) / 2)
    np.testing.assert_allclose(ln_mean(a, b), want, rtol=1e-10)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2)])
def test_sbp_property(dim, order):
    """Q_r + Q_r^T = E^T B_r E on the reference element (the summation-by-
    parts property the hybridized operator relies on)."""
    pb = Problem(dim, order, (3,) * dim, 0.0, 1.0, (True,) * dim, own_matrices=True)
    el = pb.odisc.el
    ho = _HybridOps(el, 2 * order)
    from oracle.disc import FACES, REF_VERTS
    # reference outward normals times face measure ratio (biunit)
    for r in range(dim):
        b = []
        for f, fv in enumerate(FACES[dim]):
            v = REF_VERTS[dim][list(fv)]
            opp = [i for i in range(dim + 1) if i not in fv][0]
            # normal of face f in reference coordinates, scaled by the ratio of
            # the face's measure to the biunit face measure
            if dim == 2:
                t = v[1] - v[0]
                n = np.array([t[1], -t[0]])
            else:
                n = np.cross(v[1] - v[0], v[2] - v[0])
            if n @ (REF_VERTS[dim][opp] - v[0]) > 0:
                n = -n
            area = np.linalg.norm(n) / (1.0 if dim == 2 else 2.0)
            n = n / np.linalg.norm(n)
            biunit = 2.0
            b.append(n[r] * (area / biunit) * ho.wf)
        bdiag = np.concatenate(b)
        lhs = ho.qr[r] + ho.qr[r].T
        rhs = ho.e.T @ (bdiag[:, None] * ho.e)
        assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(rhs).max()


def _entropy_rate(pb, q, rhs):
    from oracle.esdg import _HybridOps
    el = pb.odisc.el
    ho = _HybridOps(el, 2 * el.order)
    uq = np.einsum("qn,fkn->fkq", ho.vq, q)
    ev = np.einsum("nq,fkq->fkn", ho.pq, entropy_vars(uq, GAMMA, pb.dim))
    mass = np.linalg.inv(ho.minv)
    mr = np.einsum("ij,fkj->fki", mass, rhs) * pb.odisc.J[None, :, None]
    return np.sum(ev * mr), np.sum(np.abs(ev) * np.abs(mr))


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2)])
def test_entropy_conservation_and_dissipation(dim, order):
    n = (4, 3) if dim == 2 else (3, 3, 3)
    pb = Problem(dim, order, n, 0.0, (1.0, 0.8, 0.9)[:dim], (True,) * dim,
                 own_matrices=True)
    q = smooth_perturbed_state(pb.x, seed=4, amp=0.15)
    rec = esdg_euler_operator(pb.odisc, pb.box, q, GAMMA, {}, dissipation=False)
    rate, scale = _entropy_rate(pb, q, rec)
    assert abs(rate) < 1e-12 * scale
    res = esdg_euler_operator(pb.odisc, pb.box, q, GAMMA, {})
    rate_s, scale_s = _entropy_rate(pb, q, res)
    assert rate_s < 1e-12 * scale_s
    assert rate_s < -1e-8 * scale_s            # strictly dissipative here
    # global conservation of mass, momentum and energy
    mass = np.linalg.inv(_HybridOps(pb.odisc.el, 2 * order).minv)
    tot = np.einsum("ij,fkj->f", mass, res * pb.odisc.J[None, :, None])
    assert np.abs(tot).max() < 1e-12 * np.abs(res).max()


@pytest.mark.parametrize("dim,order", [(2, 2), (3, 2)])
def test_free_stream(dim, order):
    n = (3, 3) if dim == 2 else (3, 3, 3)
    pb = Problem(dim, order, n, 0.0, 1.0, (True,) * dim, own_matrices=True)
    q = np.zeros((2 + dim,) + pb.x.shape[1:])
    q[0], q[1], q[2] = 1.2, 2.5, 0.3
    r = esdg_euler_operator(pb.odisc, pb.box, q, GAMMA, {})
    assert np.abs(r).max() < 1e-11
