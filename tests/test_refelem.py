"""Reference element: the product's operators against the oracle's
independently built ones, and against what the mathematics fixes.

The reference gets these from modepy (warp-and-blend nodes, orthonormal
simplex basis) inside grudge; modepy is not importable here, so node
coordinates and ordering are parity unpinned against it, while the operators
on a given node set are pinned by polynomial exactness, exact integrals and
integration by parts."""

import itertools
import math

import numpy as np
import pytest

from mirgecom_amd.refelem import (
    FACE_VERTICES, REF_VERTICES, get_reference_element, legendre_gauss_lobatto)
from oracle.disc import OracleElement

CASES = [(d, p) for d in (2, 3) for p in (1, 2, 3, 4)]


@pytest.mark.parametrize("dim,order", CASES)
def test_matrices_match_oracle(dim, order):
    re = get_reference_element(dim, order)
    oe = OracleElement(dim, order, re.nodes)
    assert np.abs(re.mass - oe.mass).max() < 1e-13
    assert np.abs(re.diff - oe.diff).max() < 1e-12
    for f in range(dim + 1):
        assert (re.face_mask[f] == oe.face_mask[f]).all()
        assert np.abs(re.face_mass[f] - oe.face_mass[f]).max() < 1e-13


@pytest.mark.parametrize("dim,order", CASES)
def test_derivative_exact_on_polynomials(dim, order):
    re = get_reference_element(dim, order)
    x = re.nodes
    for exps in itertools.product(range(order + 1), repeat=dim):
        if sum(exps) > order:
            continue
        u = np.prod(x**np.array(exps), axis=1)
        for a in range(dim):
            if exps[a] == 0:
                du = 0 * u
            else:
                e2 = list(exps)
                e2[a] -= 1
                du = exps[a] * np.prod(x**np.array(e2), axis=1)
            assert np.abs(re.diff[a] @ u - du).max() < 1e-11


def _simplex_monomial_integral(exps):
    """int over the unit simplex of prod y_i^a_i = prod a_i! / (sum a + d)!"""
    d = len(exps)
    return (np.prod([math.factorial(a) for a in exps])
            / math.factorial(sum(exps) + d))


@pytest.mark.parametrize("dim,order", CASES)
def test_mass_integrates_exactly(dim, order):
    """1^T M u = int u for polynomials of degree <= p, and
    u^T M v = int u v for degree <= 2p, via the closed form."""
    re = get_reference_element(dim, order)
    y = (1 + re.nodes) / 2            # biunit -> unit simplex, J = 2^dim
    for exps in itertools.product(range(2 * order + 1), repeat=dim):
        if sum(exps) > 2 * order:
            continue
        e1 = [min(a, order) for a in exps]
        if sum(e1) > order:
            continue
        e2 = [a - b for a, b in zip(exps, e1)]
        if sum(e2) > order:
            continue
        u = np.prod(y**np.array(e1), axis=1)
        v = np.prod(y**np.array(e2), axis=1)
        exact = 2.0**dim * _simplex_monomial_integral(exps)
        assert abs(u @ re.mass @ v - exact) < 1e-13


@pytest.mark.parametrize("dim,order", CASES)
def test_integration_by_parts(dim, order):
    """M D_r + D_r^T M = sum_f E_f n_{f,r} |f| / 2 on the reference element:
    pins the face masses, the mass matrix and the derivative jointly."""
    re = get_reference_element(dim, order)
    verts = REF_VERTICES[dim]
    for a in range(dim):
        lhs = re.mass @ re.diff[a] + re.diff[a].T @ re.mass
        rhs = np.zeros_like(lhs)
        for f, fv in enumerate(FACE_VERTICES[dim]):
            pv = verts[list(fv)]
            opp = [v for v in range(dim + 1) if v not in fv][0]
            if dim == 2:
                t = pv[1] - pv[0]
                nrm = np.array([t[1], -t[0]])
                meas = np.linalg.norm(t)
            else:
                nrm = np.cross(pv[1] - pv[0], pv[2] - pv[0])
                meas = 0.5 * np.linalg.norm(nrm)
            nrm = nrm / np.linalg.norm(nrm)
            if nrm @ (verts[opp] - pv[0]) > 0:
                nrm = -nrm
            m = re.face_mask[f]
            rhs[np.ix_(m, m)] += re.face_mass[f] * nrm[a] * meas / 2
        assert np.abs(lhs - rhs).max() < 1e-12


@pytest.mark.parametrize("dim,order", CASES)
def test_lift_and_weak_diff_definitions(dim, order):
    re = get_reference_element(dim, order)
    nfp = re.nface_nodes
    emb = np.zeros((re.nnodes, re.nfaces * nfp))
    for f in range(re.nfaces):
        emb[re.face_mask[f], f * nfp:(f + 1) * nfp] = re.face_mass[f]
    assert np.abs(re.mass @ re.lift - emb).max() < 1e-12
    for a in range(dim):
        assert np.abs(re.mass @ re.weak_diff[a] - re.diff[a].T @ re.mass).max() < 1e-12


@pytest.mark.parametrize("dim,order", CASES)
def test_nodes_warp_and_blend_properties(dim, order):
    """Vertices are nodes; every edge carries the 1-D GLL points; the set is
    invariant under the vertex permutations of the simplex; order <= 2 is
    equispaced."""
    re = get_reference_element(dim, order)
    x = re.nodes
    verts = REF_VERTICES[dim]
    for v in verts:
        assert np.min(np.linalg.norm(x - v, axis=1)) < 1e-14
    gll = legendre_gauss_lobatto(order)
    for i, j in itertools.combinations(range(dim + 1), 2):
        pts = verts[i] + np.outer((1 + gll) / 2, verts[j] - verts[i])
        for pt in pts:
            assert np.min(np.linalg.norm(x - pt, axis=1)) < 1e-12
    bary = np.concatenate([1 - ((1 + x) / 2).sum(axis=1, keepdims=True),
                           (1 + x) / 2], axis=1)
    key = sorted(map(tuple, np.round(bary, 10)))
    for perm in itertools.permutations(range(dim + 1)):
        pkey = sorted(map(tuple, np.round(bary[:, perm], 10)))
        assert np.allclose(key, pkey, atol=1e-9)
    if order <= 2:
        tup = re.tuples
        assert np.abs(x - (-1 + 2 * tup / order)).max() < 1e-15


def test_node_counts():
    for dim in (2, 3):
        for p in (1, 2, 3, 4):
            re = get_reference_element(dim, p)
            assert re.nnodes == math.comb(p + dim, dim)
            assert re.nface_nodes == math.comb(p + dim - 1, dim - 1)
