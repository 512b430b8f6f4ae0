"""Oracle discretization: element matrices, affine geometry, face maps.

Restates what grudge/meshmode/modepy provide to the reference's operators
(mirgecom/operators.py:34-95 calls ``op.weak_local_grad``,
``op.weak_local_div``, ``op.face_mass``, ``op.inverse_mass``), built
independently of the product's code:

* Lagrange basis through the monomial basis in ``y = (1 + x)/2``;
* mass and face-mass matrices by collapsed-coordinate (Duffy) Gauss-Jacobi
  quadrature, exact to polynomial degree ``2p``;
* derivative matrices by exact differentiation of monomials;
* connectivity by matching physical face-node coordinates (with periodic
  wrap), not by combinatorics;
* outward normals from the face's vertex cross product oriented away from
  the opposite vertex.

Parity: pinned by polynomial exactness, volume/area integrals and the
coordinate-matching test of connectivity (tests/test_oracle_disc.py); the
node coordinates of modepy's warp-and-blend set and meshmode's element
ordering are inputs and remain parity unpinned against the reference.
"""

from __future__ import annotations

import itertools
import math

import numpy as np
from scipy.special import roots_jacobi

# Face numbering of the all-faces array (a layout convention shared with the
# product so that face arrays can be compared entry by entry).
FACES = {
    2: ((0, 1), (1, 2), (0, 2)),
    3: ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)),
}

REF_VERTS = {
    2: np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]]),
    3: np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0],
                 [-1.0, -1.0, 1.0]]),
}


def _exponents(dim, order):
    return [a for a in itertools.product(range(order + 1), repeat=dim)
            if sum(a) <= order]


def _mono(y, exps):
    return np.stack([np.prod(y**np.array(a)[None, :], axis=1) for a in exps],
                    axis=1)


def unit_simplex_quadrature(dim, n):
    """Collapsed-coordinate Gauss rule on ``{y >= 0, sum y <= 1}`` exact for
    polynomials of degree ``<= 2n-1``.  Returns ``(points (nq, dim), w)``."""
    if dim == 0:
        return np.zeros((1, 0)), np.array([1.0])
    if dim == 1:
        x, w = roots_jacobi(n, 0.0, 0.0)
        return ((1 + x) / 2)[:, None], w / 2
    pts = []
    wts = []
    # u-direction with weight (1-u)^(dim-1)
    xu, wu = roots_jacobi(n, dim - 1.0, 0.0)
    u = (1 + xu) / 2
    wu = wu / 2**dim
    sub_p, sub_w = unit_simplex_quadrature(dim - 1, n)
    for ui, wi in zip(u, wu):
        pts.append(np.concatenate(
            [np.full((len(sub_w), 1), ui), (1 - ui) * sub_p], axis=1))
        wts.append(wi * sub_w)
    return np.concatenate(pts), np.concatenate(wts)


class OracleElement:
    """Lagrange element on the biunit simplex for given node coordinates."""

    def __init__(self, dim, order, nodes):
        self.dim = dim
        self.order = order
        self.nodes = np.asarray(nodes, dtype=np.float64)
        self.np = self.nodes.shape[0]
        exps = _exponents(dim, order)
        assert len(exps) == self.np
        self.exps = exps
        y = (1 + self.nodes) / 2
        vm = _mono(y, exps)
        self.coef = np.linalg.inv(vm)          # l_j = sum_a coef[a, j] m_a
        # mass via quadrature on the unit simplex, scaled to the biunit one
        qp, qw = unit_simplex_quadrature(dim, order + 1)
        phi = _mono(qp, exps) @ self.coef      # (nq, Np)
        self.mass = (2.0**dim) * (phi.T * qw) @ phi
        # derivative matrices: d/dx_i = (1/2) d/dy_i
        self.diff = np.zeros((dim, self.np, self.np))
        for i in range(dim):
            dm = np.zeros((self.np, len(exps)))
            for k, a in enumerate(exps):
                if a[i] == 0:
                    continue
                b = list(a)
                b[i] -= 1
                dm[:, k] = 0.5 * a[i] * np.prod(y**np.array(b)[None, :], axis=1)
            self.diff[i] = dm @ self.coef
        # faces: node sets on each face of the simplex, by coordinates
        verts = REF_VERTS[dim]
        bary = self._barycentric(self.nodes)
        self.faces = []
        self.face_mask = []
        self.face_mass = []
        qf, wf = unit_simplex_quadrature(dim - 1, order + 1)
        for fverts in FACES[dim]:
            opp = [v for v in range(dim + 1) if v not in fverts][0]
            self.faces.append(tuple(fverts))
            mask = np.nonzero(np.abs(bary[:, opp]) < 1e-12)[0]
            self.face_mask.append(mask)
            # quadrature points on the face (affine image of unit simplex)
            vf = verts[list(fverts)]
            xq = vf[0] + qf @ (vf[1:] - vf[0])
            phiq = _mono((1 + xq) / 2, exps) @ self.coef  # (nq, Np)
            phiq = phiq[:, mask]
            # integral over a face scaled to a face of measure 2 (biunit)
            meas_unit = 1.0 / math.factorial(dim - 1)
            self.face_mass.append((phiq.T * wf) @ phiq * (2.0 / meas_unit))
        self.face_mask = np.array(self.face_mask)
        self.nfp = self.face_mask.shape[1]

    @staticmethod
    def _barycentric(x):
        lam = (1 + x) / 2
        return np.concatenate([1 - lam.sum(axis=1, keepdims=True), lam], axis=1)


class OracleDisc:
    """Affine geometry and connectivity for a simplicial mesh.

    ``vertices`` (nv, d), ``elements`` (K, d+1), ``box`` = (a, b, periodic)
    for periodic wrap; ``ref_nodes`` the element's node coordinates.
    """

    def __init__(self, vertices, elements, ref_nodes, order, box):
        vertices = np.asarray(vertices, dtype=np.float64)
        elements = np.asarray(elements)
        dim = vertices.shape[1]
        self.dim = dim
        self.el = OracleElement(dim, order, ref_nodes)
        el = self.el
        self.K = elements.shape[0]
        v = vertices[elements]
        self.v = v
        # x = v0 + sum_a (v_{a+1} - v0)(1 + r_a)/2
        jm = 0.5 * (v[:, 1:] - v[:, :1])               # [k, a, i] = dx_i/dr_a
        self.J = np.linalg.det(jm)
        assert np.all(self.J > 0)
        # dr_a/dx_i = inverse of dx_i/dr_a
        self.drdx = np.linalg.inv(np.transpose(jm, (0, 2, 1)))   # [k, a, i]
        self.x = (v[:, None, 0, :] + np.einsum(
            "kai,na->kni", v[:, 1:] - v[:, :1], (1 + el.nodes) / 2))  # (K,Np,d)
        # faces
        nf = dim + 1
        self.nf = nf
        self.normal = np.zeros((self.K, nf, dim))
        self.sJ = np.zeros((self.K, nf))
        for f, fv in enumerate(el.faces):
            opp = [i for i in range(dim + 1) if i not in fv][0]
            pv = v[:, list(fv)]
            if dim == 2:
                t = pv[:, 1] - pv[:, 0]
                nrm = np.stack([t[:, 1], -t[:, 0]], axis=1)
                area = np.linalg.norm(t, axis=1)
            else:
                nrm = np.cross(pv[:, 1] - pv[:, 0], pv[:, 2] - pv[:, 0])
                area = 0.5 * np.linalg.norm(nrm, axis=1)
            nrm = nrm / np.linalg.norm(nrm, axis=1)[:, None]
            inward = v[:, opp] - pv[:, 0]
            sgn = np.sign(np.einsum("ki,ki->k", nrm, inward))
            self.normal[:, f] = -sgn[:, None] * nrm
            self.sJ[:, f] = area / 2.0
        # connectivity by coordinates
        a, b, periodic = box
        self.period = np.array([(b[i] - a[i]) if periodic[i] else 0.0
                                for i in range(dim)])
        self._connect()

    def use_element_matrices(self, mass, diff, face_mass):
        """Replace the oracle's own element matrices by given ones (same
        node set and face numbering).  The element matrices are an input of
        grudge's operators (modepy builds them); injecting the product's
        isolates the operator assembly from matrix-construction roundoff,
        which alone moves p=3 right-hand sides by ~1e-12 relative."""
        el = self.el
        assert np.allclose(mass, el.mass, rtol=0, atol=1e-12)
        assert np.allclose(diff, el.diff, rtol=0, atol=1e-11)
        el.mass = np.array(mass, dtype=np.float64)
        el.diff = np.array(diff, dtype=np.float64)
        el.face_mass = [np.array(m, dtype=np.float64) for m in face_mass]

    def _wrap(self, x):
        y = x.copy()
        for i in range(self.dim):
            if self.period[i] > 0:
                y[..., i] = np.mod(y[..., i], self.period[i])
                # snap values next to the period back to 0
                close = np.abs(y[..., i] - self.period[i]) < 1e-9 * self.period[i]
                y[..., i][close] = 0.0
        return y

    def _connect(self):
        el = self.el
        K, nf, nfp = self.K, self.nf, el.nfp
        fx = np.stack([self.x[:, el.face_mask[f]] for f in range(nf)],
                      axis=1)                       # (K, Nf, Nfp, d)
        # a face node is identified by its face (wrapped centroid) and its
        # own wrapped coordinates: vertex/edge nodes lie on many faces
        cen = np.repeat(fx.mean(axis=2, keepdims=True), el.nfp, axis=2)
        wx = np.concatenate([self._wrap(cen), self._wrap(fx)],
                            axis=-1).reshape(-1, 2 * self.dim)
        scale = max(1.0, np.abs(wx).max())
        q = np.round(wx / scale * 2**30).astype(np.int64)
        order = np.lexsort(q.T[::-1])
        sq = q[order]
        # tolerance-robust matching: compare neighbours after sort
        nbr = np.full(q.shape[0], -1, dtype=np.int64)
        i = 0
        n = q.shape[0]
        while i < n:
            j = i + 1
            while j < n and np.all(np.abs(sq[j] - sq[i]) <= 2):
                j += 1
            grp = order[i:j]
            if len(grp) == 2:
                nbr[grp[0]] = grp[1]
                nbr[grp[1]] = grp[0]
            elif len(grp) > 2:
                raise ValueError("face node shared by more than two faces")
            i = j
        self.nbr = nbr.reshape(K, nf, nfp)          # flat face-node index
        self.is_bnd = self.nbr < 0
        # volume node index of each face node
        vol = (np.arange(K)[:, None, None] * el.np + el.face_mask[None])
        self.vmapM = vol
        flat_vol = vol.reshape(-1)
        self.vmapP = np.where(self.nbr >= 0, flat_vol[np.maximum(self.nbr, 0)],
                              -1)
        self.face_x = fx

    # {{{ grudge-semantics element-local operators

    def inverse_mass(self, u):
        """``M^-1 u`` per element (affine: reference mass times J)."""
        minv = np.linalg.inv(self.el.mass)
        return np.einsum("ij,...kj->...ki", minv, u) / self.J[:, None]

    def weak_local_grad(self, u):
        """``(d phi_j / d x_i, u)`` per element: ``(dim, ..., K, Np)``."""
        el = self.el
        # int d_a phi_j u = (D_a^T M u)_j
        dtm = np.stack([el.diff[a].T @ el.mass for a in range(self.dim)])
        ref = np.einsum("aij,...kj->a...ki", dtm, u)     # (a, ..., K, Np)
        out = np.einsum("a...ki,kax->x...ki", ref, self.drdx)
        return out * self.J[:, None]

    def weak_local_div(self, v):
        """``sum_i (d phi / d x_i, v_i)``; ``v`` has a trailing-dim axis
        first: ``(dim, ..., K, Np)``."""
        el = self.el
        dtm = np.stack([el.diff[a].T @ el.mass for a in range(self.dim)])
        # sum_i drdx[a, i] v_i
        vr = np.einsum("kai,i...kn->a...kn", self.drdx, v)
        out = np.einsum("aij,a...kj->...ki", dtm, vr)
        return out * self.J[:, None]

    def face_mass(self, flux):
        """``sum_f sJ_f E_f flux_f``: ``flux`` is ``(..., K, Nf, Nfp)``."""
        el = self.el
        out = np.zeros(flux.shape[:-2] + (el.np,))
        for f in range(self.nf):
            contrib = np.einsum("mn,...kn->...km", el.face_mass[f],
                                flux[..., f, :]) * self.sJ[:, f][:, None]
            out[..., el.face_mask[f]] += contrib
        return out

    # }}}

    def trace(self, u):
        """Interior trace at face nodes: ``(..., K, Np)`` -> ``(..., K, Nf,
        Nfp)``."""
        flat = u.reshape(u.shape[:-2] + (-1,))
        return flat[..., self.vmapM]

    def exterior(self, u):
        """Neighbour trace (interior faces); boundary faces get NaN."""
        flat = u.reshape(u.shape[:-2] + (-1,))
        out = flat[..., np.maximum(self.vmapP, 0)]
        return np.where(self.is_bnd, np.nan, out)
