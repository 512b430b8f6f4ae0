"""Reference simplex elements: warp-and-blend nodes and the element-local
operator matrices of a nodal DG discretization.

This is the host-side (one-time, FP64) setup that the reference obtains from
modepy + meshmode through ``create_discretization_collection``
(mirgecom/discretization.py:41-83), whose base group factory is
``InterpolatoryEdgeClusteredGroupFactory(order)``: warp-and-blend
("edge-clustered") nodes on the biunit simplex.

The node construction follows the published method of Hesthaven & Warburton,
*Nodal Discontinuous Galerkin Methods* (2008), Sec. 6.1 (triangles) and
Sec. 10.1 (tetrahedra), including their optimised blend parameters.  The
polynomial basis is the orthonormal Koornwinder/Dubiner basis of the same
book (Sec. 6.2, 10.1), used to form the Vandermonde matrix, the mass matrix
``M = (V V^T)^-1`` and the differentiation matrices ``D_r = V_r V^-1``.

Reference simplex (biunit): vertices
  2D: v0=(-1,-1), v1=(1,-1), v2=(-1,1)
  3D: v0=(-1,-1,-1), v1=(1,-1,-1), v2=(-1,1,-1), v3=(-1,-1,1)
Face f is listed by its vertex tuple in ``FACE_VERTICES[dim][f]``.

Node order: lattice tuples (i, j[, k]) with the first index varying fastest,
i.e. ``r`` fastest, then ``s``, then ``t`` (the order modepy's
``node_tuples_for_space`` yields).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

REF_VERTICES = {
    1: np.array([[-1.0], [1.0]]),
    2: np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]]),
    3: np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0],
                 [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]]),
}

# faces as tuples of reference-vertex indices
FACE_VERTICES = {
    2: ((0, 1), (1, 2), (0, 2)),
    3: ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)),
}

# Hesthaven & Warburton optimised blend parameters (index = order)
_ALPHA_OPT_2D = [0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999,
                 1.2832, 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223,
                 1.6258]
_ALPHA_OPT_3D = [0.0, 0.0, 0.0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                 1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655]


# {{{ 1D orthonormal Jacobi polynomials

def jacobi_p(x, alpha, beta, n):
    """Orthonormal Jacobi polynomial P_n^{(alpha,beta)} on [-1, 1]."""
    x = np.asarray(x, dtype=np.float64)
    pl = np.zeros((n + 1,) + x.shape)
    gamma0 = (2.0**(alpha + beta + 1) / (alpha + beta + 1)
              * math.gamma(alpha + 1) * math.gamma(beta + 1)
              / math.gamma(alpha + beta + 1))
    pl[0] = 1.0 / math.sqrt(gamma0)
    if n == 0:
        return pl[0]
    gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0
    pl[1] = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(gamma1)
    if n == 1:
        return pl[1]
    aold = 2 / (2 + alpha + beta) * math.sqrt(
        (alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, n):
        h1 = 2 * i + alpha + beta
        anew = 2 / (h1 + 2) * math.sqrt(
            (i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta)
            / (h1 + 1) / (h1 + 3))
        bnew = -(alpha**2 - beta**2) / h1 / (h1 + 2)
        pl[i + 1] = 1 / anew * (-aold * pl[i - 1] + (x - bnew) * pl[i])
        aold = anew
    return pl[n]


def grad_jacobi_p(x, alpha, beta, n):
    """Derivative of the orthonormal Jacobi polynomial."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.zeros_like(x)
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(
        x, alpha + 1, beta + 1, n - 1)


def legendre_gauss_lobatto(n):
    """The n+1 Gauss-Lobatto-Legendre points on [-1, 1], ascending."""
    if n == 1:
        return np.array([-1.0, 1.0])
    inner = np.polynomial.legendre.Legendre.basis(n).deriv().roots()
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])

# }}}


# {{{ warp & blend nodes

def _warp_factor(n, rout):
    """1D warp from equispaced to GLL nodes, blended (H&W Warpfactor)."""
    lgl = legendre_gauss_lobatto(n)
    req = np.linspace(-1, 1, n + 1)
    veq = np.stack([jacobi_p(req, 0, 0, i) for i in range(n + 1)], axis=1)
    pmat = np.stack([jacobi_p(rout, 0, 0, i) for i in range(n + 1)], axis=0)
    lmat = np.linalg.solve(veq.T, pmat)
    warp = lmat.T @ (lgl - req)
    zerof = (np.abs(rout) < 1.0 - 1.0e-10).astype(np.float64)
    sf = 1.0 - (zerof * rout)**2
    return warp / sf + warp * (zerof - 1)


def lattice_tuples(dim, order):
    """Integer lattice tuples of the order-``order`` simplex, first index
    fastest."""
    if dim == 1:
        return [(i,) for i in range(order + 1)]
    if dim == 2:
        return [(i, j) for j in range(order + 1) for i in range(order + 1 - j)]
    if dim == 3:
        return [(i, j, k) for k in range(order + 1)
                for j in range(order + 1 - k)
                for i in range(order + 1 - j - k)]
    raise ValueError(f"unsupported dimension {dim}")


def _nodes_2d(n):
    alpha = _ALPHA_OPT_2D[n] if n < 16 else 5.0 / 3.0
    tup = np.array(lattice_tuples(2, n), dtype=np.float64)
    l3 = tup[:, 0] / n      # r direction
    l1 = tup[:, 1] / n      # s direction
    l2 = 1.0 - l1 - l3
    x = -l2 + l3
    y = (-l2 - l3 + 2 * l1) / math.sqrt(3.0)
    blend1 = 4 * l2 * l3
    blend2 = 4 * l1 * l3
    blend3 = 4 * l1 * l2
    warpf1 = _warp_factor(n, l3 - l2)
    warpf2 = _warp_factor(n, l1 - l3)
    warpf3 = _warp_factor(n, l2 - l1)
    warp1 = blend1 * warpf1 * (1 + (alpha * l1)**2)
    warp2 = blend2 * warpf2 * (1 + (alpha * l2)**2)
    warp3 = blend3 * warpf3 * (1 + (alpha * l3)**2)
    x = x + warp1 + math.cos(2 * math.pi / 3) * warp2 \
        + math.cos(4 * math.pi / 3) * warp3
    y = y + math.sin(2 * math.pi / 3) * warp2 + math.sin(4 * math.pi / 3) * warp3
    # equilateral -> biunit right triangle
    bl1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    bl2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    bl3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    r = -bl2 + bl3 - bl1
    s = -bl2 - bl3 + bl1
    return np.stack([r, s], axis=1)


def _evalwarp(p, xnodes, xout):
    warp = np.zeros_like(xout)
    xeq = np.array([-1 + 2 * (p - i) / p for i in range(p + 1)])
    for i in range(p + 1):
        d = np.full_like(xout, xnodes[i] - xeq[i])
        for j in range(1, p):
            if i != j:
                d = d * (xout - xeq[j]) / (xeq[i] - xeq[j])
        if i != 0:
            d = -d / (xeq[i] - xeq[0])
        if i != p:
            d = d / (xeq[i] - xeq[p])
        warp = warp + d
    return warp


def _evalshift(p, pval, l1, l2, l3):
    gauss_x = -legendre_gauss_lobatto(p)
    blend1 = l2 * l3
    blend2 = l1 * l3
    blend3 = l1 * l2
    warpfactor1 = 4 * _evalwarp(p, gauss_x, l3 - l2)
    warpfactor2 = 4 * _evalwarp(p, gauss_x, l1 - l3)
    warpfactor3 = 4 * _evalwarp(p, gauss_x, l2 - l1)
    warp1 = blend1 * warpfactor1 * (1 + (pval * l1)**2)
    warp2 = blend2 * warpfactor2 * (1 + (pval * l2)**2)
    warp3 = blend3 * warpfactor3 * (1 + (pval * l3)**2)
    dx = warp1 + math.cos(2 * math.pi / 3) * warp2 \
        + math.cos(4 * math.pi / 3) * warp3
    dy = math.sin(2 * math.pi / 3) * warp2 + math.sin(4 * math.pi / 3) * warp3
    return dx, dy


def _nodes_3d(n):
    alpha = _ALPHA_OPT_3D[n] if n < 15 else 1.0
    tol = 1e-10
    tup = np.array(lattice_tuples(3, n), dtype=np.float64)
    r = -1 + 2 * tup[:, 0] / n
    s = -1 + 2 * tup[:, 1] / n
    t = -1 + 2 * tup[:, 2] / n
    l1 = (1 + t) / 2
    l2 = (1 + s) / 2
    l3 = -(1 + r + s + t) / 2
    l4 = (1 + r) / 2
    v1 = np.array([-1, -1 / math.sqrt(3), -1 / math.sqrt(6)])
    v2 = np.array([1, -1 / math.sqrt(3), -1 / math.sqrt(6)])
    v3 = np.array([0, 2 / math.sqrt(3), -1 / math.sqrt(6)])
    v4 = np.array([0, 0, 3 / math.sqrt(6)])
    t1 = np.array([v2 - v1, v2 - v1, v3 - v2, v3 - v1])
    t2 = np.array([v3 - 0.5 * (v1 + v2), v4 - 0.5 * (v1 + v2),
                   v4 - 0.5 * (v2 + v3), v4 - 0.5 * (v1 + v3)])
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    t2 /= np.linalg.norm(t2, axis=1)[:, None]
    xyz = np.outer(l3, v1) + np.outer(l4, v2) + np.outer(l2, v3) \
        + np.outer(l1, v4)
    shift = np.zeros_like(xyz)
    for face in range(4):
        if face == 0:
            la, lb, lc, ld = l1, l2, l3, l4
        elif face == 1:
            la, lb, lc, ld = l2, l1, l3, l4
        elif face == 2:
            la, lb, lc, ld = l3, l1, l4, l2
        else:
            la, lb, lc, ld = l4, l1, l3, l2
        warp1, warp2 = _evalshift(n, alpha, lb, lc, ld)
        blend = lb * lc * ld
        denom = (lb + 0.5 * la) * (lc + 0.5 * la) * (ld + 0.5 * la)
        ids = denom > tol
        blend[ids] = (1 + (alpha * la[ids])**2) * blend[ids] / denom[ids]
        shift = shift + np.outer(blend * warp1, t1[face]) \
            + np.outer(blend * warp2, t2[face])
        ids = (la < tol) & (((lb > tol).astype(int) + (lc > tol).astype(int)
                             + (ld > tol).astype(int)) < 3)
        shift[ids] = np.outer(warp1[ids], t1[face]) \
            + np.outer(warp2[ids], t2[face])
    xyz = xyz + shift
    # equilateral -> biunit
    rhs = xyz.T - 0.5 * (v2 + v3 + v4 - v1)[:, None]
    amat = np.stack([0.5 * (v2 - v1), 0.5 * (v3 - v1), 0.5 * (v4 - v1)], axis=1)
    rst = np.linalg.solve(amat, rhs)
    return rst.T.copy()


@lru_cache(maxsize=None)
def warp_and_blend_nodes(dim, order):
    """Warp-and-blend nodes on the biunit simplex, shape ``(Np, dim)``."""
    if order == 0:
        return np.zeros((1, dim)) - 1.0 / (dim + 1)
    if dim == 1:
        return legendre_gauss_lobatto(order)[:, None]
    if order <= 2 and dim in (2, 3):
        # warp & blend is the identity map for order <= 2
        tup = np.array(lattice_tuples(dim, order), dtype=np.float64)
        return -1 + 2 * tup / order
    if dim == 2:
        return _nodes_2d(order)
    if dim == 3:
        return _nodes_3d(order)
    raise ValueError(f"unsupported dimension {dim}")

# }}}


# {{{ orthonormal simplex basis

def _rs_to_ab(r, s):
    a = np.where(np.abs(1 - s) > 1e-14, 2 * (1 + r) / np.where(
        np.abs(1 - s) > 1e-14, 1 - s, 1.0) - 1, -1.0)
    return a, s


def _rst_to_abc(r, s, t):
    den_a = -s - t
    a = np.where(np.abs(den_a) > 1e-14,
                 2 * (1 + r) / np.where(np.abs(den_a) > 1e-14, den_a, 1.0) - 1,
                 -1.0)
    b = np.where(np.abs(1 - t) > 1e-14,
                 2 * (1 + s) / np.where(np.abs(1 - t) > 1e-14, 1 - t, 1.0) - 1,
                 -1.0)
    return a, b, t


def basis_indices(dim, order):
    """Mode index tuples (total degree <= order), in lattice order."""
    return lattice_tuples(dim, order)


def simplex_basis(dim, order, x):
    """Orthonormal basis values, shape ``(npoints, nmodes)``; ``x`` is
    ``(npoints, dim)`` on the biunit simplex."""
    x = np.asarray(x, dtype=np.float64)
    idx = basis_indices(dim, order)
    out = np.zeros((x.shape[0], len(idx)))
    if dim == 1:
        for m, (i,) in enumerate(idx):
            out[:, m] = jacobi_p(x[:, 0], 0, 0, i)
    elif dim == 2:
        a, b = _rs_to_ab(x[:, 0], x[:, 1])
        for m, (i, j) in enumerate(idx):
            out[:, m] = (math.sqrt(2.0) * jacobi_p(a, 0, 0, i)
                         * jacobi_p(b, 2 * i + 1, 0, j) * (1 - b)**i)
    elif dim == 3:
        a, b, c = _rst_to_abc(x[:, 0], x[:, 1], x[:, 2])
        for m, (i, j, k) in enumerate(idx):
            out[:, m] = (2 * math.sqrt(2.0) * jacobi_p(a, 0, 0, i)
                         * jacobi_p(b, 2 * i + 1, 0, j) * (1 - b)**i
                         * jacobi_p(c, 2 * (i + j) + 2, 0, k)
                         * (1 - c)**(i + j))
    return out


def simplex_basis_grad(dim, order, x):
    """Gradients of the orthonormal basis, shape ``(dim, npoints, nmodes)``."""
    x = np.asarray(x, dtype=np.float64)
    idx = basis_indices(dim, order)
    out = np.zeros((dim, x.shape[0], len(idx)))
    if dim == 1:
        for m, (i,) in enumerate(idx):
            out[0, :, m] = grad_jacobi_p(x[:, 0], 0, 0, i)
    elif dim == 2:
        a, b = _rs_to_ab(x[:, 0], x[:, 1])
        for m, (i, j) in enumerate(idx):
            fa = jacobi_p(a, 0, 0, i)
            dfa = grad_jacobi_p(a, 0, 0, i)
            gb = jacobi_p(b, 2 * i + 1, 0, j)
            dgb = grad_jacobi_p(b, 2 * i + 1, 0, j)
            ddr = dfa * gb
            if i > 0:
                ddr = ddr * (0.5 * (1 - b))**(i - 1)
            dds = dfa * (gb * (0.5 * (1 + a)))
            if i > 0:
                dds = dds * (0.5 * (1 - b))**(i - 1)
            tmp = dgb * (0.5 * (1 - b))**i
            if i > 0:
                tmp = tmp - 0.5 * i * gb * (0.5 * (1 - b))**(i - 1)
            dds = dds + fa * tmp
            out[0, :, m] = 2**(i + 0.5) * ddr
            out[1, :, m] = 2**(i + 0.5) * dds
    elif dim == 3:
        a, b, c = _rst_to_abc(x[:, 0], x[:, 1], x[:, 2])
        for m, (i, j, k) in enumerate(idx):
            fa = jacobi_p(a, 0, 0, i)
            dfa = grad_jacobi_p(a, 0, 0, i)
            gb = jacobi_p(b, 2 * i + 1, 0, j)
            dgb = grad_jacobi_p(b, 2 * i + 1, 0, j)
            hc = jacobi_p(c, 2 * (i + j) + 2, 0, k)
            dhc = grad_jacobi_p(c, 2 * (i + j) + 2, 0, k)
            vr = dfa * (gb * hc)
            if i > 0:
                vr = vr * (0.5 * (1 - b))**(i - 1)
            if i + j > 0:
                vr = vr * (0.5 * (1 - c))**(i + j - 1)
            vs = 0.5 * (1 + a) * vr
            tmp = dgb * (0.5 * (1 - b))**i
            if i > 0:
                tmp = tmp + (-0.5 * i) * (gb * (0.5 * (1 - b))**(i - 1))
            if i + j > 0:
                tmp = tmp * (0.5 * (1 - c))**(i + j - 1)
            tmp = fa * (tmp * hc)
            vs = vs + tmp
            vt = 0.5 * (1 + a) * vr + 0.5 * (1 + b) * tmp
            tmp = dhc * (0.5 * (1 - c))**(i + j)
            if i + j > 0:
                tmp = tmp - 0.5 * (i + j) * (hc * (0.5 * (1 - c))**(i + j - 1))
            tmp = fa * (gb * tmp)
            tmp = tmp * (0.5 * (1 - b))**i
            vt = vt + tmp
            scale = 2**(2 * i + j + 1.5)
            out[0, :, m] = vr * scale
            out[1, :, m] = vs * scale
            out[2, :, m] = vt * scale
    return out

# }}}


def face_local_coords(dim, face, x):
    """Coordinates of points ``x`` (on face ``face`` of the reference
    simplex) in the biunit (dim-1)-simplex of that face."""
    verts = REF_VERTICES[dim][list(FACE_VERTICES[dim][face])]
    x = np.asarray(x, dtype=np.float64)
    # x = va * (1 - sum xi_k)/... : x - va = sum_k (v_{k+1} - va) (1+xi_k)/2
    edges = (verts[1:] - verts[0]).T           # (dim, dim-1)
    coeff, *_ = np.linalg.lstsq(edges, (x - verts[0]).T, rcond=None)
    return (2 * coeff - 1).T


@dataclass(frozen=True)
class ReferenceElement:
    """Element-local operators of a nodal simplex of order ``order``.

    Attributes (``Np`` nodes, ``Nf`` faces, ``Nfp`` nodes per face):

    * ``nodes``: ``(Np, dim)`` reference coordinates
    * ``tuples``: ``(Np, dim)`` integer lattice tuples
    * ``vdm``: Vandermonde ``V[i, n] = psi_n(x_i)``
    * ``mass``: ``M = (V V^T)^-1``, ``inv_mass``
    * ``diff``: ``(dim, Np, Np)`` strong derivatives ``D_r = V_r V^-1``
    * ``weak_diff``: ``(dim, Np, Np)`` ``M^-1 D_r^T M`` -- the element-local
      weak derivative ``M^-1 (d phi_j/dr, u)`` applied to nodal values
    * ``face_mask``: ``(Nf, Nfp)`` volume node index of each face node
    * ``face_mass``: ``(Nf, Nfp, Nfp)`` face mass matrices on the biunit face
    * ``lift``: ``(Np, Nf*Nfp)`` ``M^-1 E`` with ``E`` the embedded face masses
    """

    dim: int
    order: int
    nodes: np.ndarray
    tuples: np.ndarray
    vdm: np.ndarray
    mass: np.ndarray
    inv_mass: np.ndarray
    diff: np.ndarray
    weak_diff: np.ndarray
    face_mask: np.ndarray
    face_mass: np.ndarray
    lift: np.ndarray

    @property
    def nnodes(self):
        return self.nodes.shape[0]

    @property
    def nfaces(self):
        return self.dim + 1

    @property
    def nface_nodes(self):
        return self.face_mask.shape[1]

    @property
    def ref_volume(self):
        return 2.0**self.dim / math.factorial(self.dim)


@lru_cache(maxsize=None)
def get_reference_element(dim: int, order: int) -> ReferenceElement:
    if dim not in (2, 3):
        raise ValueError("only triangles and tetrahedra are supported")
    if order < 1:
        raise ValueError("order must be >= 1")
    nodes = warp_and_blend_nodes(dim, order)
    tuples = np.array(lattice_tuples(dim, order), dtype=np.int64)
    vdm = simplex_basis(dim, order, nodes)
    vinv = np.linalg.inv(vdm)
    mass = vinv.T @ vinv
    inv_mass = vdm @ vdm.T
    grads = simplex_basis_grad(dim, order, nodes)
    diff = np.stack([grads[a] @ vinv for a in range(dim)])
    weak_diff = np.stack([inv_mass @ diff[a].T @ mass for a in range(dim)])

    # barycentric lattice coordinate of every node w.r.t. each vertex
    bary = np.concatenate(
        [(order - tuples.sum(axis=1))[:, None], tuples], axis=1)
    nf = dim + 1
    face_mask = []
    for f in range(nf):
        opposite = [v for v in range(dim + 1) if v not in FACE_VERTICES[dim][f]]
        assert len(opposite) == 1
        face_mask.append(np.nonzero(bary[:, opposite[0]] == 0)[0])
    face_mask = np.array(face_mask, dtype=np.int64)
    nfp = face_mask.shape[1]

    face_mass = np.zeros((nf, nfp, nfp))
    emb = np.zeros((len(nodes), nf * nfp))
    for f in range(nf):
        xf = face_local_coords(dim, f, nodes[face_mask[f]])
        vf = simplex_basis(dim - 1, order, xf)
        mf = np.linalg.inv(vf @ vf.T)
        face_mass[f] = mf
        emb[face_mask[f], f * nfp:(f + 1) * nfp] = mf
    lift = inv_mass @ emb
    return ReferenceElement(
        dim=dim, order=order, nodes=nodes, tuples=tuples, vdm=vdm, mass=mass,
        inv_mass=inv_mass, diff=diff, weak_diff=weak_diff,
        face_mask=face_mask, face_mass=face_mass, lift=lift)


# {{{ quadrature (overintegration)

def simplex_quadrature(dim, degree):
    """Collapsed Gauss-Jacobi (Stroud conical product) rule on the biunit
    ``dim``-simplex, exact for total degree ``<= degree``: ``n =
    ceil((degree + 1) / 2)`` points per collapsed direction with the Jacobi
    weights ``(1-b)^1``, ``(1-c)^2`` absorbing the Duffy Jacobian.  Returns
    ``(x (nq, dim), w (nq,))``; ``sum(w)`` is the simplex volume."""
    from scipy.special import roots_jacobi
    n = max(1, (degree + 2) // 2)
    a, wa = roots_jacobi(n, 0.0, 0.0)
    if dim == 1:
        return a[:, None], wa
    b, wb = roots_jacobi(n, 1.0, 0.0)
    if dim == 2:
        aa, bb = np.meshgrid(a, b, indexing="ij")
        r = 0.5 * (1 + aa) * (1 - bb) - 1
        w = 0.5 * np.outer(wa, wb)
        return np.stack([r.ravel(), bb.ravel()], axis=1), w.ravel()
    if dim == 3:
        c, wc = roots_jacobi(n, 2.0, 0.0)
        aa, bb, cc = np.meshgrid(a, b, c, indexing="ij")
        s = 0.5 * (1 + bb) * (1 - cc) - 1
        r = 0.25 * (1 + aa) * (1 - bb) * (1 - cc) - 1
        w = 0.125 * np.einsum("i,j,k->ijk", wa, wb, wc)
        return np.stack([r.ravel(), s.ravel(), cc.ravel()], axis=1), w.ravel()
    raise ValueError("dim must be 1, 2 or 3")


def face_quadrature(dim, degree):
    """Rule on the biunit (dim-1)-simplex of an element face, exact to
    ``degree`` and invariant under every permutation of the face's vertices,
    so that the two elements sharing a face evaluate the numerical flux at
    the same physical points whatever their local face orientation:
    Gauss-Legendre on segments; on triangles the collapsed rule (already
    symmetric in the two vertices of its base edge) taken with each of the
    three vertices as the collapsed one, weights divided by three."""
    x, w = simplex_quadrature(dim - 1, degree)
    if dim - 1 < 2:
        return x, w
    lam = np.stack([-(x[:, 0] + x[:, 1]) / 2, (1 + x[:, 0]) / 2,
                    (1 + x[:, 1]) / 2], axis=1)             # barycentric
    verts = np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]])
    pts = [np.roll(lam, k, axis=1) @ verts for k in range(3)]
    return np.concatenate(pts), np.concatenate([w / 3.0] * 3)


@dataclass(frozen=True)
class QuadratureOperators:
    """Element operators of the quadrature discretization (grudge's
    DISCR_TAG_QUAD, mirgecom/discretization.py:65-83) for a nodal element:

    * ``points (nq, dim)``, ``weights (nq,)``: the volume rule
    * ``interp (nq, Np)``: nodal values -> quadrature points
    * ``weak_diff (dim, Np, nq)``: ``M^-1 D_q,r^T W`` -- applied to flux
      values at the points, the weak derivative ``M^-1 (d phi/dr, F)``
    * ``face_points (nfq, dim-1)``, ``face_weights (nfq,)``: the face rule on
      the biunit (dim-1)-simplex (face-local coordinates; ``face_quadrature``)
    * ``face_interp (Nf, nfq, Nfp)``: face-node values (in face_mask order)
      -> face quadrature points
    * ``face_xref (Nf, nfq, dim)``: those points in element coordinates
    * ``lift (Np, Nf*nfq)``: ``M^-1 E_q`` with ``E_q[i, (f, m)] = w_m
      phi_i(x_fm)`` on the biunit face
    """

    degree: int
    points: np.ndarray
    weights: np.ndarray
    interp: np.ndarray
    weak_diff: np.ndarray
    face_points: np.ndarray
    face_weights: np.ndarray
    face_interp: np.ndarray
    face_xref: np.ndarray
    lift: np.ndarray

    @property
    def nq(self):
        return self.points.shape[0]

    @property
    def nfq(self):
        return self.face_points.shape[0]


@lru_cache(maxsize=None)
def get_quadrature_operators(dim: int, order: int,
                             degree: int) -> QuadratureOperators:
    re = get_reference_element(dim, order)
    vinv = np.linalg.inv(re.vdm)
    xq, wq = simplex_quadrature(dim, degree)
    interp = simplex_basis(dim, order, xq) @ vinv
    gq = simplex_basis_grad(dim, order, xq)
    weak_diff = np.stack([re.inv_mass @ (gq[r] @ vinv).T @ np.diag(wq)
                          for r in range(dim)])
    xf, wf = face_quadrature(dim, degree)
    nf, nfq = dim + 1, xf.shape[0]
    face_interp = np.zeros((nf, nfq, re.nface_nodes))
    face_xref = np.zeros((nf, nfq, dim))
    emb = np.zeros((re.nnodes, nf * nfq))
    for f in range(nf):
        verts = REF_VERTICES[dim][list(FACE_VERTICES[dim][f])]
        xe = verts[0] + 0.5 * (1 + xf) @ (verts[1:] - verts[0])
        face_xref[f] = xe
        xn = face_local_coords(dim, f, re.nodes[re.face_mask[f]])
        face_interp[f] = (simplex_basis(dim - 1, order, xf)
                          @ np.linalg.inv(simplex_basis(dim - 1, order, xn)))
        emb[:, f * nfq:(f + 1) * nfq] = (simplex_basis(dim, order, xe)
                                         @ vinv).T * wf[None, :]
    return QuadratureOperators(
        degree=degree, points=xq, weights=wq, interp=interp,
        weak_diff=weak_diff, face_points=xf, face_weights=wf,
        face_interp=face_interp, face_xref=face_xref, lift=re.inv_mass @ emb)

# }}}


# {{{ hybridized SBP operators (entropy-stable DG)

@dataclass(frozen=True)
class HybridOperators:
    """Reference-element pieces of the hybridized summation-by-parts
    operators that grudge's flux differencing uses (Chan, J. Comput. Phys.
    362, 2018), on the quadrature of ``get_quadrature_operators``:

    * ``vh (nh, Np)``: nodal values -> the ``nh = nq + Nf*nfq`` hybridized
      points (volume points, then each face's points)
    * ``proj (Np, nq)``: ``P_q = M^-1 V_q^T W`` (L2 projection)
    * ``skew (dim, nq, nq)``: ``Q_r - Q_r^T`` with ``Q_r = W V_q D_r P_q``
    * ``extrap (Nf*nfq, nq)``: ``E = V_f P_q``
    * ``apply (Np, nh)``: ``M^-1 V_h^T``

    ``skew_diag`` and ``extrap_diag`` hold ``skew`` and ``extrap`` along
    their cyclic diagonals (the order in which the device kernel pairs the
    points).
    """

    vh: np.ndarray
    proj: np.ndarray
    skew: np.ndarray
    extrap: np.ndarray
    apply: np.ndarray

    @property
    def nh(self):
        return self.vh.shape[0]

    @property
    def skew_diag(self):
        """``(dim, nq, nq)``: ``[r][s][h] = skew[r][h][(h + s) mod nq]``."""
        nq = self.skew.shape[1]
        h = np.arange(nq)[None, :]
        return self.skew[:, h, (h + np.arange(nq)[:, None]) % nq]

    @property
    def extrap_diag(self):
        """``(nq, Nf*nfq)``: ``[s][m] = extrap[m][(m + s) mod nq]``."""
        nft, nq = self.extrap.shape
        m = np.arange(nft)[None, :]
        return self.extrap[m, (m + np.arange(nq)[:, None]) % nq]


@lru_cache(maxsize=None)
def get_hybrid_operators(dim: int, order: int, degree: int) -> HybridOperators:
    re = get_reference_element(dim, order)
    qo = get_quadrature_operators(dim, order, degree)
    vinv = np.linalg.inv(re.vdm)
    vq = qo.interp
    vf = np.concatenate([simplex_basis(dim, order, qo.face_xref[f]) @ vinv
                         for f in range(dim + 1)])
    proj = re.inv_mass @ vq.T * qo.weights[None, :]
    qr = np.stack([(qo.weights[:, None] * vq) @ re.diff[r] @ proj
                   for r in range(dim)])
    vh = np.concatenate([vq, vf])
    return HybridOperators(vh=vh, proj=proj, skew=qr - np.transpose(qr, (0, 2, 1)),
                           extrap=vf @ proj, apply=re.inv_mass @ vh.T)

# }}}
