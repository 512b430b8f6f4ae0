"""Oracle Euler operator on a quadrature discretization (NumPy FP64; test
infrastructure only): ``euler_operator(..., quadrature_tag=DISCR_TAG_QUAD)``
of mirgecom/euler.py:367-416 with grudge's quadrature-based weak forms
(mirgecom/operators.py:34-95 on ``dd_vol_quad`` / ``dd_allfaces_quad``,
mirgecom/discretization.py:65-83 for the quadrature group).

    dQ/dt = M^-1 [ sum_q w_q J grad(phi_i)(x_q) . F(u(x_q))
                   - sum_f sum_m w_m sJ_f phi_i(x_fm) F*(u-(x_fm), u+(x_fm)) ]

* the state is interpolated to the points (``project_fluid_state``) and the
  flux formed there; both face traces are the degree-p face polynomials
  evaluated at the face points; boundary exterior states come from the
  boundary conditions evaluated at the boundary faces' points;
* volume rule: the collapsed (Stroud) Gauss-Jacobi rule of degree ``2p``,
  ``n = p + 1`` points per collapsed direction, coordinates
  ``s = b, r = (1+a)(1-b)/2 - 1`` (2-D) and
  ``t = c, s = (1+b)(1-c)/2 - 1, r = (1+a)(1-b)(1-c)/4 - 1`` (3-D);
* face rule: Gauss-Legendre on segments; on triangles the collapsed rule
  taken with each vertex as the collapsed one (weights / 3), a point set
  invariant under every face-vertex permutation.

grudge uses modepy's Xiao-Gimbutas rules for the same degree; the weak
forms agree exactly wherever the integrands are polynomials of degree
``<= 2p`` and differ by quadrature error otherwise (parity with the
reference's rule unpinned: modepy is absent).  Pinned by: rule exactness to
degree ``2p``; uniform states give 0; for fluxes of degree ``<= p`` the
result equals the nodal operator's; for a degree-``2p`` momentum flux the
mass/momentum rows equal the brute-force high-degree evaluation of the same
weak form (tests/test_oracle_overintegration.py).
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import roots_jacobi

from oracle import physics as ph
from oracle.disc import FACES, REF_VERTS, _mono
from oracle.operators import _boundary_groups, _sub_state


def stroud_rule(dim, degree):
    """Collapsed Gauss-Jacobi rule on the biunit ``dim``-simplex (see the
    module docstring for the coordinate convention)."""
    n = max(1, (degree + 2) // 2)
    a, wa = roots_jacobi(n, 0.0, 0.0)
    if dim == 1:
        return a[:, None], wa
    b, wb = roots_jacobi(n, 1.0, 0.0)
    if dim == 2:
        pts, wts = [], []
        for ai, wai in zip(a, wa):
            for bj, wbj in zip(b, wb):
                pts.append([(1 + ai) * (1 - bj) / 2 - 1, bj])
                wts.append(wai * wbj / 2)
        return np.array(pts), np.array(wts)
    c, wc = roots_jacobi(n, 2.0, 0.0)
    pts, wts = [], []
    for ai, wai in zip(a, wa):
        for bj, wbj in zip(b, wb):
            for ck, wck in zip(c, wc):
                pts.append([(1 + ai) * (1 - bj) * (1 - ck) / 4 - 1,
                            (1 + bj) * (1 - ck) / 2 - 1, ck])
                wts.append(wai * wbj * wck / 8)
    return np.array(pts), np.array(wts)


def face_rule(dim, degree):
    """Rule on the biunit (dim-1)-simplex of a face, invariant under the
    face's vertex permutations."""
    x, w = stroud_rule(dim - 1, degree)
    if dim - 1 < 2:
        return x, w
    v = REF_VERTS[2]
    bary = np.stack([-(x[:, 0] + x[:, 1]) / 2, (1 + x[:, 0]) / 2,
                     (1 + x[:, 1]) / 2], axis=1)
    pts = np.concatenate([bary[:, [(k + i) % 3 for i in range(3)]] @ v
                          for k in range(3)])
    return pts, np.concatenate([w / 3] * 3)


def _lagrange(el, x):
    """Lagrange basis of ``el`` at biunit points ``x (n, dim)`` -> (n, Np)."""
    return _mono((1 + x) / 2, el.exps) @ el.coef


def _lagrange_grad(el, x):
    """Biunit-coordinate gradients ``(dim, n, Np)``."""
    y = (1 + x) / 2
    out = np.zeros((el.dim, x.shape[0], el.np))
    for i in range(el.dim):
        dm = np.zeros((x.shape[0], len(el.exps)))
        for k, a in enumerate(el.exps):
            if a[i] == 0:
                continue
            b = list(a)
            b[i] -= 1
            dm[:, k] = 0.5 * a[i] * np.prod(y ** np.array(b)[None, :], axis=1)
        out[i] = dm @ el.coef
    return out


def euler_operator_quad(disc, box, cv, gas_model, boundaries, time=0.0,
                        numerical_flux_func=ph.inviscid_facial_flux_rusanov,
                        degree=None, temperature=None):
    """Overintegrated Euler RHS ``(F, K, Np)`` of the conserved state ``cv``;
    for a mixture, ``temperature (K, Np)`` is interpolated to every point as
    the seed of the Newton inversion there (gas_model.py:494-588)."""
    el = disc.el
    dim, nf, nfp = disc.dim, disc.nf, el.nfp
    degree = 2 * el.order if degree is None else degree
    nvar = cv.shape[0]
    # volume
    xq, wq = stroud_rule(dim, degree)
    phi = _lagrange(el, xq)
    dphi = _lagrange_grad(el, xq)
    mix = temperature is not None
    uq = np.einsum("qn,fkn->fkq", phi, cv)
    tq = np.einsum("qn,kn->kq", phi, temperature).reshape(-1) if mix else None
    sq = ph.make_fluid_state(uq.reshape(nvar, -1), gas_model, dim,
                             temperature_seed=tq)
    fq = ph.inviscid_flux(sq).reshape(nvar, dim, disc.K, -1)   # (F, d, K, q)
    vol = np.einsum("q,aqi,kad,fdkq->fki", wq, dphi, disc.drdx, fq)
    vol = vol * disc.J[None, :, None]
    # faces
    xf, wf = face_rule(dim, degree)
    nfq = len(wf)
    ext_nodes = disc.exterior(cv)                              # (F, K, Nf, Nfp)
    own_nodes = disc.trace(cv)
    um = np.zeros((nvar, disc.K, nf, nfq))
    up = np.zeros((nvar, disc.K, nf, nfq))
    tm = np.zeros((disc.K, nf, nfq))
    tp = np.zeros((disc.K, nf, nfq))
    if mix:
        t_own = disc.trace(temperature[None])[0]                   # (K, Nf, Nfp)
        t_ext = np.nan_to_num(disc.exterior(temperature[None])[0], nan=1.0)
    phif = []
    xref = []
    for f, fv in enumerate(FACES[dim]):
        vf = REF_VERTS[dim][list(fv)]
        xe = vf[0] + 0.5 * (1 + xf) @ (vf[1:] - vf[0])
        xref.append(xe)
        pf = _lagrange(el, xe)                                  # (nfq, Np)
        phif.append(pf)
        trace_basis = pf[:, el.face_mask[f]]                    # (nfq, Nfp)
        um[:, :, f] = np.einsum("mj,fkj->fkm", trace_basis, own_nodes[:, :, f])
        up[:, :, f] = np.einsum("mj,fkj->fkm", trace_basis,
                                np.nan_to_num(ext_nodes[:, :, f]))
        if mix:
            tm[:, f] = np.einsum("mj,kj->km", trace_basis, t_own[:, f])
            tp[:, f] = np.einsum("mj,kj->km", trace_basis, t_ext[:, f])
    normal = np.repeat(disc.normal[:, :, None, :], nfq, axis=2)  # (K, Nf, q, d)
    flat = lambda a: a.reshape(a.shape[0], -1)                 # noqa: E731
    nrm = np.moveaxis(normal.reshape(-1, dim), 1, 0)
    sm = ph.make_fluid_state(flat(um), gas_model, dim,
                             temperature_seed=tm.reshape(-1) if mix else None)
    fstar = np.zeros((nvar, disc.K * nf * nfq))
    inter = np.nonzero(~disc.is_bnd[..., 0].reshape(-1).repeat(nfq))[0]
    sp_int = ph.make_fluid_state(flat(up)[:, inter], gas_model, dim,
                                 temperature_seed=(tp.reshape(-1)[inter]
                                                   if mix else None))
    fstar[:, inter] = numerical_flux_func(_sub_state(sm, inter), sp_int,
                                          gas_model, nrm[:, inter])
    # boundary faces: conditions evaluated at the faces' physical points
    xphys = (disc.v[:, None, None, 0, :]
             + np.einsum("fmi,kid->kfmd", (1 + np.array(xref)) / 2,
                         disc.v[:, 1:, :] - disc.v[:, :1, :]))
    xphys = np.moveaxis(xphys.reshape(-1, dim), 1, 0)
    for bdry, idx in _boundary_groups(disc, box, boundaries):
        faces = np.unique(idx // nfp)
        qidx = (faces[:, None] * nfq + np.arange(nfq)[None]).reshape(-1)
        smb = _sub_state(sm, qidx)
        n = nrm[:, qidx]
        if bdry.kind == "dummy":
            fstar[:, qidx] = ph._dot_normal(ph.inviscid_flux(smb), n)
            continue
        if bdry.kind == "prescribed":
            ext = bdry.state_func(xphys[:, qidx], smb, time)
        else:
            ext = bdry.state_plus(smb, n, gas_model)
        fstar[:, qidx] = numerical_flux_func(smb, ext, gas_model, n)
    fstar = fstar.reshape(nvar, disc.K, nf, nfq)
    face = np.zeros_like(vol)
    for f in range(nf):
        face += np.einsum("m,mi,fkm->fki", wf, phif[f], fstar[:, :, f]) \
            * disc.sJ[None, :, f, None]
    return disc.inverse_mass(vol - face)
