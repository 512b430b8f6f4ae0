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


class QuadStates:
    """``make_operator_fluid_states`` on the quadrature discretization
    (mirgecom/gas_model.py:728-904 with ``quadrature_tag``): the state
    interpolated to the volume points and to both sides' face points, the
    EOS evaluated there (``project_fluid_state``, gas_model.py:494-588;
    mixtures seed Newton with the interpolated nodal temperature), and each
    boundary's exterior state evaluated at its faces' points.  Flat point
    axes: volume ``K*nq``, faces ``K*Nf*nfq``."""

    def __init__(self, disc, box, cv, gas_model, boundaries, time, degree,
                 temperature=None):
        el = disc.el
        dim, nf = disc.dim, disc.nf
        self.disc, self.gas_model = disc, gas_model
        self.mix = temperature is not None
        self.xq, self.wq = stroud_rule(dim, degree)
        self.phi = _lagrange(el, self.xq)
        self.dphi = _lagrange_grad(el, self.xq)
        self.nq = len(self.wq)
        self.xf, self.wf = face_rule(dim, degree)
        self.nfq = len(self.wf)
        self.phif, self.tbasis, xref = [], [], []
        for f, fv in enumerate(FACES[dim]):
            vf = REF_VERTS[dim][list(fv)]
            xe = vf[0] + 0.5 * (1 + self.xf) @ (vf[1:] - vf[0])
            xref.append(xe)
            pf = _lagrange(el, xe)                              # (nfq, Np)
            self.phif.append(pf)
            self.tbasis.append(pf[:, el.face_mask[f]])          # (nfq, Nfp)
        nvar = cv.shape[0]
        uq = self.to_vol(cv)
        tq = self.to_vol(temperature).reshape(-1) if self.mix else None
        self.vol = ph.make_fluid_state(uq.reshape(nvar, -1), gas_model, dim,
                                       temperature_seed=tq)
        um = self.to_face(cv)
        tm = tp = None
        if self.mix:
            tm = self.to_face(temperature).reshape(-1)
            tp = self.to_face(np.nan_to_num(
                disc.exterior(temperature[None])[0], nan=1.0),
                exterior=True).reshape(-1)
        up = self.ext_to_face(cv)
        flat = lambda a: a.reshape(a.shape[0], -1)             # noqa: E731
        normal = np.repeat(disc.normal[:, :, None, :], self.nfq, axis=2)
        self.nrm = np.moveaxis(normal.reshape(-1, dim), 1, 0)  # (dim, Nfq)
        self.minus = ph.make_fluid_state(flat(um), gas_model, dim,
                                         temperature_seed=tm)
        self.inter = np.nonzero(
            ~disc.is_bnd[..., 0].reshape(-1).repeat(self.nfq))[0]
        self.plus_int = ph.make_fluid_state(
            flat(up)[:, self.inter], gas_model, dim,
            temperature_seed=tp[self.inter] if self.mix else None)
        # boundary faces: conditions evaluated at the faces' physical points
        xphys = (disc.v[:, None, None, 0, :]
                 + np.einsum("fmi,kid->kfmd", (1 + np.array(xref)) / 2,
                             disc.v[:, 1:, :] - disc.v[:, :1, :]))
        xphys = np.moveaxis(xphys.reshape(-1, dim), 1, 0)
        self.groups = []
        nfp = el.nfp
        for bdry, idx in _boundary_groups(disc, box, boundaries):
            faces = np.unique(idx // nfp)
            qidx = (faces[:, None] * self.nfq
                    + np.arange(self.nfq)[None]).reshape(-1)
            smb = _sub_state(self.minus, qidx)
            n = self.nrm[:, qidx]
            if bdry.kind == "dummy":
                ext = None
            elif bdry.kind == "prescribed":
                ext = bdry.state_func(xphys[:, qidx], smb, time)
            else:
                ext = bdry.state_plus(smb, n, gas_model)
            self.groups.append((bdry, qidx, smb, ext))
        self.nfq_all = disc.K * nf * self.nfq

    def to_vol(self, u):
        """Nodal ``(..., K, Np)`` -> volume points ``(..., K, nq)``."""
        return np.einsum("qn,...kn->...kq", self.phi, u)

    def to_face(self, u, exterior=False):
        """Nodal ``(..., K, Np)`` (or, with ``exterior``, face-node values
        ``(..., K, Nf, Nfp)``) -> face points ``(..., K, Nf, nfq)``."""
        tr = u if exterior else self.disc.trace(u)
        return np.stack([np.einsum("mj,...kj->...km", self.tbasis[f],
                                   tr[..., f, :])
                         for f in range(self.disc.nf)], axis=-2)

    def ext_to_face(self, u):
        """The face neighbours' polynomials at this element's face points
        (boundary faces 0)."""
        return self.to_face(np.nan_to_num(self.disc.exterior(u)), exterior=True)

    def weak_div(self, vol_flux, face_flux):
        """``M^-1 (sum_q w_q J grad(phi_i) . F_q - sum_f sum_m w_m sJ phi_i
        f_m)``; ``vol_flux (F, dim, K*nq)``, ``face_flux (F, K*Nf*nfq)`` (the
        normal flux) -> ``(F, K, Np)``."""
        d = self.disc
        nvar = vol_flux.shape[0]
        fq = vol_flux.reshape(nvar, d.dim, d.K, self.nq)
        vol = np.einsum("q,aqi,kad,fdkq->fki", self.wq, self.dphi, d.drdx, fq)
        vol = vol * d.J[None, :, None]
        fs = face_flux.reshape(nvar, d.K, d.nf, self.nfq)
        face = np.zeros_like(vol)
        for f in range(d.nf):
            face += np.einsum("m,mi,fkm->fki", self.wf, self.phif[f],
                              fs[:, :, f]) * d.sJ[None, :, f, None]
        return d.inverse_mass(vol - face)

    def weak_grad(self, u_vol, u_face):
        """grudge's weak gradient on the points (mirgecom/operators.py:34-63):
        ``M^-1 (sum_f sum_m w_m sJ phi_i u*_m n - sum_q w_q J grad(phi_i)
        u_q)``; ``u_vol (..., K*nq)``, ``u_face (..., K*Nf*nfq)`` the face
        values -> ``(..., dim, K, Np)``."""
        d = self.disc
        lead = u_vol.shape[:-1]
        uq = u_vol.reshape(lead + (d.K, self.nq))
        vol = np.einsum("q,aqi,kad,...kq->...dki", self.wq, self.dphi, d.drdx,
                        uq) * d.J[:, None]
        us = u_face.reshape(lead + (d.K, d.nf, self.nfq))
        face = np.zeros_like(vol)
        for f in range(d.nf):
            face += np.einsum("m,mi,...km,kd->...dki", self.wf, self.phif[f],
                              us[..., f, :],
                              d.normal[:, f, :] * d.sJ[:, f, None])
        return d.inverse_mass(face - vol)


def _inviscid_face_flux(qs, numerical_flux_func):
    """mirgecom/inviscid.py:290-376 at the face points -> ``(F, K*Nf*nfq)``."""
    nvar = qs.minus.cv.shape[0]
    out = np.zeros((nvar, qs.nfq_all))
    ii = qs.inter
    out[:, ii] = numerical_flux_func(_sub_state(qs.minus, ii), qs.plus_int,
                                     qs.gas_model, qs.nrm[:, ii])
    for bdry, qidx, smb, ext in qs.groups:
        n = qs.nrm[:, qidx]
        if bdry.kind == "dummy":
            out[:, qidx] = ph._dot_normal(ph.inviscid_flux(smb), n)
        else:
            out[:, qidx] = numerical_flux_func(smb, ext, qs.gas_model, n)
    return out


def euler_operator_quad(disc, box, cv, gas_model, boundaries, time=0.0,
                        numerical_flux_func=ph.inviscid_facial_flux_rusanov,
                        degree=None, temperature=None):
    """Overintegrated Euler RHS ``(F, K, Np)`` of the conserved state ``cv``;
    for a mixture, ``temperature (K, Np)`` is interpolated to every point as
    the seed of the Newton inversion there (gas_model.py:494-588)."""
    degree = 2 * disc.el.order if degree is None else degree
    qs = QuadStates(disc, box, cv, gas_model, boundaries, time, degree,
                    temperature)
    fq = ph.inviscid_flux(qs.vol)                              # (F, d, K*nq)
    return qs.weak_div(fq, _inviscid_face_flux(qs, numerical_flux_func))


def _check_state_boundaries(qs):
    for bdry, _, _, _ in qs.groups:
        if bdry.kind not in ("dummy", "prescribed"):
            raise NotImplementedError(
                "the quadrature Navier-Stokes oracle covers prescribed and "
                "dummy boundaries")


def grad_operators_quad(qs):
    """``grad_cv_operator`` / ``grad_t_operator`` with ``quadrature_tag``
    (mirgecom/navierstokes.py:129-347): the weak gradient of the conserved
    state and of the temperature at the volume points, central flux at the
    interior face points, the boundary's gradient flux at the boundary
    points (prescribed: the exterior value, boundary.py:70-72 and 895-908;
    dummy: the interior value, boundary.py:1031-1035).  Returns the nodal
    ``grad_cv (F, dim, K, Np)`` and ``grad_t (dim, K, Np)``."""
    _check_state_boundaries(qs)
    nvar = qs.vol.cv.shape[0]
    u_vol = np.concatenate([qs.vol.cv, qs.vol.temperature[None]])
    u_face = np.zeros((nvar + 1, qs.nfq_all))
    ii = qs.inter
    um = np.concatenate([qs.minus.cv, qs.minus.temperature[None]])
    up = np.concatenate([qs.plus_int.cv, qs.plus_int.temperature[None]])
    u_face[:, ii] = ph.num_flux_central(um[:, ii], up)
    for bdry, qidx, smb, ext in qs.groups:
        src = smb if bdry.kind == "dummy" else ext
        u_face[:, qidx] = np.concatenate([src.cv, src.temperature[None]])
    g = qs.weak_grad(u_vol, u_face)
    return g[:-1], g[-1]


def ns_operator_quad(disc, box, cv, gas_model, boundaries, time=0.0,
                     inviscid_numerical_flux_func=ph.inviscid_facial_flux_rusanov,
                     degree=None, temperature=None, inviscid_terms_on=True,
                     return_gradients=False):
    """Overintegrated Navier-Stokes RHS (mirgecom/navierstokes.py:350-575
    with ``quadrature_tag``): the nodal gradients from
    ``grad_operators_quad``; the viscous flux at the volume points from the
    point states and the gradients interpolated there (``op.project``); at
    the face points the central viscous flux of both sides' states and
    interpolated gradient traces (viscous.py:296-349, 431-544; boundaries:
    the interior gradients on both sides, the dummy boundary's one-sided
    flux); ``div_operator(F_V - F_I, f_V* - f_I*)`` on the points."""
    degree = 2 * disc.el.order if degree is None else degree
    qs = QuadStates(disc, box, cv, gas_model, boundaries, time, degree,
                    temperature)
    grad_cv, grad_t = grad_operators_quad(qs)
    nvar, dim = cv.shape[0], disc.dim
    gcv_q = qs.to_vol(grad_cv).reshape(nvar, dim, -1)
    gt_q = qs.to_vol(grad_t).reshape(dim, -1)
    vol = ph.viscous_flux(qs.vol, gcv_q, gt_q)
    gcv_m = qs.to_face(grad_cv).reshape(nvar, dim, -1)
    gcv_p = qs.ext_to_face(grad_cv).reshape(nvar, dim, -1)
    gt_m = qs.to_face(grad_t).reshape(dim, -1)
    gt_p = qs.ext_to_face(grad_t).reshape(dim, -1)
    bnd = np.zeros((nvar, qs.nfq_all))
    ii = qs.inter
    bnd[:, ii] = ph.viscous_facial_flux_central(
        _sub_state(qs.minus, ii), qs.plus_int, gcv_m[..., ii], gcv_p[..., ii],
        gt_m[..., ii], gt_p[..., ii], qs.nrm[:, ii])
    for bdry, qidx, smb, ext in qs.groups:
        n = qs.nrm[:, qidx]
        if bdry.kind == "dummy":
            bnd[:, qidx] = ph._dot_normal(
                ph.viscous_flux(smb, gcv_m[..., qidx], gt_m[..., qidx]), n)
        else:
            bnd[:, qidx] = ph.viscous_facial_flux_central(
                smb, ext, gcv_m[..., qidx], gcv_m[..., qidx], gt_m[..., qidx],
                gt_m[..., qidx], n)
    if inviscid_terms_on:
        vol = vol - ph.inviscid_flux(qs.vol)
        bnd = bnd - _inviscid_face_flux(qs, inviscid_numerical_flux_func)
    # div_operator (mirgecom/operators.py:66-95): -M^-1 (S^T v - face_mass)
    rhs = -qs.weak_div(vol, bnd)
    if return_gradients:
        return rhs, grad_cv, grad_t
    return rhs
