"""Oracle entropy-stable (flux-differencing) Euler operator (NumPy FP64;
test infrastructure only): ``euler_operator(..., use_esdg=True,
quadrature_tag=DISCR_TAG_QUAD)`` of mirgecom/euler.py:108-310 for the ideal
gas, with the reference's entropy variables (gas_model.py:1026-1100,
Chandrashekar 2013 eq. 4.5; Hughes, Franca & Mallet 1986 for the inverse),
its entropy-conserving two-point flux (inviscid.py:451-518, Chandrashekar
2013 eqs. 4.12-4.14) and entropy-stable Rusanov face flux
(inviscid.py:601-711), and grudge's hybridized-SBP flux differencing as
documented by grudge and by Chan (J. Comput. Phys. 362, 2018):

    V_q, V_f: nodal values -> volume / face quadrature points
    P_q = M^-1 V_q^T W,   E = V_f P_q,   Q_r = W V_q D_r P_q
    Q_x = J sum_r (dr/dx) Q_r,   B_x = diag(n_x sJ w_f)
    Q_h = [[Q_x - E^T B_x E / 2,  E^T B_x / 2], [-B_x E / 2,  B_x / 2]]
    dQ/dt = -M^-1 [ sum_x [V_q; V_f]^T ((Q_h - Q_h^T) o F_x(u_i, u_j)) 1
                    + V_f^T diag(sJ w_f) f* ]

with u the conserved state at the hybridized points obtained from the
L2-projected entropy variables (``volume_quadrature_project`` then
``volume_and_surface_quadrature_interpolation``).  Faces use the exterior
element's projected entropy variables at the same (vertex-symmetric) face
points; boundary faces the interpolated interior state and the boundary
condition's exterior state, as the reference does.

grudge itself is absent, so parity with it is unpinned beyond these
definitions; the restatement is pinned by the properties that define the
scheme (tests/test_oracle_esdg.py): discrete entropy conservation with the
entropy-conserving face flux, entropy dissipation with the Rusanov one,
free-stream preservation, conservation, the two-point flux's consistency
and symmetry, and the entropy-variable round trip.
"""

from __future__ import annotations

import numpy as np

from oracle import physics as ph
from oracle.disc import FACES, REF_VERTS, _mono
from oracle.operators import _boundary_groups
from oracle.overintegration import _lagrange, face_rule, stroud_rule


# {{{ entropy variables and fluxes

def entropy_vars(q, gamma, dim):
    """gas_model.py:1026-1060: ``(F, ...)`` conserved -> entropy variables."""
    rho = q[0]
    u = q[2:2 + dim] / rho
    u2 = np.sum(u * u, axis=0)
    p = (gamma - 1) * (q[1] - 0.5 * rho * u2)
    s = np.log(p) - gamma * np.log(rho)
    rho_p = rho / p
    ev = [(gamma - s) / (gamma - 1) - 0.5 * rho_p * u2, -rho_p]
    ev += [rho_p * u[d] for d in range(dim)]
    ev += [q[k] / rho / (gamma - 1) for k in range(2 + dim, q.shape[0])]
    return np.stack(ev)


def conserved_vars(ev, gamma, dim):
    """gas_model.py:1063-1100: entropy variables -> conserved."""
    v = ev * (gamma - 1)
    v1, v5 = v[0], v[1]
    v234 = v[2:2 + dim]
    s = gamma - v1 + np.sum(v234 * v234, axis=0) / (2 * v5)
    iota = ((gamma - 1) / (-v5) ** gamma) ** (1 / (gamma - 1))
    rho_iota = iota * np.exp(-s / (gamma - 1))
    mass = -rho_iota * v5
    out = [mass, rho_iota * (1 - np.sum(v234 * v234, axis=0) / (2 * v5))]
    out += [rho_iota * v234[d] for d in range(dim)]
    out += [mass * v[k] for k in range(2 + dim, ev.shape[0])]
    return np.stack(out)


def ln_mean(x, y, epsilon=1e-4):
    """inviscid.py:470-476: logarithmic mean with its series near x = y."""
    f2 = (x * (x - 2 * y) + y * y) / (x * (x + 2 * y) + y * y)
    safe = np.where(f2 < epsilon, 2.0, y / x)
    return np.where(f2 < epsilon,
                    (x + y) / (2 + f2 * 2 / 3 + f2 * f2 * 2 / 5 + f2 ** 3 * 2 / 7),
                    (y - x) / np.log(safe))


def chandrashekar(ql, qr, gamma, dim):
    """inviscid.py:451-518: two-point flux ``(F, dim, ...)``."""
    def prim(q):
        rho = q[0]
        u = q[2:2 + dim] / rho
        p = (gamma - 1) * (q[1] - 0.5 * rho * np.sum(u * u, axis=0))
        return rho, u, p, q[2 + dim:] / rho
    rl, ul, pl, yl = prim(ql)
    rr, ur, pr, yr = prim(qr)
    beta_l, beta_r = 0.5 * rl / pl, 0.5 * rr / pr
    kin_l = 0.5 * np.sum(ul * ul, axis=0)
    kin_r = 0.5 * np.sum(ur * ur, axis=0)
    rho_avg = 0.5 * (rl + rr)
    rho_mean = ln_mean(rl, rr)
    beta_mean = ln_mean(beta_l, beta_r)
    beta_avg = 0.5 * (beta_l + beta_r)
    u_avg = 0.5 * (ul + ur)
    p_mean = 0.5 * rho_avg / beta_avg
    vsq_avg = kin_l + kin_r
    mass = rho_mean * u_avg                                   # (dim, ...)
    mom = mass[:, None] * u_avg[None, :] + np.eye(dim).reshape(
        (dim, dim) + (1,) * (mass.ndim - 1)) * p_mean        # [i][d]
    energy = (mass * 0.5 * (1 / (gamma - 1) / beta_mean - vsq_avg)
              + np.einsum("id...,i...->d...", mom, u_avg))
    spec = [rho_mean * 0.5 * (yl[k] + yr[k]) * u_avg for k in range(yl.shape[0])]
    return np.stack([mass, energy] + [mom[i] for i in range(dim)] + spec)


def wavespeed(q, gamma, dim):
    rho = q[0]
    u = q[2:2 + dim] / rho
    u2 = np.sum(u * u, axis=0)
    p = (gamma - 1) * (q[1] - 0.5 * rho * u2)
    return np.sqrt(u2) + np.sqrt(gamma * p / rho)


def es_facial_flux(qm, qp, normal, gamma, dim, dissipation=True):
    """inviscid.py:601-711: ``F_S(q-, q+) . n - alpha/2 (q+ - q-)``, alpha
    the larger wavespeed (Rusanov), or no penalty."""
    f = np.einsum("fd...,d...->f...", chandrashekar(qm, qp, gamma, dim), normal)
    if dissipation:
        alpha = np.maximum(wavespeed(qm, gamma, dim), wavespeed(qp, gamma, dim))
        f = f - 0.5 * alpha * (qp - qm)
    return f

# }}}


class _HybridOps:
    """Reference-element hybridized SBP pieces of the oracle element."""

    def __init__(self, el, degree):
        dim = el.dim
        xq, wq = stroud_rule(dim, degree)
        self.vq = _lagrange(el, xq)                               # (nq, Np)
        xf, wf = face_rule(dim, degree)
        self.nfq = len(wf)
        self.wf = wf
        vf = []
        for fv in FACES[dim]:
            v = REF_VERTS[dim][list(fv)]
            vf.append(_lagrange(el, v[0] + 0.5 * (1 + xf) @ (v[1:] - v[0])))
        self.vf = np.concatenate(vf)                              # (Nf*nfq, Np)
        mass = self.vq.T @ (wq[:, None] * self.vq)
        self.minv = np.linalg.inv(mass)
        self.pq = self.minv @ self.vq.T * wq[None, :]             # (Np, nq)
        self.e = self.vf @ self.pq                                # (Nf*nfq, nq)
        self.qr = np.stack([(wq[:, None] * self.vq) @ el.diff[r] @ self.pq
                            for r in range(dim)])                 # (dim, nq, nq)
        self.vh = np.concatenate([self.vq, self.vf])              # (nh, Np)
        self.nq = len(wq)


def esdg_euler_operator(disc, box, cv, gamma, boundaries, time=0.0,
                        dissipation=True, degree=None):
    """Entropy-stable Euler RHS ``(F, K, Np)`` (ideal gas, ratio ``gamma``)."""
    el = disc.el
    dim, nf, nfp, K = disc.dim, disc.nf, el.nfp, disc.K
    degree = 2 * el.order if degree is None else degree
    ho = _HybridOps(el, degree)
    nq, nfq, nft = ho.nq, ho.nfq, nf * ho.nfq
    nvar = cv.shape[0]
    # entropy projection
    uq = np.einsum("qn,fkn->fkq", ho.vq, cv)
    ev_nodal = np.einsum("nq,fkq->fkn", ho.pq, entropy_vars(uq, gamma, dim))
    uh = conserved_vars(np.einsum("hn,fkn->fkh", ho.vh, ev_nodal), gamma, dim)
    # volume flux differencing, element by element
    vol = np.zeros((nvar, K, el.np))
    sj = disc.sJ                                                  # (K, Nf)
    for k in range(K):
        fmat = chandrashekar(uh[:, k, :, None], uh[:, k, None, :], gamma, dim)
        for d in range(dim):
            qx = disc.J[k] * np.einsum("r,rij->ij", disc.drdx[k][:, d], ho.qr)
            bx = np.repeat(disc.normal[k, :, d] * sj[k], nfq) * np.tile(ho.wf, nf)
            qh = np.zeros((nq + nft, nq + nft))
            qh[:nq, :nq] = qx - 0.5 * ho.e.T @ (bx[:, None] * ho.e)
            qh[:nq, nq:] = 0.5 * ho.e.T * bx[None, :]
            qh[nq:, :nq] = -0.5 * bx[:, None] * ho.e
            qh[nq:, nq:] = 0.5 * np.diag(bx)
            skew = qh - qh.T
            vol[:, k] += np.einsum("hn,fhj->fn", ho.vh,
                                   skew[None] * fmat[:, d])
    # face fluxes at the face points
    uf_m = uh[:, :, nq:].reshape(nvar, K, nf, nfq)
    ev_ext = disc.exterior(ev_nodal)                              # (F, K, Nf, Nfp)
    up = np.zeros_like(uf_m)
    for f in range(nf):
        tb = ho.vf[f * nfq:(f + 1) * nfq][:, el.face_mask[f]]    # (nfq, Nfp)
        evp = np.einsum("mj,fkj->fkm", tb, np.nan_to_num(ev_ext[:, :, f], nan=-1.0))
        up[:, :, f] = conserved_vars(evp, gamma, dim)
    normal = np.moveaxis(np.repeat(disc.normal[:, :, None, :], nfq, axis=2),
                         -1, 0)                                   # (d, K, Nf, q)
    fstar = es_facial_flux(uf_m, up, normal, gamma, dim, dissipation)
    # boundary faces: interpolated interior state, the condition's exterior
    if boundaries:
        trace_q = np.einsum("hn,fkn->fkh", ho.vf, cv).reshape(nvar, K, nf, nfq)
        xf_ref = []
        xf, _ = face_rule(dim, degree)
        for fv in FACES[dim]:
            v = REF_VERTS[dim][list(fv)]
            xf_ref.append(v[0] + 0.5 * (1 + xf) @ (v[1:] - v[0]))
        xphys = (disc.v[:, None, None, 0, :]
                 + np.einsum("fmi,kid->kfmd", (1 + np.array(xf_ref)) / 2,
                             disc.v[:, 1:, :] - disc.v[:, :1, :]))
        gm = ph.GasModel(ph.IdealSingleGas(gamma=gamma))
        for bdry, idx in _boundary_groups(disc, box, boundaries):
            faces = np.unique(idx // nfp)
            kk, ff = faces // nf, faces % nf
            qm = conserved_vars(entropy_vars(trace_q[:, kk, ff], gamma, dim),
                                gamma, dim)                       # (F, nb, q)
            n = normal[:, kk, ff]
            if bdry.kind == "dummy":
                qp = qm
            else:
                sm = ph.make_fluid_state(qm.reshape(nvar, -1), gm, dim)
                ext = bdry.state_func(np.moveaxis(xphys[kk, ff], -1, 0)
                                      .reshape(dim, -1), sm, time)
                qp = ext.cv.reshape(qm.shape)
            fstar[:, kk, ff] = es_facial_flux(qm, qp, n, gamma, dim, dissipation)
    face = np.einsum("hn,fkh->fkn", ho.vf,
                     (fstar * (sj[:, :, None] * ho.wf[None, None, :])[None])
                     .reshape(nvar, K, nft))
    return -np.einsum("ij,fkj->fki", ho.minv, vol + face) / disc.J[None, :, None]


# {{{ mixtures: Renac's two-point flux and the state-based pipeline

def entropy_vars_state(q, p, gamma, dim):
    """gas_model.py:1026-1060 with the state's own pressure ``p`` and ratio
    ``gamma`` (for a mixture, ``p = rho R T`` from the Newton temperature,
    not ``(gamma - 1) rho e``)."""
    rho = q[0]
    u = q[2:2 + dim] / rho
    u2 = np.sum(u * u, axis=0)
    s = np.log(p) - gamma * np.log(rho)
    rho_p = rho / p
    ev = [(gamma - s) / (gamma - 1) - 0.5 * rho_p * u2, -rho_p]
    ev += [rho_p * u[d] for d in range(dim)]
    ev += [q[k] / rho / (gamma - 1) for k in range(2 + dim, q.shape[0])]
    return np.stack(ev)


def renac(sl, sr, eos, dim):
    """inviscid.py:521-598: the entropy-conserving, kinetic-energy-preserving
    two-point flux of Renac (2021, eq. 24) between states ``sl`` and ``sr``
    (``oracle.physics.State``), ``(F, dim, ...)``.  The energy flux takes the
    mixture's c_v at the logarithmic-mean temperature and the averaged mass
    fractions."""
    tl, tr = sl.temperature, sr.temperature
    pl, pr = sl.pressure, sr.pressure
    gl, gr = eos.gamma(sl.cv, tl), eos.gamma(sr.cv, tr)
    thl, thr = 1.0 / tl, 1.0 / tr
    t_avg = 0.5 * (tl + tr)
    pot_avg = 0.5 * (pl * thl + pr * thr)
    t_mean = 1.0 / ln_mean(thl, thr)
    pec_avg = pot_avg * t_avg
    p_mean = ln_mean(pl, pr)
    rl, rr = sl.mass, sr.mass
    ul, ur = sl.velocity, sr.velocity
    yl, yr = sl.species_mass_fractions, sr.species_mass_fractions
    kin_comb = 0.5 * np.sum(ul * ur, axis=0)
    rho_mean = ln_mean(rl, rr)
    spec_mean = rho_mean * 0.5 * (yl + yr)
    u_avg = 0.5 * (ul + ur)
    mass = rho_mean * u_avg                                    # (dim, ...)
    mom = mass[:, None] * u_avg[None, :] + np.eye(dim).reshape(
        (dim, dim) + (1,) * (mass.ndim - 1)) * pec_avg         # [i][d]
    gamma_avg = 0.5 * (gl + gr)
    ener_es = p_mean / (gamma_avg - 1) + 0.5 * rho_mean * np.sum(u_avg * u_avg,
                                                                 axis=0)
    cv_es = np.concatenate([rho_mean[None], ener_es[None], rho_mean * u_avg,
                            spec_mean])
    cvm = eos.heat_capacity_cv(cv_es, t_mean)
    energy = (cvm * t_mean + kin_comb) * mass + pec_avg * u_avg
    spec = [spec_mean[k] * u_avg for k in range(spec_mean.shape[0])]
    return np.stack([mass, energy] + [mom[i] for i in range(dim)] + spec)


def _es_projected_state(u, seed, gas_model, dim):
    """project_fluid_state(..., entropy_stable=True) at a set of points
    (gas_model.py:494-588): the state made from ``u``, its entropy variables
    and back with the state's own ratio, and the state made again from
    those conserved variables (mixtures: both Newton solves seeded with
    ``seed``)."""
    eos = gas_model.eos
    s0 = ph.make_fluid_state(u, gas_model, dim, temperature_seed=seed)
    g0 = eos.gamma(s0.cv, s0.temperature)
    cv_mod = conserved_vars(entropy_vars_state(s0.cv, s0.pressure, g0, dim), g0,
                            dim)
    return ph.make_fluid_state(cv_mod, gas_model, dim, temperature_seed=seed)


def _flat_state(s, shape):
    """A state of ``(F, ...)`` arrays reshaped to ``(F,) + shape``."""
    def r(a):
        return None if a is None else a.reshape(a.shape[:1] + shape) \
            if a.ndim > len(s.temperature.shape) else a.reshape(shape)
    return ph.State(cv=s.cv.reshape(s.cv.shape[:1] + shape), dim=s.dim,
                    temperature=s.temperature.reshape(shape),
                    pressure=s.pressure.reshape(shape),
                    speed_of_sound=s.speed_of_sound.reshape(shape),
                    species_enthalpies=(None if s.species_enthalpies is None else
                                        s.species_enthalpies.reshape(
                                            s.species_enthalpies.shape[:1] + shape)))


def esdg_operator(disc, box, cv, gas_model, boundaries, time=0.0,
                  dissipation=True, degree=None, temperature=None,
                  ec_flux=None, es_volume=True):
    """Entropy-stable Euler RHS ``(F, K, Np)`` through fluid states
    (mirgecom/euler.py:108-310 step by step), for the ideal gas and, with
    the nodal ``temperature (K, Np)``, the uiuc_7sp mixture:

    * volume points: ``project_fluid_state(entropy_stable=True)`` (with
      ``es_volume=False`` the plain projected state, which is what
      ns_operator hands its inviscid operator, mirgecom/navierstokes.py:
      564-570); entropy variables with the state's pressure and ratio
      gamma_q, L2-projected;
    * hybridized points: the projected variables and the nodal ratio
      ``gamma(cv, T)`` interpolated, ``entropy_to_conservative_vars``, the
      state made there (Newton seeded with the interpolated nodal T);
    * interior faces: the neighbour's projected variables at the face
      points turned into conserved variables with this element's
      interpolated ratio, seeded with the neighbour's nodal T;
    * boundary faces: ``project_fluid_state(entropy_stable=True)`` of the
      interior and the condition's exterior state from it;
    * two-point flux ``ec_flux(sl, sr)``: Renac's for mixtures, else
      Chandrashekar's; faces ``F_S . n - alpha/2 (q+ - q-)`` (Rusanov,
      ``alpha`` the larger of ``|u| + c``) or no penalty.
    """
    el = disc.el
    dim, nf, nfp, K = disc.dim, disc.nf, el.nfp, disc.K
    degree = 2 * el.order if degree is None else degree
    eos = gas_model.eos
    mix = temperature is not None
    if ec_flux is None:
        if mix:
            def ec_flux(a, b):
                return renac(a, b, eos, dim)
        else:
            def ec_flux(a, b):
                return chandrashekar(a.cv, b.cv, eos.gamma(), dim)
    ho = _HybridOps(el, degree)
    nq, nfq, nft = ho.nq, ho.nfq, nf * ho.nfq
    nh = nq + nft
    nvar = cv.shape[0]

    def interp(mat, a):           # (n, Np) x (..., K, Np) -> (..., K, n)
        return np.einsum("hn,...kn->...kh", mat, a)

    seed = (lambda a: interp(a, temperature).reshape(-1)) if mix else \
        (lambda a: None)
    # volume points: projected entropy variables
    if es_volume:
        sq = _es_projected_state(interp(ho.vq, cv).reshape(nvar, -1),
                                 seed(ho.vq), gas_model, dim)
    else:
        sq = ph.make_fluid_state(interp(ho.vq, cv).reshape(nvar, -1), gas_model,
                                 dim, temperature_seed=seed(ho.vq))
    gq = eos.gamma(sq.cv, sq.temperature)
    ev_q = entropy_vars_state(sq.cv, sq.pressure, gq, dim).reshape(nvar, K, nq)
    ev_nodal = np.einsum("nq,fkq->fkn", ho.pq, ev_q)
    # hybridized points
    g_base = eos.gamma(cv, temperature) * np.ones((K, el.np))
    gh = interp(ho.vh, g_base)                                    # (K, nh)
    uh = conserved_vars(interp(ho.vh, ev_nodal), gh, dim)
    sh = ph.make_fluid_state(uh.reshape(nvar, -1), gas_model, dim,
                             temperature_seed=seed(ho.vh))
    sh = _flat_state(sh, (K, nh))
    # volume flux differencing, element by element
    vol = np.zeros((nvar, K, el.np))
    sj = disc.sJ

    def pick(s, k, sl):
        """The state at ``[k] + sl`` of its ``(K, nh)`` point axes."""
        idx = (k,) + (sl if isinstance(sl, tuple) else (sl,))
        full = (slice(None),) + idx
        return ph.State(cv=s.cv[full], dim=dim,
                        temperature=s.temperature[idx],
                        pressure=s.pressure[idx],
                        speed_of_sound=s.speed_of_sound[idx],
                        species_enthalpies=(None if s.species_enthalpies is None
                                            else s.species_enthalpies[full]))

    for k in range(K):
        a = pick(sh, k, (slice(None), None))
        b = pick(sh, k, (None, slice(None)))
        fmat = ec_flux(a, b)                                      # (F, d, h, j)
        for d in range(dim):
            qx = disc.J[k] * np.einsum("r,rij->ij", disc.drdx[k][:, d], ho.qr)
            bx = np.repeat(disc.normal[k, :, d] * sj[k], nfq) * np.tile(ho.wf, nf)
            qh = np.zeros((nh, nh))
            qh[:nq, :nq] = qx - 0.5 * ho.e.T @ (bx[:, None] * ho.e)
            qh[:nq, nq:] = 0.5 * ho.e.T * bx[None, :]
            qh[nq:, :nq] = -0.5 * bx[:, None] * ho.e
            qh[nq:, nq:] = 0.5 * np.diag(bx)
            skew = qh - qh.T
            vol[:, k] += np.einsum("hn,fhj->fn", ho.vh, skew[None] * fmat[:, d])
    # faces: own states from the hybridized face points
    sm = pick(sh, slice(None), slice(nq, None))                   # (.., K, nft)
    sm = _flat_state(sm, (K, nf, nfq))
    # boundary faces (no neighbour) take their own values here; their flux
    # is replaced below
    ev_ext = disc.exterior(ev_nodal)                              # (F, K, Nf, Nfp)
    ev_ext = np.where(np.isnan(ev_ext), disc.trace(ev_nodal), ev_ext)
    t_ext = None
    if mix:
        t_ext = disc.exterior(temperature[None])[0]
        t_ext = np.where(np.isnan(t_ext), disc.trace(temperature[None])[0], t_ext)
    evp = np.zeros((nvar, K, nf, nfq))
    tp = np.zeros((K, nf, nfq))
    for f in range(nf):
        tb = ho.vf[f * nfq:(f + 1) * nfq][:, el.face_mask[f]]    # (nfq, Nfp)
        evp[:, :, f] = np.einsum("mj,fkj->fkm", tb, ev_ext[:, :, f])
        if mix:
            tp[:, f] = np.einsum("mj,kj->km", tb, t_ext[:, f])
    gface = gh[:, nq:].reshape(K, nf, nfq)
    up = conserved_vars(evp, gface, dim)
    sp = ph.make_fluid_state(up.reshape(nvar, -1), gas_model, dim,
                             temperature_seed=tp.reshape(-1) if mix else None)
    sp = _flat_state(sp, (K, nf, nfq))
    normal = np.moveaxis(np.repeat(disc.normal[:, :, None, :], nfq, axis=2),
                         -1, 0)                                   # (d, K, Nf, q)

    def face_flux(a, b, n):
        f = np.einsum("fd...,d...->f...", ec_flux(a, b), n)
        if dissipation:
            alpha = np.maximum(a.wavespeed, b.wavespeed)
            f = f - 0.5 * alpha * (b.cv - a.cv)
        return f

    fstar = face_flux(sm, sp, normal)
    if boundaries:
        trace_q = interp(ho.vf, cv).reshape(nvar, K, nf, nfq)
        trace_t = interp(ho.vf, temperature).reshape(K, nf, nfq) if mix else None
        xf, _ = face_rule(dim, degree)
        xf_ref = []
        for fv in FACES[dim]:
            v = REF_VERTS[dim][list(fv)]
            xf_ref.append(v[0] + 0.5 * (1 + xf) @ (v[1:] - v[0]))
        xphys = (disc.v[:, None, None, 0, :]
                 + np.einsum("fmi,kid->kfmd", (1 + np.array(xf_ref)) / 2,
                             disc.v[:, 1:, :] - disc.v[:, :1, :]))
        for bdry, idx in _boundary_groups(disc, box, boundaries):
            faces = np.unique(idx // nfp)
            kk, ff = faces // nf, faces % nf
            nb = len(faces)
            smb = _es_projected_state(
                trace_q[:, kk, ff].reshape(nvar, -1), gas_model=gas_model,
                seed=trace_t[kk, ff].reshape(-1) if mix else None, dim=dim)
            if bdry.kind == "dummy":
                spb = smb
            elif bdry.kind == "prescribed":
                spb = bdry.state_func(np.moveaxis(xphys[kk, ff], -1, 0)
                                      .reshape(dim, -1), smb, time)
            else:
                raise NotImplementedError("the entropy-stable oracle covers "
                                          "prescribed and dummy boundaries")
            n = normal[:, kk, ff].reshape(dim, -1)
            fstar[:, kk, ff] = face_flux(smb, spb, n).reshape(nvar, nb, nfq)
    face = np.einsum("hn,fkh->fkn", ho.vf,
                     (fstar * (sj[:, :, None] * ho.wf[None, None, :])[None])
                     .reshape(nvar, K, nft))
    return -np.einsum("ij,fkj->fki", ho.minv, vol + face) / disc.J[None, :, None]

# }}}
