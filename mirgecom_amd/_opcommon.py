"""Shared host-side preparation of the GPU operator calls."""

from __future__ import annotations

from warnings import warn

import torch

from mirgecom_amd import _lib
from mirgecom_amd.boundary import MengaldoBoundaryCondition
from mirgecom_amd.device import get_device_discretization
from mirgecom_amd.dof_desc import (
    BTAG_ALL, DISCR_TAG_BASE, DISCR_TAG_QUAD, BoundaryDomainTag, DOFDesc,
    VolumeDomainTag)
from mirgecom_amd.fluid import ConservedVars
from mirgecom_amd.gas_model import FluidState, _LazyDV
from mirgecom_amd.utils import normalize_boundaries


def check_dd(dd):
    """mirgecom/euler.py:369-372"""
    if not isinstance(dd.domain_tag, VolumeDomainTag):
        raise TypeError("dd must represent a volume")
    if dd.discretization_tag != DISCR_TAG_BASE:
        raise ValueError("dd must belong to the base discretization")


def check_unsupported(quadrature_tag, use_esdg, operator_states_quad,
                      limiter_func, entropy_min, comm_tag=None,
                      allow_quad=False, state=None, boundaries=None):
    if use_esdg and not (allow_quad and quadrature_tag == DISCR_TAG_QUAD):
        raise NotImplementedError("the entropy-stable (ESDG) operator is built "
                                  "with quadrature_tag=DISCR_TAG_QUAD")
    if quadrature_tag not in ((DISCR_TAG_BASE, DISCR_TAG_QUAD) if allow_quad
                              else (DISCR_TAG_BASE,)):
        raise NotImplementedError(f"quadrature_tag {quadrature_tag!r} is not "
                                  "supported here")
    if operator_states_quad is not None:
        # the kernels form the face states themselves: accept the states of
        # make_operator_fluid_states made for this very call
        from mirgecom_amd.gas_model import OperatorFluidStates
        if not isinstance(operator_states_quad, OperatorFluidStates):
            raise NotImplementedError(
                "operator_states_quad must come from mirgecom_amd.gas_model."
                "make_operator_fluid_states")
        operator_states_quad.check(state, quadrature_tag, boundaries)
    if limiter_func is not None or entropy_min is not None:
        raise NotImplementedError("limiters are not part of the device path")


class Prepared:
    """Device pointers and boundary exterior states for one operator call."""

    def __init__(self, dcoll, state, gas_model, boundaries, time):
        cv = state.cv
        q = cv.q
        if not q.is_cuda:
            raise RuntimeError("the GPU operators need the state on a HIP "
                               "device")
        if q.dtype != torch.float64:
            raise TypeError("the state must be float64")
        self.dev = get_device_discretization(dcoll, q.device)
        dev = self.dev
        host = dcoll
        K, npe = host.nelements, host.refel.nnodes
        if tuple(q.shape[1:]) != (K, npe):
            raise ValueError(f"state shape {tuple(q.shape)} does not match the "
                             f"discretization ({K}, {npe})")
        self.q = q.contiguous()
        self.dim = cv.dim
        self.nspecies = cv.nspecies
        self.mix = bool(getattr(gas_model.eos, "is_mixture", False))
        self.temp = None
        if self.mix:
            self.temp = state.temperature.contiguous()
        self.gas = gas_model.device_params(cv.nspecies)
        boundaries = normalize_boundaries(boundaries)
        nb = host.nbnd_nodes
        self.qb = None
        self.tb = None
        self.wall = None
        if nb > 0:
            self.qb = torch.empty((q.shape[0], nb), dtype=q.dtype,
                                  device=q.device)
            if self.mix:
                self.tb = torch.empty(nb, dtype=q.dtype, device=q.device)
            walls = any(isinstance(b, MengaldoBoundaryCondition)
                        for b in boundaries.values())
            if walls:
                # Mengaldo data: CV_bc, T_bc, T(state_bc), flags (mg_wall)
                self._wq = torch.empty_like(self.qb)
                self._wt = torch.empty(nb, dtype=q.dtype, device=q.device)
                self._wts = (torch.empty(nb, dtype=q.dtype, device=q.device)
                             if self.mix else None)
                self._wflag = torch.zeros(nb, dtype=torch.int32,
                                          device=q.device)
            covered = torch.zeros(nb, dtype=torch.int32, device=q.device)
            flat_q = self.q.reshape(q.shape[0], -1)
            for btag, bdry in boundaries.items():
                tag = btag.tag
                idx = dev.boundary_index(tag)
                if idx.numel() == 0:
                    continue
                covered[idx] += 1
                vidx = dev.bnd_vol_index[idx]
                cv_m = ConservedVars(self.dim, flat_q[:, vidx])
                t_m = (self.temp.reshape(-1)[vidx] if self.mix else None)
                state_m = FluidState(cv_m, _LazyDV(cv_m, gas_model.eos, t_m),
                                     gas_model)
                dd_bdry = DOFDesc(BoundaryDomainTag(tag), DISCR_TAG_BASE)
                ext = bdry.exterior_state(dcoll, dd_bdry, gas_model, state_m,
                                          time=time)
                self.qb[:, idx] = ext.cv.q.to(q.dtype)
                if self.mix:
                    self.tb[idx] = ext.temperature
                if walls:
                    if isinstance(bdry, MengaldoBoundaryCondition):
                        sbc = bdry.state_bc(dcoll, dd_bdry, gas_model, state_m,
                                            time=time)
                        self._wq[:, idx] = sbc.cv.q
                        self._wt[idx] = bdry.temperature_bc(dcoll, dd_bdry,
                                                            state_m, time=time)
                        if self.mix:
                            self._wts[idx] = sbc.temperature
                        self._wflag[idx] = bdry.device_flags
                    else:
                        self._wq[:, idx] = ext.cv.q
                        self._wt[idx] = ext.temperature
                        if self.mix:
                            self._wts[idx] = ext.temperature
            # the coverage depends only on the tags: check it once per tag
            # set (the check synchronises, which a captured call must not)
            key = tuple(sorted(b.tag for b in boundaries))
            if key not in dev.checked_tag_sets:
                if int(covered.min()) != 1 or int(covered.max()) != 1:
                    raise ValueError("every boundary face needs exactly one "
                                     "boundary condition")
                dev.checked_tag_sets.add(key)
            if walls:
                self._wwork = torch.empty(((q.shape[0] + 1) * self.dim, nb),
                                          dtype=q.dtype, device=q.device)
                self.wall = _lib.MgWall(
                    q_bc=self._wq.data_ptr(), t_bc=self._wt.data_ptr(),
                    t_state_bc=(self._wts.data_ptr() if self.mix else None),
                    flags=self._wflag.data_ptr(),
                    vol_index=dev.bnd_vol_index32.data_ptr(),
                    normals=dev.bnd_normals.data_ptr(),
                    work=self._wwork.data_ptr())
        elif boundaries:
            unknown = [b.tag for b in boundaries
                       if b.tag != BTAG_ALL
                       and b.tag not in host.boundary_tags()]
            if unknown:
                raise ValueError(f"unknown boundary tags {unknown}")
        self.stream = _lib.stream_handle()


def warn_mixture_limiter(state, limiter_func):
    if state.is_mixture and limiter_func is None:
        warn("Mixtures often require species limiting, and a non-limited "
             "state is being created for this operator. For mixtures, "
             "one should pass the operator_states_quad argument with "
             "limited states or pass a limiter_func to this operator.",
             stacklevel=3)


def local_timestep(dcoll, state, viscous):
    """Node-local stable timestep ``(K, Np)`` from ``mg_local_timestep``."""
    q = state.cv.q
    if not q.is_cuda or q.dtype != torch.float64:
        raise RuntimeError("the GPU timestep needs a float64 state on a HIP "
                           "device")
    dev = get_device_discretization(dcoll, q.device)
    K, npe = dcoll.nelements, dcoll.refel.nnodes
    if tuple(q.shape[1:]) != (K, npe):
        raise ValueError(f"state shape {tuple(q.shape)} does not match the "
                         f"discretization ({K}, {npe})")
    q = q.contiguous()
    gas = state.gas_model.device_params(state.cv.nspecies,
                                        with_transport=viscous)
    temp = None
    if gas.eos_kind == 1:
        temp = state.temperature.contiguous()
    out = torch.empty((K, npe), dtype=q.dtype, device=q.device)
    _lib.call("mg_local_timestep", dev.mesh, gas, _lib.ptr(q), _lib.ptr(temp),
              _lib.ptr(dev.lengthscale), 1 if viscous else 0, _lib.ptr(out),
              _lib.stream_handle())
    return out


def quad_boundary_states(dcoll, dev, q, dim, gas_model, boundaries, time,
                         temperature=None, entropy_stable=False):
    """Exterior states ``(F, Nbf*nfq)`` at the boundary faces' quadrature
    points (mirgecom/euler.py:260-301: the interior state projected to the
    quadrature points of each boundary, ``project_fluid_state``, and the
    boundary's exterior state evaluated there) and, for mixtures (nodal
    ``temperature`` given), the exterior temperatures; ``(None, None)``
    without boundary faces.  Mixture interior states get their temperature
    by Newton from the interpolated nodal temperature; with
    ``entropy_stable`` (mixtures) the interior state is first taken through
    its entropy variables and back (``project_fluid_state(...,
    entropy_stable=True)``, gas_model.py:494-588)."""
    boundaries = normalize_boundaries(boundaries)
    nb = dcoll.nbnd_nodes
    if nb == 0:
        if boundaries:
            unknown = [b.tag for b in boundaries if b.tag not in (BTAG_ALL,)]
            if unknown:
                raise ValueError(f"boundary tags {unknown} are not on the mesh")
        return None, None
    nf = q.shape[0]
    interior = dev.to_bnd_quad(q.reshape(nf, -1)[:, dev.bnd_vol_index])
    t_int = None
    tbq = None
    if temperature is not None:
        seed = dev.to_bnd_quad(temperature.reshape(-1)[dev.bnd_vol_index])
        t_int = gas_model.eos.temperature(ConservedVars(dim, interior),
                                          temperature_seed=seed)
        if entropy_stable:
            from mirgecom_amd.gas_model import (
                conservative_to_entropy_vars, entropy_to_conservative_vars)
            cv0 = ConservedVars(dim, interior)
            s0 = FluidState(cv0, _LazyDV(cv0, gas_model.eos, t_int), gas_model)
            g0 = gas_model.eos.gamma(cv0, t_int)
            interior = entropy_to_conservative_vars(
                g0, conservative_to_entropy_vars(g0, s0)).q
            t_int = gas_model.eos.temperature(ConservedVars(dim, interior),
                                              temperature_seed=seed)
        tbq = torch.empty_like(t_int)
    qbq = torch.empty_like(interior)
    covered = torch.zeros(interior.shape[1], dtype=torch.int32, device=q.device)
    for btag, bdry in boundaries.items():
        idx = dev.boundary_index(btag.tag, True)
        if idx.numel() == 0:
            continue
        covered[idx] += 1
        cv_m = ConservedVars(dim, interior[:, idx])
        t_m = t_int[idx] if t_int is not None else None
        state_m = FluidState(cv_m, _LazyDV(cv_m, gas_model.eos, t_m), gas_model)
        dd_bdry = DOFDesc(BoundaryDomainTag(btag.tag), DISCR_TAG_QUAD)
        ext = bdry.exterior_state(dcoll, dd_bdry, gas_model, state_m, time=time)
        qbq[:, idx] = ext.cv.q.to(q.dtype)
        if tbq is not None:
            tbq[idx] = ext.temperature
    # the coverage depends only on the tags: check it once per tag set (the
    # check synchronises, which a captured call must not)
    key = ("quad",) + tuple(sorted(b.tag for b in boundaries))
    if key not in dev.checked_tag_sets:
        if int(covered.min()) != 1 or int(covered.max()) != 1:
            raise ValueError("every boundary face needs exactly one boundary "
                             "condition")
        dev.checked_tag_sets.add(key)
    return qbq, tbq
