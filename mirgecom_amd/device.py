"""HBM-resident copy of a discretization and its ``mg_mesh`` descriptor."""

from __future__ import annotations

import numpy as np
import torch

from mirgecom_amd import _lib
from mirgecom_amd.discretization import BTAG_ALL, Discretization


class DeviceDiscretization:
    """Geometry, connectivity and reference operators uploaded once.

    Besides the kernel inputs it keeps the node coordinates and, for every
    boundary tag, the boundary-node indices and coordinates that boundary
    callbacks need."""

    def __init__(self, dcoll: Discretization, device):
        self.host = dcoll
        self.device = torch.device(device)
        f64 = dict(dtype=torch.float64, device=self.device)
        re = dcoll.refel

        def t64(a):
            return torch.as_tensor(np.ascontiguousarray(a), **f64)

        def t32(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32),
                                   device=self.device)

        self.rx = t64(dcoll.rx)
        self.normals = t64(dcoll.normals)
        self.fscale = t64(dcoll.fscale)
        self.vmap_p = t32(dcoll.vmap_p)
        self.weak_diff = t64(re.weak_diff)
        self.lift = t64(re.lift)
        self.face_mask = t32(re.face_mask)
        self.nodes = t64(dcoll.host_nodes)
        self.lengthscale = t64(dcoll.lengthscales)
        nfp = re.nface_nodes
        bf = dcoll.bnd_faces
        # flat volume node index of every boundary node
        if len(bf):
            vol_idx = (bf[:, 0:1] * re.nnodes + re.face_mask[bf[:, 1]]).reshape(-1)
        else:
            vol_idx = np.zeros(0, dtype=np.int64)
        self.bnd_vol_index = torch.as_tensor(vol_idx, dtype=torch.int64,
                                             device=self.device)
        self.bnd_vol_index32 = torch.as_tensor(vol_idx, dtype=torch.int32,
                                               device=self.device)
        self.bnd_nodes = t64(dcoll.bnd_nodes)
        self.bnd_normals = t64(dcoll.bnd_normals)
        self._tag_index = {}
        self.checked_tag_sets = set()
        if dcoll.traversal == "morton":
            from mirgecom_amd.discretization import element_traversal_rank
            self.elem_rank = t32(element_traversal_rank(dcoll.mesh))
        elif dcoll.traversal == "natural":
            self.elem_rank = None
        else:
            raise ValueError(f"unknown traversal {dcoll.traversal!r}")
        self.mesh = _lib.MgMesh(
            dim=dcoll.dim, order=re.order, np=re.nnodes, nfp=nfp,
            nfaces=re.nfaces, pad0=0, nelements=dcoll.nelements,
            nbnd_nodes=dcoll.nbnd_nodes,
            rx=self.rx.data_ptr(), normals=self.normals.data_ptr(),
            fscale=self.fscale.data_ptr(), vmap_p=self.vmap_p.data_ptr(),
            weak_diff=self.weak_diff.data_ptr(), lift=self.lift.data_ptr(),
            face_mask=self.face_mask.data_ptr(), workspace=None,
            elem_rank=None if self.elem_rank is None else self.elem_rank.data_ptr())
        lib = _lib.load()
        nbytes = lib.mg_mesh_workspace_bytes(self.mesh)
        if nbytes <= 0:
            raise NotImplementedError(
                f"dim={dcoll.dim} order={re.order} has no device kernels")
        self.workspace = torch.empty((nbytes + 7) // 8, **f64)
        _lib.check(lib.mg_mesh_prepare(self.mesh, _lib.ptr(self.workspace),
                                       _lib.stream_handle()))

    def boundary_index(self, tag, quad=False):
        """Indices (into the boundary-node arrays, or the boundary
        quadrature-point arrays) of a boundary tag."""
        key = (tag, quad)
        if key not in self._tag_index:
            mask = self.host.boundary_node_slice_mask(tag, quad)
            self._tag_index[key] = torch.as_tensor(
                np.nonzero(mask)[0], dtype=torch.int64, device=self.device)
        return self._tag_index[key]

    # {{{ quadrature discretization (DISCR_TAG_QUAD)

    def _quad(self):
        if not hasattr(self, "_qcache"):
            qo = self.host.quadrature
            f64 = dict(dtype=torch.float64, device=self.device)
            c = {k: torch.as_tensor(np.ascontiguousarray(getattr(qo, k)), **f64)
                 for k in ("interp", "weak_diff", "face_interp", "lift")}
            # the kernel's layouts: transposed so that lanes read consecutively
            c["interp_t"] = c["interp"].T.contiguous()
            c["weak_diff_t"] = c["weak_diff"].transpose(1, 2).contiguous()
            c["face_interp_t"] = c["face_interp"].transpose(1, 2).contiguous()
            c["lift_t"] = c["lift"].T.contiguous()
            xq, nq = self.host.quad_boundary_geometry()
            c["bnd_x"] = torch.as_tensor(np.ascontiguousarray(xq), **f64)
            c["bnd_n"] = torch.as_tensor(np.ascontiguousarray(nq), **f64)
            bf = self.host.bnd_faces
            c["bnd_face"] = torch.as_tensor(
                bf[:, 1] if len(bf) else np.zeros(0, dtype=np.int64),
                dtype=torch.int64, device=self.device)
            c["ops"] = _lib.MgQuad(
                nq=qo.nq, nfq=qo.nfq, interp=c["interp_t"].data_ptr(),
                weak_diff=c["weak_diff_t"].data_ptr(),
                face_interp=c["face_interp_t"].data_ptr(),
                lift=c["lift_t"].data_ptr())
            self._qcache = c
        return self._qcache

    def esdg_ops(self):
        """The ``mg_esdg`` struct of the hybridized SBP operators."""
        c = self._quad()
        if "esdg" not in c:
            from mirgecom_amd.refelem import get_hybrid_operators
            host = self.host
            ho = get_hybrid_operators(host.dim, host.order, host.quadrature_order)
            f64 = dict(dtype=torch.float64, device=self.device)

            def t(a):
                return torch.as_tensor(np.ascontiguousarray(a), **f64)
            c["esdg_arrays"] = arr = {
                "proj": t(ho.proj.T), "vh": t(ho.vh.T),
                "skew": t(np.transpose(ho.skew, (0, 2, 1))),
                "ext": t(ho.extrap), "ext_col": t(ho.extrap.T),
                "apply": t(ho.apply.T),
                "face_weights": t(host.quadrature.face_weights),
                "skew_diag": t(ho.skew_diag), "ext_diag": t(ho.extrap_diag)}
            c["esdg"] = _lib.MgEsdg(
                nh=ho.nh, pad0=0, **{k: v.data_ptr() for k, v in arr.items()})
        return c["esdg"]

    def quad_ops(self):
        """The ``mg_quad`` struct of the quadrature operators."""
        return self._quad()["ops"]

    def quad_nodes(self):
        """Quadrature point coordinates ``(dim, K, nq)``."""
        return torch.einsum("qn,dkn->dkq", self._quad()["interp"], self.nodes)

    def bnd_quad_geometry(self):
        c = self._quad()
        return c["bnd_x"], c["bnd_n"]

    def to_bnd_quad(self, trace):
        """Boundary-node values ``(..., Nbf*Nfp)`` -> boundary quadrature
        points ``(..., Nbf*nfq)`` (each face's own interpolation)."""
        c = self._quad()
        nfp = self.host.refel.nface_nodes
        lead = trace.shape[:-1]
        t = trace.reshape(lead + (-1, nfp))
        out = torch.einsum("bmj,...bj->...bm", c["face_interp"][c["bnd_face"]], t)
        return out.reshape(lead + (-1,))

    # }}}


def get_device_discretization(dcoll: Discretization, device=None):
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    key = str(torch.device(device))
    if key not in dcoll._device_cache:
        dcoll._device_cache[key] = DeviceDiscretization(dcoll, device)
    return dcoll._device_cache[key]


__all__ = ["DeviceDiscretization", "get_device_discretization", "BTAG_ALL"]
