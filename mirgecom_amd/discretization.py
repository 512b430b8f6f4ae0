"""Box meshes of simplices, affine geometric factors and face connectivity.

Host-side, one-time FP64 setup that stands in for the reference's
``get_box_mesh`` (mirgecom/simutil.py:141-200, meshmode's
``generate_regular_rect_mesh``) and ``create_discretization_collection``
(mirgecom/discretization.py:41-83).  Everything built here is uploaded once to
HBM and read by the HIP kernels of the right-hand side.

Face connectivity is combinatorial and exact: a face node is identified by the
(periodically identified) global vertex ids of its face together with its
integer barycentric lattice coordinates with respect to those vertices.  Two
face nodes of neighbouring elements match iff these keys are equal, so the
int32 neighbour map ``vmap_p`` is built without any floating-point
comparison.

Layouts (K elements, Np nodes/element, Nf faces, Nfp nodes/face):

* nodal fields: ``(K, Np)`` row-major per field (the DOFArray layout of
  the reference, doc/support/arrays.rst:96-103)
* ``rx``: ``(K, dim, dim)``, ``rx[e, a, i] = d r_a / d x_i``
* ``normals``: ``(K, Nf, dim)`` outward unit normals (constant per face)
* ``face_jac``: ``(K, Nf)`` physical face measure / reference face measure
* ``fscale``: ``(K, Nf)`` ``face_jac / jac``
* ``vmap_p``: ``(K, Nf, Nfp)`` int32: global node index ``e'*Np + i'`` of the
  matching node on the neighbour, or ``-(b+1)`` for boundary node ``b``
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from mirgecom_amd.refelem import (
    FACE_VERTICES, REF_VERTICES, ReferenceElement, get_reference_element)


BTAG_ALL = "all"


@dataclass
class BoxMesh:
    """A simplicial mesh of an axis-aligned box.

    * ``vertices``: ``(nv, dim)``
    * ``elements``: ``(K, dim+1)`` vertex indices, positively oriented
    * ``vertex_pid``: ``(nv,)`` vertex id after periodic identification
    * ``a``, ``b``, ``n``, ``periodic``: the box description
    """

    dim: int
    vertices: np.ndarray
    elements: np.ndarray
    vertex_pid: np.ndarray
    a: tuple
    b: tuple
    n: tuple
    periodic: tuple

    @property
    def nelements(self):
        return self.elements.shape[0]


def _as_tuple(x, dim):
    return tuple(x) if np.ndim(x) else (x,) * dim


def generate_box_mesh(dim, a, b, n, periodic=None) -> BoxMesh:
    """Box ``[a, b]`` cut into ``n`` cells per axis; each square is split into
    2 triangles, each cube into 6 tetrahedra (Kuhn triangulation, all cells
    cut the same way so that the mesh is conforming, also across periodic
    boundaries).  Boundary faces are tagged ``"-i"``/``"+i"`` (``i`` 1-based
    axis), as in ``get_box_mesh`` (mirgecom/simutil.py:185-189)."""
    a = tuple(float(v) for v in _as_tuple(a, dim))
    b = tuple(float(v) for v in _as_tuple(b, dim))
    n = tuple(int(v) for v in _as_tuple(n, dim))
    periodic = (False,) * dim if periodic is None else tuple(
        bool(v) for v in _as_tuple(periodic, dim))
    for i in range(dim):
        # with 2 cells, two distinct faces can share all periodic vertex ids
        if periodic[i] and n[i] < 3:
            raise ValueError("a periodic axis needs at least 3 cells")
    axes = [np.linspace(a[i], b[i], n[i] + 1) for i in range(dim)]
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
    vertices = grid.reshape(-1, dim)
    shape = tuple(ni + 1 for ni in n)
    ids = np.arange(vertices.shape[0]).reshape(shape)
    pid_idx = np.meshgrid(*[np.arange(ni + 1) % ni if periodic[i]
                            else np.arange(ni + 1) for i, ni in enumerate(n)],
                          indexing="ij")
    vertex_pid = np.ravel_multi_index(
        [p.ravel() for p in pid_idx], shape).astype(np.int64)

    cells = np.stack(np.meshgrid(*[np.arange(ni) for ni in n], indexing="ij"),
                     axis=-1).reshape(-1, dim)
    elems = []
    if dim == 2:
        v00 = ids[cells[:, 0], cells[:, 1]]
        v10 = ids[cells[:, 0] + 1, cells[:, 1]]
        v01 = ids[cells[:, 0], cells[:, 1] + 1]
        v11 = ids[cells[:, 0] + 1, cells[:, 1] + 1]
        elems.append(np.stack([v00, v10, v11], axis=1))
        elems.append(np.stack([v00, v11, v01], axis=1))
        elements = np.stack(elems, axis=1).reshape(-1, 3)
    elif dim == 3:
        import itertools
        for perm in itertools.permutations(range(3)):
            path = [cells.copy()]
            cur = cells.copy()
            for ax in perm:
                cur = cur.copy()
                cur[:, ax] += 1
                path.append(cur)
            elems.append(np.stack(
                [ids[p[:, 0], p[:, 1], p[:, 2]] for p in path], axis=1))
        elements = np.stack(elems, axis=1).reshape(-1, 4)
    else:
        raise ValueError("dim must be 2 or 3")

    # positive orientation
    v = vertices[elements]
    jac = (v[:, 1:] - v[:, :1])
    det = np.linalg.det(jac)
    neg = det < 0
    elements[neg, 1], elements[neg, 2] = (
        elements[neg, 2].copy(), elements[neg, 1].copy())
    return BoxMesh(dim=dim, vertices=vertices, elements=elements.astype(np.int64),
                   vertex_pid=vertex_pid, a=a, b=b, n=n, periodic=periodic)


def get_box_mesh(dim, a, b, n, t=None, periodic=None):
    """Call surface of ``mirgecom.simutil.get_box_mesh``
    (mirgecom/simutil.py:141-200); ``t`` (mesh type) is accepted for
    signature compatibility and ignored."""
    return generate_box_mesh(dim, a, b, n, periodic=periodic)


@dataclass
class Discretization:
    """The discretization collection: reference element, mesh, geometric
    factors, connectivity and node coordinates, all FP64/int32 on the host.
    ``to_device`` uploads them to HBM."""

    mesh: BoxMesh
    refel: ReferenceElement
    host_nodes: np.ndarray     # (dim, K, Np)
    rx: np.ndarray             # (K, dim, dim)
    jac: np.ndarray            # (K,)
    normals: np.ndarray        # (K, Nf, dim)
    face_jac: np.ndarray       # (K, Nf)
    fscale: np.ndarray         # (K, Nf)
    vmap_p: np.ndarray         # (K, Nf, Nfp) int32
    bnd_faces: np.ndarray      # (Nbf, 2) [element, face]
    bnd_face_tags: list        # tag of each boundary face
    bnd_nodes: np.ndarray      # (dim, Nbf*Nfp)
    bnd_normals: np.ndarray    # (dim, Nbf*Nfp)
    quadrature_order: int = 0  # degree of DISCR_TAG_QUAD (2p by default)
    # order in which the device kernels take element blocks: "morton"
    # (element_traversal_rank) or "natural" (element index order)
    traversal: str = "morton"
    _device_cache: dict = field(default_factory=dict, repr=False)
    _host_cache: dict = field(default_factory=dict, repr=False)

    @property
    def dim(self):
        return self.mesh.dim

    @property
    def order(self):
        return self.refel.order

    @property
    def nelements(self):
        return self.mesh.nelements

    @property
    def nnodes(self):
        return self.nelements * self.refel.nnodes

    @property
    def lengthscales(self):
        """Element characteristic lengths ``(K,)`` (see
        ``characteristic_lengthscales``)."""
        if "h" not in self._host_cache:
            self._host_cache["h"] = characteristic_lengthscales(self)
        return self._host_cache["h"]

    @property
    def nbnd_nodes(self):
        return self.bnd_nodes.shape[1]

    @property
    def quadrature(self):
        """Element operators of the quadrature discretization
        (``refelem.QuadratureOperators`` of degree ``quadrature_order``)."""
        from mirgecom_amd.refelem import get_quadrature_operators
        return get_quadrature_operators(self.dim, self.order,
                                        self.quadrature_order)

    def quad_boundary_geometry(self):
        """Coordinates and unit normals ``(dim, Nbf*nfq)`` of the boundary
        faces' quadrature points, face by face in the order of the boundary
        nodes."""
        if "bq" not in self._host_cache:
            qo, re = self.quadrature, self.refel
            bf = self.bnd_faces
            if len(bf) == 0:
                z = np.zeros((self.dim, 0))
                self._host_cache["bq"] = (z, z)
            else:
                xf = self.host_nodes[:, bf[:, 0:1], re.face_mask[bf[:, 1]]]
                xq = np.einsum("bmj,dbj->dbm", qo.face_interp[bf[:, 1]], xf)
                nq = np.repeat(self.normals[bf[:, 0], bf[:, 1]][:, None, :],
                               qo.nfq, axis=1)
                self._host_cache["bq"] = (
                    xq.reshape(self.dim, -1),
                    np.moveaxis(nq, -1, 0).reshape(self.dim, -1))
        return self._host_cache["bq"]

    def boundary_tags(self):
        return sorted(set(self.bnd_face_tags))

    def nodes(self, dd=None):
        """Node coordinates on the device: ``(dim, K, Np)`` for the volume,
        ``(dim, n)`` for a boundary ``DOFDesc`` (the boundary nodes of that
        tag, in the order the boundary callbacks receive states)."""
        from mirgecom_amd.device import get_device_discretization
        dev = get_device_discretization(self)
        tag = _boundary_tag_of(dd)
        quad = _is_quad(dd)
        if tag is None:
            return dev.quad_nodes() if quad else dev.nodes
        if quad:
            return dev.bnd_quad_geometry()[0][:, dev.boundary_index(tag, True)]
        return dev.bnd_nodes[:, dev.boundary_index(tag)]

    def normal(self, dd):
        """Outward unit normals ``(dim, n)`` on the nodes of boundary ``dd``."""
        from mirgecom_amd.device import get_device_discretization
        dev = get_device_discretization(self)
        tag = _boundary_tag_of(dd)
        if _is_quad(dd):
            return dev.bnd_quad_geometry()[1][:, dev.boundary_index(tag, True)]
        return dev.bnd_normals[:, dev.boundary_index(tag)]

    def discr_from_dd(self, dd):
        """Object with a ``nodes()`` method (grudge's ``discr_from_dd``)."""
        dcoll = self

        class _D:
            def nodes(self):
                return dcoll.nodes(dd)
        return _D()

    def boundary_node_slice_mask(self, tag, quad=False):
        """Boolean mask over boundary nodes (or boundary quadrature points)
        for a boundary tag."""
        nfp = self.quadrature.nfq if quad else self.refel.nface_nodes
        if tag == BTAG_ALL:
            fm = np.ones(len(self.bnd_face_tags), dtype=bool)
        else:
            fm = np.array([t == tag for t in self.bnd_face_tags], dtype=bool)
        return np.repeat(fm, nfp)


def _is_quad(dd):
    from mirgecom_amd.dof_desc import DISCR_TAG_QUAD
    return getattr(dd, "discretization_tag", None) == DISCR_TAG_QUAD


def _boundary_tag_of(dd):
    if dd is None:
        return None
    dt = getattr(dd, "domain_tag", dd)
    if type(dt).__name__ == "VolumeDomainTag":
        return None
    return getattr(dt, "tag", dt)


def _geometry(mesh: BoxMesh, refel: ReferenceElement):
    dim = mesh.dim
    v = mesh.vertices[mesh.elements]             # (K, d+1, d)
    dxdr = 0.5 * (v[:, 1:] - v[:, :1])           # (K, a, i) = d x_i / d r_a
    dxdr = np.transpose(dxdr, (0, 2, 1))         # (K, i, a)
    jac = np.linalg.det(dxdr)
    if np.any(jac <= 0):
        raise ValueError("mesh contains inverted elements")
    rx = np.linalg.inv(dxdr)                     # (K, a, i) = d r_a / d x_i

    # nodes: x = v0 + sum_a (v_{a+1} - v0) (1 + r_a)/2
    r = refel.nodes                              # (Np, d)
    nodes = v[:, None, 0, :] + np.einsum(
        "kad,na->knd", v[:, 1:] - v[:, :1], 0.5 * (1 + r))
    nodes = np.transpose(nodes, (2, 0, 1)).copy()   # (d, K, Np)

    # gradients of barycentric coordinates w.r.t. reference coordinates
    gbar = np.zeros((dim + 1, dim))
    gbar[0, :] = -0.5
    for a in range(dim):
        gbar[a + 1, a] = 0.5
    nf = dim + 1
    normals = np.zeros((mesh.nelements, nf, dim))
    face_jac = np.zeros((mesh.nelements, nf))
    for f, fverts in enumerate(FACE_VERTICES[dim]):
        opp = [i for i in range(dim + 1) if i not in fverts][0]
        grad_l = np.einsum("a,kai->ki", gbar[opp], rx)
        nrm = -grad_l / np.linalg.norm(grad_l, axis=1)[:, None]
        normals[:, f] = nrm
        fv = v[:, list(fverts)]
        if dim == 2:
            meas = np.linalg.norm(fv[:, 1] - fv[:, 0], axis=1)
            ref_meas = 2.0
        else:
            meas = 0.5 * np.linalg.norm(
                np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0]), axis=1)
            ref_meas = 2.0
        face_jac[:, f] = meas / ref_meas
    jac_ref = jac  # d x / d r determinant: physical volume / reference volume
    fscale = face_jac / jac_ref[:, None]
    return nodes, rx, jac_ref, normals, face_jac, fscale


def _face_node_keys(mesh: BoxMesh, refel: ReferenceElement):
    """Integer keys identifying every face node, ``(K*Nf*Nfp, 2*dim)``:
    the face's periodic vertex ids (sorted) followed by the node's
    barycentric lattice coordinates in that vertex order."""
    dim = mesh.dim
    order = refel.order
    tup = refel.tuples
    bary = np.concatenate([(order - tup.sum(axis=1))[:, None], tup], axis=1)
    pid = mesh.vertex_pid[mesh.elements]          # (K, d+1)
    K = mesh.nelements
    nfp = refel.nface_nodes
    keys = np.zeros((K, dim + 1, nfp, 2 * dim), dtype=np.int64)
    for f, fverts in enumerate(FACE_VERTICES[dim]):
        fv = np.array(fverts)
        gids = pid[:, fv]                         # (K, dim)
        b = bary[refel.face_mask[f]][:, fv]       # (Nfp, dim)
        order_idx = np.argsort(gids, axis=1, kind="stable")
        sg = np.take_along_axis(gids, order_idx, axis=1)
        sb = b[None, :, :][np.zeros(K, dtype=int)]   # (K, Nfp, dim)
        sb = np.take_along_axis(sb, order_idx[:, None, :].repeat(nfp, 1), axis=2)
        keys[:, f, :, :dim] = sg[:, None, :]
        keys[:, f, :, dim:] = sb
    return keys.reshape(-1, 2 * dim)


def _boundary_tag(mesh: BoxMesh, face_centroid):
    dim = mesh.dim
    scale = max(mesh.b[i] - mesh.a[i] for i in range(dim))
    tol = 1e-9 * scale
    for i in range(dim):
        if abs(face_centroid[i] - mesh.a[i]) < tol:
            return f"-{i + 1}"
        if abs(face_centroid[i] - mesh.b[i]) < tol:
            return f"+{i + 1}"
    raise ValueError("boundary face not on the box boundary")


def _connectivity(mesh: BoxMesh, refel: ReferenceElement, nodes, normals):
    dim = mesh.dim
    K = mesh.nelements
    nf = dim + 1
    nfp = refel.nface_nodes
    np_ = refel.nnodes
    keys = _face_node_keys(mesh, refel)
    n_fn = keys.shape[0]
    order = np.lexsort(keys.T[::-1])
    sk = keys[order]
    same_next = np.all(sk[1:] == sk[:-1], axis=1)
    if np.any(same_next[1:] & same_next[:-1]):
        raise ValueError("a face node key occurs more than twice")
    vol_idx = (np.arange(K)[:, None, None] * np_
               + refel.face_mask[None, :, :]).reshape(-1)
    vmap_p = np.full(n_fn, np.iinfo(np.int32).min, dtype=np.int64)
    pairs = np.nonzero(same_next)[0]
    i0 = order[pairs]
    i1 = order[pairs + 1]
    vmap_p[i0] = vol_idx[i1]
    vmap_p[i1] = vol_idx[i0]

    # boundary faces: every face whose nodes found no partner
    unmatched = vmap_p == np.iinfo(np.int32).min
    face_unmatched = unmatched.reshape(K, nf, nfp)
    if np.any(face_unmatched.any(axis=2) != face_unmatched.all(axis=2)):
        raise ValueError("partially matched face")
    bf = np.argwhere(face_unmatched.all(axis=2))      # (Nbf, 2), sorted
    vmap_p = vmap_p.reshape(K, nf, nfp)
    nbf = bf.shape[0]
    bnodes = np.zeros((dim, nbf * nfp))
    bnormals = np.zeros((dim, nbf * nfp))
    tags = []
    for ib, (e, f) in enumerate(bf):
        vmap_p[e, f, :] = -(ib * nfp + np.arange(nfp) + 1)
        idx = refel.face_mask[f]
        bnodes[:, ib * nfp:(ib + 1) * nfp] = nodes[:, e, idx]
        bnormals[:, ib * nfp:(ib + 1) * nfp] = normals[e, f][:, None]
        tags.append(_boundary_tag(mesh, nodes[:, e, idx].mean(axis=1)))
    if np.abs(vmap_p).max() >= 2**31:
        raise ValueError("mesh too large for int32 node indices")
    return vmap_p.astype(np.int32), bf.astype(np.int64), tags, bnodes, bnormals


def create_discretization_collection(mesh: BoxMesh, order: int, *,
                                     quadrature_order: int = -1) -> Discretization:
    """Build the nodal DG discretization of ``mesh`` at polynomial ``order``
    (call surface of mirgecom/discretization.py:41-83); the quadrature
    discretization (DISCR_TAG_QUAD) integrates to degree ``quadrature_order``,
    ``2 order`` by default as in the reference."""
    refel = get_reference_element(mesh.dim, order)
    nodes, rx, jac, normals, face_jac, fscale = _geometry(mesh, refel)
    vmap_p, bf, tags, bnodes, bnormals = _connectivity(mesh, refel, nodes, normals)
    return Discretization(
        mesh=mesh, refel=refel, host_nodes=nodes, rx=rx, jac=jac, normals=normals,
        face_jac=face_jac, fscale=fscale, vmap_p=vmap_p, bnd_faces=bf,
        bnd_face_tags=tags, bnd_nodes=bnodes, bnd_normals=bnormals,
        quadrature_order=2 * order if quadrature_order < 0 else int(quadrature_order))


def characteristic_lengthscales(dcoll: Discretization):
    """Per-element characteristic length ``h_e = min_{i != j} |r_i - r_j| * r_D``
    (shape ``(K,)``): the smallest distance between two distinct nodes of the
    biunit reference simplex times the in-radius ``r_D = d V / S`` of the
    element (volume ``V``, total face measure ``S``), the estimate of
    Hesthaven & Warburton, Nodal DG Methods, sec. 6.4, that the reference takes
    from grudge's ``characteristic_lengthscales`` (mirgecom/inviscid.py:415-421,
    mirgecom/viscous.py:594-607)."""
    refel = dcoll.refel
    dim = dcoll.dim
    vol = dcoll.jac * refel.ref_volume
    # every face of the biunit simplex has reference measure 2
    face_meas = dcoll.face_jac.sum(axis=1) * 2.0
    inradius = dim * vol / face_meas
    x = refel.nodes
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    d[np.eye(len(d), dtype=bool)] = np.inf
    return inradius * d.min()


def element_traversal_rank(mesh: BoxMesh):
    """Position ``(K,)`` int32 of every element in the Z-order (Morton)
    traversal of the element centroids: the centroids are quantised on a
    ``2^b`` lattice over their bounding box (b = 10 bits per axis in 3-D, 15
    in 2-D), the lattice coordinates' bits interleaved, and the elements
    sorted by the interleaved code (ties by element index).  Consecutive
    ranks are mesh neighbours except across the curve's jumps, which is what
    the device kernels' block visiting order needs (it changes no result)."""
    dim = mesh.dim
    c = mesh.vertices[mesh.elements].mean(axis=1)
    lo, hi = c.min(axis=0), c.max(axis=0)
    bits = 10 if dim == 3 else 15
    span = np.where(hi > lo, hi - lo, 1.0)
    q = np.rint((c - lo) / span * ((1 << bits) - 1)).astype(np.int64)
    code = np.zeros(len(c), dtype=np.int64)
    for b in range(bits):
        for d in range(dim):
            code |= ((q[:, d] >> b) & 1) << (dim * b + d)
    order = np.argsort(code, kind="stable")
    rank = np.empty(len(c), dtype=np.int32)
    rank[order] = np.arange(len(c), dtype=np.int32)
    return rank
