"""Element traversal rank (discretization.element_traversal_rank) and the
workspace size the visiting-order tables add (host-only checks)."""

import ctypes

import numpy as np
import pytest

from mirgecom_amd.discretization import element_traversal_rank, generate_box_mesh

EPBS = [4, 5, 6, 7, 8, 9, 10, 12, 16]


@pytest.mark.parametrize("dim,n", [(3, 6), (2, 9), (3, (3, 5, 4))])
def test_rank_is_permutation(dim, n):
    mesh = generate_box_mesh(dim, 0.0, 1.0, n if isinstance(n, tuple) else (n,) * dim)
    r = element_traversal_rank(mesh)
    assert r.dtype == np.int32
    assert np.array_equal(np.sort(r), np.arange(mesh.nelements))


def test_rank_follows_z_order():
    """On a 2^3 lattice of cubes the Z-order visits the 8 octants of each
    level in turn: the first 6 * 8 ranks fill the lower corner's 2x2x2
    cubes, and consecutive ranks stay close (mean centroid step well under
    the natural order's, whose k-steps jump a whole layer)."""
    n = 8
    mesh = generate_box_mesh(3, 0.0, 1.0, (n,) * 3)
    r = element_traversal_rank(mesh)
    c = mesh.vertices[mesh.elements].mean(axis=1)
    order = np.argsort(r)
    first = c[order[:48]]
    assert first.max() < 2.0 / n
    step_z = np.linalg.norm(np.diff(c[order], axis=0), axis=1)
    step_nat = np.linalg.norm(np.diff(c, axis=0), axis=1)
    assert step_z.mean() < 1.5 / n
    assert np.percentile(step_z, 99) < np.percentile(step_nat, 99.9)


def test_workspace_grows_by_tables():
    from mirgecom_amd import _lib
    lib = _lib.load()
    K = 1000
    m = _lib.MgMesh(dim=3, order=3, np=20, nfp=10, nfaces=4, nelements=K)
    base = lib.mg_mesh_workspace_bytes(m)
    assert base > 0 and base % 256 == 0
    rank = (ctypes.c_int32 * K)()
    m.elem_rank = ctypes.addressof(rank)
    ints = K + sum((K + e - 1) // e for e in EPBS)
    assert lib.mg_mesh_workspace_bytes(m) == base + 4 * ints
