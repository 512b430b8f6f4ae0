"""VTK output of volume fields: the visualizer that
mirgecom/simutil.py:309-386 (``write_visfile``) drives.

The reference obtains it from ``grudge.shortcuts.make_visualizer`` (a
meshmode ``Visualizer``): fields are interpolated from the DG nodes to an
equispaced lattice of order ``vis_order`` in every element, each element's
lattice is split into ``vis_order^dim`` linear simplices, and the result is
written as a VTK unstructured grid (``.vtu``, one per rank) with a ``.pvtu``
manifest.  The same is done here on the host (the data leaves the GPU for
the file anyway): interpolation by ``V_eq V^-1`` of the element basis
(exact for the degree-``p`` fields), points duplicated per element (the
fields are discontinuous), linear ``VTK_TRIANGLE`` / ``VTK_TETRA`` cells,
base64 binary data arrays.
"""

from __future__ import annotations

import base64
import itertools
import os
from xml.sax.saxutils import quoteattr

import numpy as np

from mirgecom_amd.refelem import simplex_basis

VTK_TRIANGLE = 5
VTK_TETRA = 10


def equispaced_lattice(dim, order):
    """Lattice multi-indices ``(n, dim)`` with ``sum <= order`` and their
    biunit coordinates ``-1 + 2 i / order``."""
    idx = [t for t in itertools.product(range(order + 1), repeat=dim)
           if sum(t) <= order]
    idx = np.array(sorted(idx, key=lambda t: (sum(t), t[::-1])), dtype=np.int64)
    return idx, -1.0 + 2.0 * idx / order


def lattice_simplices(dim, order):
    """Split of the order-``order`` lattice into ``order^dim`` positively
    oriented linear simplices: ``(ncells, dim+1)`` lattice point indices."""
    idx, x = equispaced_lattice(dim, order)
    pos = {tuple(t): n for n, t in enumerate(idx)}
    e = np.eye(dim, dtype=np.int64)
    cells = []

    def at(v, *offs):
        t = np.array(v) + sum(offs, np.zeros(dim, dtype=np.int64))
        return pos[tuple(t)]

    for v in idx:
        s = v.sum()
        if dim == 2:
            if s <= order - 1:
                cells.append([at(v), at(v, e[0]), at(v, e[1])])
            if s <= order - 2:
                cells.append([at(v, e[0]), at(v, e[0], e[1]), at(v, e[1])])
        else:
            if s <= order - 1:
                cells.append([at(v), at(v, e[0]), at(v, e[1]), at(v, e[2])])
            if s <= order - 2:
                # octahedron around the diagonal (v+e0, v+e1+e2)
                a, f = at(v, e[0]), at(v, e[1], e[2])
                ring = [at(v, e[1]), at(v, e[0], e[1]), at(v, e[0], e[2]),
                        at(v, e[2])]
                for r in range(4):
                    cells.append([a, f, ring[r], ring[(r + 1) % 4]])
            if s <= order - 3:
                cells.append([at(v, e[0], e[1]), at(v, e[0], e[2]),
                              at(v, e[1], e[2]), at(v, e[0], e[1], e[2])])
    cells = np.array(cells, dtype=np.int64)
    # orient positively
    for c in cells:
        m = (x[c[1:]] - x[c[0]]).T
        if np.linalg.det(m) < 0:
            c[[1, 2]] = c[[2, 1]]
    return cells


def _b64(arr):
    raw = np.ascontiguousarray(arr).tobytes()
    head = np.array([len(raw)], dtype=np.uint64).tobytes()
    return base64.b64encode(head + raw).decode("ascii")


_VTK_TYPE = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
             np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32",
             np.dtype(np.uint8): "UInt8"}


def _data_array(name, arr, ncomp):
    arr = np.ascontiguousarray(arr)
    return (f'<DataArray type="{_VTK_TYPE[arr.dtype]}" Name={quoteattr(name)} '
            f'NumberOfComponents="{ncomp}" format="binary">{_b64(arr)}'
            '</DataArray>\n')


def _host(a):
    try:
        import torch
        if isinstance(a, torch.Tensor):
            return a.detach().to("cpu", dtype=torch.float64).numpy()
    except ImportError:
        pass
    return np.asarray(a, dtype=np.float64)


def flatten_fields(names_and_fields):
    """``(name, field)`` pairs -> ``(name, array (ncomp, K, Np))``: tensors or
    arrays ``(K, Np)`` are scalars, ``(n, K, Np)`` vectors of n components,
    object arrays of ``(K, Np)`` fields vectors; a ``ConservedVars`` becomes
    ``name_mass``, ``name_energy``, ``name_momentum`` and, with species,
    ``name_species_mass`` (as meshmode flattens array containers)."""
    from mirgecom_amd.fluid import ConservedVars
    out = []
    for name, fld in names_and_fields:
        if isinstance(fld, ConservedVars):
            sub = [("mass", fld.mass), ("energy", fld.energy),
                   ("momentum", fld.momentum)]
            if fld.nspecies:
                sub.append(("species_mass", fld.species_mass))
            out.extend(flatten_fields([(f"{name}_{s}", f) for s, f in sub]))
            continue
        if isinstance(fld, np.ndarray) and fld.dtype == object:
            arr = np.stack([_host(x) for x in fld.ravel()])
        else:
            arr = _host(fld)
        if arr.ndim == 2:
            arr = arr[None]
        if arr.ndim != 3:
            raise ValueError(f"field {name!r}: expected (K, Np) or (n, K, Np)")
        out.append((name, arr))
    return out


class Visualizer:
    """Writes volume fields of a discretization as subdivided linear cells
    (meshmode.discretization.visualization.Visualizer's default mode)."""

    def __init__(self, dcoll, vis_order=None):
        re = dcoll.refel
        self.dcoll = dcoll
        self.dim = dcoll.dim
        self.vis_order = re.order if vis_order is None else int(vis_order)
        if self.vis_order < 1:
            raise ValueError("vis_order must be at least 1")
        _, xeq = equispaced_lattice(self.dim, self.vis_order)
        # interpolation DG nodes -> lattice: V_eq V^-1 (same polynomial space)
        veq = simplex_basis(self.dim, re.order, xeq)
        self.interp = veq @ np.linalg.inv(re.vdm)          # (nvis, Np)
        self.cells = lattice_simplices(self.dim, self.vis_order)
        self.nvis = xeq.shape[0]

    def _points(self):
        x = np.einsum("vn,dkn->kvd", self.interp, self.dcoll.host_nodes)
        pts = np.zeros(x.shape[:2] + (3,))
        pts[..., :self.dim] = x
        return pts.reshape(-1, 3)

    def _connectivity(self):
        k = self.dcoll.nelements
        conn = (self.cells[None] + self.nvis * np.arange(k)[:, None, None])
        nv = self.dim + 1
        offsets = nv * np.arange(1, conn.shape[0] * conn.shape[1] + 1)
        ctype = VTK_TRIANGLE if self.dim == 2 else VTK_TETRA
        types = np.full(offsets.size, ctype, dtype=np.uint8)
        return conn.reshape(-1), offsets.astype(np.int64), types

    def vis_fields(self, names_and_fields):
        """Fields at the visualization points: ``(name, (npts, ncomp))``."""
        out = []
        for name, arr in flatten_fields(names_and_fields):
            v = np.einsum("vn,ckn->kvc", self.interp, arr)
            out.append((name, v.reshape(-1, arr.shape[0])))
        return out

    def write_vtk_file(self, file_name, names_and_fields, overwrite=False,
                       real_only=False, par_manifest_filename=None):
        if os.path.exists(file_name) and not overwrite:
            raise FileExistsError(f"output file '{file_name}' already exists")
        pts = self._points()
        conn, offsets, types = self._connectivity()
        fields = self.vis_fields(names_and_fields)
        with open(file_name, "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" '
                    'version="1.0" byte_order="LittleEndian" '
                    'header_type="UInt64">\n<UnstructuredGrid>\n')
            f.write(f'<Piece NumberOfPoints="{pts.shape[0]}" '
                    f'NumberOfCells="{offsets.size}">\n')
            f.write("<Points>\n" + _data_array("points", pts, 3) + "</Points>\n")
            f.write("<Cells>\n" + _data_array("connectivity", conn, 1)
                    + _data_array("offsets", offsets, 1)
                    + _data_array("types", types, 1) + "</Cells>\n")
            f.write("<PointData>\n")
            for name, v in fields:
                f.write(_data_array(name, v, v.shape[1]))
            f.write("</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
        if par_manifest_filename is not None:
            self._write_manifest(par_manifest_filename, [file_name], fields,
                                 overwrite)

    def write_parallel_vtk_file(self, mpi_comm, file_name_pattern,
                                names_and_fields, overwrite=False,
                                par_manifest_filename=None):
        """One ``.vtu`` per rank (``file_name_pattern.format(rank=r)``) and,
        from rank 0, the ``.pvtu`` manifest listing all of them."""
        rank = mpi_comm.Get_rank() if mpi_comm is not None else 0
        nranks = mpi_comm.Get_size() if mpi_comm is not None else 1
        fname = file_name_pattern.format(rank=rank)
        self.write_vtk_file(fname, names_and_fields, overwrite=overwrite)
        if rank == 0 and par_manifest_filename is not None:
            pieces = [file_name_pattern.format(rank=r) for r in range(nranks)]
            self._write_manifest(par_manifest_filename, pieces,
                                 self.vis_fields(names_and_fields), overwrite)

    @staticmethod
    def _write_manifest(path, pieces, fields, overwrite):
        if os.path.exists(path) and not overwrite:
            raise FileExistsError(f"output file '{path}' already exists")
        base = os.path.dirname(path)
        with open(path, "w") as f:
            f.write('<?xml version="1.0"?>\n<VTKFile type="PUnstructuredGrid" '
                    'version="1.0" byte_order="LittleEndian" '
                    'header_type="UInt64">\n<PUnstructuredGrid GhostLevel="0">\n')
            f.write('<PPoints><PDataArray type="Float64" '
                    'NumberOfComponents="3"/></PPoints>\n<PPointData>\n')
            for name, v in fields:
                f.write(f'<PDataArray type="Float64" Name={quoteattr(name)} '
                        f'NumberOfComponents="{v.shape[1]}"/>\n')
            f.write("</PPointData>\n")
            for p in pieces:
                f.write(f"<Piece Source={quoteattr(os.path.relpath(p, base or '.'))}/>\n")
            f.write("</PUnstructuredGrid>\n</VTKFile>\n")


def make_visualizer(dcoll, vis_order=None, **kwargs):
    """grudge.shortcuts.make_visualizer (as used by mirgecom's drivers)."""
    return Visualizer(dcoll, vis_order=vis_order)


__all__ = ["Visualizer", "make_visualizer"]
