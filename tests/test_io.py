"""Restart files (mirgecom/restart.py:54-78, test/test_restart.py:43-88
restated) and VTK output (mirgecom/simutil.py:309-386): exact round trips,
the reference's directory rules, and VTU files whose cells tile each element
and whose point data is the exact interpolant of the DG fields."""

import base64
import math
import os
import xml.etree.ElementTree as ET
import zipfile

import numpy as np
import pytest
import torch

from mirgecom_amd.discretization import (
    create_discretization_collection, generate_box_mesh)
from mirgecom_amd.fluid import ConservedVars, make_conserved
from mirgecom_amd.restart import PathError, read_restart_data, write_restart_file
from mirgecom_amd.visualization import (
    equispaced_lattice, lattice_simplices, make_visualizer)


@pytest.mark.parametrize("nspecies", [0, 10])
def test_restart_cv(tmp_path, nspecies):
    """test/test_restart.py:43-88: a CV container survives the round trip
    bit for bit (with the rest of a driver's restart dictionary)."""
    dim = 3
    mesh = generate_box_mesh(dim, -0.5, 0.5, (4,) * dim)
    dcoll = create_discretization_collection(mesh, 3)
    x = torch.as_tensor(dcoll.host_nodes)
    mass = x[0]
    mom = torch.stack([x[2] * (i + 3) for i in range(dim)])
    sm = (torch.stack([i * x[0] for i in range(nspecies)]) * mass
          if nspecies else None)
    cv = make_conserved(dim, mass=mass, energy=x[1], momentum=mom,
                        species_mass=sm)
    data = {"state": cv, "t": 0.1 + 1e-17, "step": 42, "order": 3,
            "local_mesh": mesh, "nparts": 1, "tseed": np.linspace(0, 1, 7),
            "nested": [1, (2.5, "x"), {"a": None, 3: True}]}
    fn = str(tmp_path / "rst" / f"test_{nspecies}.pkl")
    write_restart_file(None, data, fn)
    back = read_restart_data(None, fn)
    assert isinstance(back["state"], ConservedVars)
    assert torch.equal(back["state"].q, cv.q)
    assert back["t"] == data["t"] and back["step"] == 42
    assert np.array_equal(back["tseed"], data["tseed"])
    assert back["nested"] == data["nested"]
    m = back["local_mesh"]
    assert np.array_equal(m.vertices, mesh.vertices)
    assert np.array_equal(m.elements, mesh.elements)
    assert m.periodic == mesh.periodic and m.n == mesh.n
    # nothing in the file needs unpickling
    with zipfile.ZipFile(fn) as zf:
        for name in zf.namelist():
            if name.endswith(".npy"):
                import io
                np.load(io.BytesIO(zf.read(name)), allow_pickle=False)


def test_restart_path_error(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    with pytest.raises(PathError):
        write_restart_file(None, {"a": 1}, str(blocker / "rst.pkl"))
    with pytest.raises(TypeError):
        write_restart_file(None, {"a": object()}, str(tmp_path / "r.pkl"))


def _read_vtu(path):
    root = ET.parse(path).getroot()
    assert root.attrib["header_type"] == "UInt64"
    piece = root.find("UnstructuredGrid/Piece")
    types = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}

    def arr(el):
        raw = base64.b64decode(el.text)
        n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
        a = np.frombuffer(raw[8:8 + n], dtype=types[el.attrib["type"]])
        return a.reshape(-1, int(el.attrib["NumberOfComponents"]))
    pts = arr(piece.find("Points/DataArray"))
    cells = {el.attrib["Name"]: arr(el) for el in piece.find("Cells")}
    data = {el.attrib["Name"]: arr(el) for el in piece.find("PointData")}
    return piece, pts, cells, data


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 3), (3, 2), (3, 4)])
def test_lattice_tiles_reference_element(dim, order):
    """order^dim positive cells; random points of the reference simplex lie
    in exactly one cell (no gaps, no overlaps)."""
    cells = lattice_simplices(dim, order)
    idx, x = equispaced_lattice(dim, order)
    assert len(cells) == order ** dim
    rng = np.random.default_rng(0)
    b = rng.dirichlet(np.ones(dim + 1), size=400)
    verts = np.vstack([-np.ones(dim), -np.ones(dim) + 2 * np.eye(dim)])
    pts = b @ verts
    count = np.zeros(len(pts), dtype=int)
    for c in cells:
        m = (x[c[1:]] - x[c[0]]).T
        assert np.linalg.det(m) > 0
        lam = np.linalg.solve(m, (pts - x[c[0]]).T)
        inside = (lam > -1e-12).all(axis=0) & (lam.sum(axis=0) < 1 + 1e-12)
        count += inside
    assert np.all(count == 1)


@pytest.mark.parametrize("dim,order,vis_order", [(2, 3, None), (3, 2, None),
                                                 (3, 3, 5)])
def test_vtu_exact_interpolant(tmp_path, dim, order, vis_order):
    from mirgecom_amd.simutil import write_visfile
    mesh = generate_box_mesh(dim, 0.0, (1.0, 0.7, 0.4)[:dim], (3, 2, 2)[:dim])
    dcoll = create_discretization_collection(mesh, order)
    x = dcoll.host_nodes

    def poly(xx):   # degree `order` in the coordinates
        return 1.0 + xx[0] ** order - 2.0 * xx[0] * xx[-1] ** (order - 1)

    f = poly(x)
    vec = np.stack([x[d] * x[0] for d in range(dim)])
    cv = make_conserved(dim, mass=torch.as_tensor(f),
                        energy=torch.as_tensor(2 * f),
                        momentum=torch.as_tensor(vec),
                        species_mass=torch.as_tensor(np.stack([f, 3 * f])))
    vis = make_visualizer(dcoll, vis_order)
    viz = str(tmp_path / "viz" / "fld")
    write_visfile(dcoll, [("f", f), ("cv", cv)], vis, vizname=viz, step=7,
                  t=0.5)
    vtu = viz + "-000000007-0000.vtu"
    piece, pts, cells, data = _read_vtu(vtu)
    p = vis.vis_order
    nvis = math.comb(p + dim, dim)
    K = dcoll.nelements
    assert int(piece.attrib["NumberOfPoints"]) == K * nvis == len(pts)
    assert int(piece.attrib["NumberOfCells"]) == K * p ** dim
    assert set(data) == {"f", "cv_mass", "cv_energy", "cv_momentum",
                         "cv_species_mass"}
    xp = pts[:, :dim].T
    np.testing.assert_allclose(data["f"][:, 0], poly(xp), rtol=0, atol=1e-12)
    np.testing.assert_allclose(data["cv_energy"][:, 0], 2 * poly(xp), atol=1e-12)
    np.testing.assert_allclose(data["cv_momentum"],
                               np.stack([xp[d] * xp[0] for d in range(dim)]).T,
                               atol=1e-12)
    assert data["cv_species_mass"].shape == (K * nvis, 2)
    # the cells tile the box
    conn = cells["connectivity"].reshape(-1, dim + 1)
    vol = 0.0
    for c in conn:
        m = (pts[c[1:], :dim] - pts[c[0], :dim]).T
        d = np.linalg.det(m)
        assert d > 0
        vol += d / math.factorial(dim)
    assert vol == pytest.approx(np.prod((1.0, 0.7, 0.4)[:dim]), rel=1e-12)
    assert np.all(cells["types"] == (5 if dim == 2 else 10))
    # manifest
    man = ET.parse(viz + "-000000007.pvtu").getroot()
    srcs = [el.attrib["Source"] for el in man.iter("Piece")]
    assert srcs == [os.path.basename(vtu)]
    with pytest.raises(FileExistsError):
        write_visfile(dcoll, [("f", f)], vis, vizname=viz, step=7)
    write_visfile(dcoll, [("f", f)], vis, vizname=viz, step=7, overwrite=True)
