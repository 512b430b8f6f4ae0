"""Block visiting order (mg_mesh.elem_rank): the tables mg_mesh_prepare
builds equal their definition (the element blocks of each elements-per-
workgroup count sorted by the rank of their first element; the natural
order for a rank array that is not a permutation), and the right-hand sides
are bit-identical with the Morton order and with the natural order (the
order only changes which workgroup takes which block)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.cases import mixture_state, smooth_perturbed_state  # noqa: E402

pytestmark = pytest.mark.gpu

EPBS = [4, 5, 6, 7, 8, 9, 10, 12, 16]


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _pair(dim, order, n, periodic):
    from mirgecom_amd.discretization import (
        create_discretization_collection, generate_box_mesh)
    mesh = generate_box_mesh(dim, 0.0, 2 * np.pi, (n,) * dim, periodic=(periodic,) * dim)
    dm = create_discretization_collection(mesh, order)
    dn = create_discretization_collection(mesh, order)
    dn.traversal = "natural"
    return mesh, dm, dn


def _expected(rank, K, epb):
    nb = (K + epb - 1) // epb
    return np.argsort(rank[np.arange(nb) * epb], kind="stable").astype(np.int32)


def _tables(dd):
    ws = dd.workspace
    off = dd.mesh.block_order - ws.data_ptr()
    assert off > 0 and off % 4 == 0
    K = dd.host.nelements
    flat = ws.view(torch.int32)[off // 4:].cpu().numpy()
    out, pos = {}, K
    for e in EPBS:
        nb = (K + e - 1) // e
        out[e] = flat[pos:pos + nb]
        pos += nb
    return out


@pytest.mark.parametrize("dim,order,n", [(3, 3, 5), (2, 2, 7), (3, 1, 4)])
def test_tables(dim, order, n):
    from mirgecom_amd.device import get_device_discretization
    from mirgecom_amd.discretization import element_traversal_rank
    mesh, dm, _ = _pair(dim, order, n, True)
    dd = get_device_discretization(dm)
    torch.cuda.synchronize()
    rank = element_traversal_rank(mesh)
    got = _tables(dd)
    for e in EPBS:
        assert np.array_equal(got[e], _expected(rank, mesh.nelements, e)), e


def test_non_permutation_falls_back():
    from mirgecom_amd import _lib
    from mirgecom_amd.device import DeviceDiscretization
    mesh, dm, _ = _pair(3, 2, 3, True)
    bad = np.zeros(mesh.nelements, dtype=np.int32)   # every rank 0
    import mirgecom_amd.discretization as disc
    orig = disc.element_traversal_rank
    try:
        disc.element_traversal_rank = lambda m: bad
        dd = DeviceDiscretization(dm, "cuda")
    finally:
        disc.element_traversal_rank = orig
    torch.cuda.synchronize()
    got = _tables(dd)
    K = mesh.nelements
    for e in EPBS:
        assert np.array_equal(got[e], np.arange((K + e - 1) // e, dtype=np.int32)), e
    assert _lib.load().mg_mesh_workspace_bytes(dd.mesh) > 0


def test_ns_and_euler_bit_identical():
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    mesh, dm, dn = _pair(3, 3, 5, True)
    x = dm.host_nodes
    q = smooth_perturbed_state(x, seed=3)
    gm = GasModel(eos=IdealSingleGas(),
                  transport=SimpleTransport(viscosity=1e-3, thermal_conductivity=1e-3))
    st = make_fluid_state(ConservedVars(3, _dev(q)), gm)
    for prec in ("fp64", "mixed"):
        kw = {} if prec == "fp64" else {"precision": "mixed"}
        a = ns_operator(dm, gm, st, {}, **kw).q
        b = ns_operator(dn, gm, st, {}, **kw).q
        assert torch.equal(a, b), prec
    a = euler_operator(dm, st, gm, {}, inviscid_numerical_flux_func=inviscid_facial_flux_rusanov).q
    b = euler_operator(dn, st, gm, {}, inviscid_numerical_flux_func=inviscid_facial_flux_rusanov).q
    assert torch.equal(a, b)


def test_mixture_bit_identical():
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport
    mesh, dm, dn = _pair(3, 2, 4, True)
    x = dm.host_nodes
    q, t = mixture_state(x, seed=21)
    tr = SimpleTransport(viscosity=2e-5, thermal_conductivity=0.05,
                         species_diffusivity=np.linspace(1e-5, 3e-5, 7))
    gm = GasModel(eos=PyrometheusMixture(), transport=tr)
    st = make_fluid_state(ConservedVars(3, _dev(q)), gm, temperature_seed=_dev(t))
    a = euler_operator(dm, st, gm, {}, inviscid_numerical_flux_func=inviscid_facial_flux_rusanov).q
    b = euler_operator(dn, st, gm, {}, inviscid_numerical_flux_func=inviscid_facial_flux_rusanov).q
    assert torch.equal(a, b)
    a = ns_operator(dm, gm, st, {}).q
    b = ns_operator(dn, gm, st, {}).q
    assert torch.equal(a, b)
