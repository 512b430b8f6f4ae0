"""The C ABI library: it loads without a GPU, exports every entry point that
include/mgdg.h declares, and rejects bad arguments with an error code and a
message before touching the device (no compute calls here)."""

import ctypes
import os
import re

import pytest

from mirgecom_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include",
                      "mgdg.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:mg_status|int64_t|int32_t|const char\*)"
                                 r"\s+(mg_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("mg_euler_rhs", "mg_ns_gradient", "mg_ns_rhs", "mg_ns_flux",
              "mg_ns_div", "mg_mixture_temperature", "mg_species_source",
              "mg_rk_axpy", "mg_mesh_prepare", "mg_mesh_workspace_bytes",
              "mg_load_mechanism", "mg_last_error", "mg_abi_version"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name
    assert lib.mg_abi_version() == 6


def test_null_arguments_rejected():
    lib = _lib.load()
    rc = lib.mg_euler_rhs(None, None, None, None, None, None, 0, None, None)
    assert rc == 1
    assert b"mesh" in lib.mg_last_error()
    rc = lib.mg_rk_axpy(10, None, None, 1.0, None, None, 0.0, None)
    assert rc == 1


def test_mesh_validation():
    lib = _lib.load()
    m = _lib.MgMesh(dim=3, order=3, np=20, nfp=10, nfaces=4, nelements=10,
                    nbnd_nodes=0)
    # matrices missing
    rc = lib.mg_ns_flux(m, _lib.MgGas(transport_kind=1), None, None, None,
                        None, None, 1, None, None, None, None)
    assert rc == 1
    m.np = 21
    assert lib.mg_euler_rhs(m, _lib.MgGas(), None, None, None, None, 0, None,
                            None) == 1
    assert b"np/nfp" in lib.mg_last_error()
    m5 = _lib.MgMesh(dim=3, order=5, np=56, nfp=21, nfaces=4, nelements=1)
    assert lib.mg_mesh_workspace_bytes(m5) == -1
    m3 = _lib.MgMesh(dim=3, order=3, np=20, nfp=10, nfaces=4, nelements=1)
    assert lib.mg_mesh_workspace_bytes(m3) > 0


def test_mechanism_size_checked():
    lib = _lib.load()
    arr = (ctypes.c_double * 3)()
    assert lib.mg_load_mechanism(arr, 3) == 1
    assert b"size" in lib.mg_last_error()
    from mirgecom_amd.mechanisms import packed_mechanism
    assert packed_mechanism().size == 130


def test_wall_validation():
    """An mg_wall with flags but without its CV_bc / T_bc arrays is rejected
    before any launch (checked after the mesh, so give a valid-looking one
    with boundary nodes and fake non-NULL device arrays)."""
    lib = _lib.load()
    m = _lib.MgMesh(dim=2, order=1, np=3, nfp=2, nfaces=3, nelements=2,
                    nbnd_nodes=4, rx=1, normals=1, fscale=1, vmap_p=1,
                    weak_diff=1, lift=1, face_mask=1, workspace=1)
    w = _lib.MgWall(q_bc=1, t_bc=1, t_state_bc=None, flags=1, vol_index=None,
                    normals=1, work=1)
    rc = lib.mg_ns_gradient(m, _lib.MgGas(gamma=1.4, gas_const=1.0), 1, None,
                            1, None, w, 1, None)
    assert rc == 1
    assert b"mg_wall" in lib.mg_last_error()


def test_gas_validation():
    lib = _lib.load()
    g = _lib.MgGas(eos_kind=1, nspecies=5)
    rc = lib.mg_mixture_temperature(g, 3, 10, None, None, 5, None, None)
    assert rc == 1
    assert b"7" in lib.mg_last_error()


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 4), (3, 1), (3, 4)])
def test_workspace_sizes(dim, order):
    import math
    lib = _lib.load()
    np_ = math.comb(order + dim, dim)
    nfp = math.comb(order + dim - 1, dim - 1)
    m = _lib.MgMesh(dim=dim, order=order, np=np_, nfp=nfp, nfaces=dim + 1,
                    nelements=1)
    rpw = (np_ + 3) // 4
    # FP64 + FP32 wave-packed operators, plus the FP32 MFMA A-fragments
    # (16-row tiles, K in steps of 4, 64 lanes) of Dw_r and each face's LIFT
    mt, ksv, ksf = (np_ + 15) // 16, (np_ + 3) // 4, (nfp + 3) // 4
    mfma = 64 * mt * (dim * ksv + (dim + 1) * ksf)
    ops = 12 * 4 * rpw * (np_ * dim + (dim + 1) * nfp) + 4 * mfma
    # rounded up to 256 bytes (the visiting-order tables follow)
    assert lib.mg_mesh_workspace_bytes(m) == (ops + 255) // 256 * 256


def test_limiter_validation():
    lib = _lib.load()
    one = ctypes.c_void_p(1)
    assert lib.mg_bound_limiter(20, 10, 1, None, one, one, 0.0, 1.0, 1,
                                None) == 1
    assert lib.mg_bound_limiter(20, 10, 1, one, one, one, 0.0, 1.0, 4,
                                None) == 1
    assert b"flags" in lib.mg_last_error()
    assert lib.mg_bound_limiter(7, 10, 1, one, one, one, 0.0, 1.0, 1,
                                None) == 2
    assert b"np" in lib.mg_last_error()
    # nothing to limit: accepted without a launch
    assert lib.mg_bound_limiter(20, 0, 1, one, one, one, 0.0, 1.0, 1,
                                None) == 0


def test_limiter_rejects_host_tensors():
    torch = pytest.importorskip("torch")
    from mirgecom_amd.discretization import (
        create_discretization_collection, generate_box_mesh)
    from mirgecom_amd.limiter import bound_preserving_limiter
    dcoll = create_discretization_collection(
        generate_box_mesh(2, 0.0, 1.0, (2, 2)), 2)
    with pytest.raises(TypeError):
        bound_preserving_limiter(dcoll, torch.zeros(dcoll.nelements, 6))


def test_transport_entry_points_validation():
    lib = _lib.load()
    arr = (ctypes.c_double * 5)()
    assert lib.mg_load_transport(arr, 5) == 1
    assert b"size" in lib.mg_last_error()
    from mirgecom_amd.transport_fits import packed_transport
    assert packed_transport().size == 308
    one = ctypes.c_void_p(1)
    # mixture-averaged transport with the ideal-gas EOS is rejected
    g = _lib.MgGas(eos_kind=0, transport_kind=3, nspecies=0)
    assert lib.mg_transport_vars(g, 3, 10, one, one, one, one, one, one,
                                 None) == 1
    assert b"mixture" in lib.mg_last_error()
    g = _lib.MgGas(eos_kind=0, transport_kind=5)
    assert lib.mg_transport_vars(g, 3, 10, one, one, one, one, one, one,
                                 None) == 1
    assert b"transport_kind" in lib.mg_last_error()
    g = _lib.MgGas(eos_kind=0, transport_kind=1, gamma=1.4, gas_const=1.0)
    assert lib.mg_transport_vars(g, 4, 10, one, None, one, one, one, None,
                                 None) == 1


def test_quadrature_entry_point_validation():
    lib = _lib.load()
    m = _lib.MgMesh(dim=2, order=2, np=6, nfp=3, nfaces=3, nelements=4,
                    nbnd_nodes=0, rx=1, normals=1, fscale=1, vmap_p=1,
                    weak_diff=1, lift=1, face_mask=1, workspace=1)
    gas = _lib.MgGas(eos_kind=0, gamma=1.4, gas_const=1.0)
    one = ctypes.c_void_p(1)
    # incomplete mg_quad
    qd = _lib.MgQuad(nq=9, nfq=3, interp=1, weak_diff=None, face_interp=1,
                     lift=1)
    assert lib.mg_euler_rhs_quad(m, gas, qd, one, None, None, None, 0, 0, one,
                                 None) == 1
    assert b"mg_quad" in lib.mg_last_error()
    qd = _lib.MgQuad(nq=9, nfq=3, interp=1, weak_diff=1, face_interp=1, lift=1)
    # a mixture needs its nodal temperature (and the mechanism)
    mix = _lib.MgGas(eos_kind=1, nspecies=7)
    assert lib.mg_euler_rhs_quad(m, mix, qd, one, None, None, None, 5, 0, one,
                                 None) == 1
    # boundary exterior states required when the mesh has boundary nodes
    m.nbnd_nodes = 6
    assert lib.mg_euler_rhs_quad(m, gas, qd, one, None, None, None, 0, 0, one,
                                 None) == 1
    assert b"boundary" in lib.mg_last_error()
    # an LDS footprint above 128 KiB per element is refused
    m.nbnd_nodes = 0
    big = _lib.MgQuad(nq=20000, nfq=3, interp=1, weak_diff=1, face_interp=1,
                      lift=1)
    assert lib.mg_euler_rhs_quad(m, gas, big, one, None, None, None, 0, 0, one,
                                 None) == 2


def test_esdg_entry_point_validation():
    lib = _lib.load()
    m = _lib.MgMesh(dim=2, order=2, np=6, nfp=3, nfaces=3, nelements=4,
                    nbnd_nodes=0, rx=1, normals=1, fscale=1, vmap_p=1,
                    weak_diff=1, lift=1, face_mask=1, workspace=1)
    gas = _lib.MgGas(eos_kind=0, gamma=1.4, gas_const=1.0)
    one = ctypes.c_void_p(1)
    qd = _lib.MgQuad(nq=9, nfq=3, interp=1, weak_diff=1, face_interp=1, lift=1)
    ok = dict(proj=1, vh=1, skew=1, ext=1, ext_col=1, apply=1, face_weights=1)
    # nh must be nq + nfaces * nfq
    bad = _lib.MgEsdg(nh=17, pad0=0, **ok)
    assert lib.mg_euler_rhs_esdg(m, gas, qd, bad, one, None, 1, one, one,
                                 None) == 1
    assert b"nh" in lib.mg_last_error()
    good = _lib.MgEsdg(nh=18, pad0=0, **ok)
    # the work buffer for the projected entropy variables is required
    assert lib.mg_euler_rhs_esdg(m, gas, qd, good, one, None, 1, None, one,
                                 None) == 1
    assert b"ev_work" in lib.mg_last_error()
    mix = _lib.MgGas(eos_kind=1, nspecies=7)
    assert lib.mg_euler_rhs_esdg(m, mix, qd, good, one, None, 1, one, one,
                                 None) == 2
