"""ctypes binding of ``libmgdg.so`` (C ABI in include/mgdg.h).

The library is built in-tree by ``mirgecom_amd.build``.  There is no fallback:
if the library is missing, loading raises, and every operator that needs it
fails loudly.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MGDG_LIB overrides the library path (a build elsewhere, e.g. for A/B timing)
LIB_PATH = os.environ.get("MGDG_LIB") or os.path.join(_HERE, "libmgdg.so")

MG_FLUX_RUSANOV = 0
MG_FLUX_HLL = 1
MG_FLUX_CENTRAL = 2


class MgMesh(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("order", ctypes.c_int32),
        ("np", ctypes.c_int32), ("nfp", ctypes.c_int32),
        ("nfaces", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("nelements", ctypes.c_int64), ("nbnd_nodes", ctypes.c_int64),
        ("rx", ctypes.c_void_p), ("normals", ctypes.c_void_p),
        ("fscale", ctypes.c_void_p), ("vmap_p", ctypes.c_void_p),
        ("weak_diff", ctypes.c_void_p), ("lift", ctypes.c_void_p),
        ("face_mask", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p),
        ("elem_rank", ctypes.c_void_p), ("block_order", ctypes.c_void_p),
    ]


class MgGas(ctypes.Structure):
    _fields_ = [
        ("eos_kind", ctypes.c_int32), ("transport_kind", ctypes.c_int32),
        ("nspecies", ctypes.c_int32), ("has_lewis", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("gas_const", ctypes.c_double),
        ("mu", ctypes.c_double), ("mu_b", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("pl_alpha", ctypes.c_double), ("pl_beta", ctypes.c_double),
        ("pl_sigma", ctypes.c_double), ("pl_n", ctypes.c_double),
        ("species_diffusivity", ctypes.c_double * 8),
        ("ma_alpha", ctypes.c_double), ("ma_factor", ctypes.c_double),
        ("ma_epsilon", ctypes.c_double),
        ("ma_singular_diffusivity", ctypes.c_double),
    ]


class MgQuad(ctypes.Structure):
    _fields_ = [
        ("nq", ctypes.c_int32), ("nfq", ctypes.c_int32),
        ("interp", ctypes.c_void_p), ("weak_diff", ctypes.c_void_p),
        ("face_interp", ctypes.c_void_p), ("lift", ctypes.c_void_p),
    ]


class MgEsdg(ctypes.Structure):
    _fields_ = [
        ("nh", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("proj", ctypes.c_void_p), ("vh", ctypes.c_void_p),
        ("skew", ctypes.c_void_p), ("ext", ctypes.c_void_p),
        ("ext_col", ctypes.c_void_p), ("apply", ctypes.c_void_p),
        ("face_weights", ctypes.c_void_p), ("skew_diag", ctypes.c_void_p),
        ("ext_diag", ctypes.c_void_p),
    ]


class MgChemNorm(ctypes.Structure):
    """``mg_chem_norm`` (include/mgdg.h): surrogate normalisation and gate."""
    _fields_ = [("in_mean", ctypes.c_double * 10),
                ("in_scale", ctypes.c_double * 10),
                ("out_mean", ctypes.c_double * 7),
                ("out_scale", ctypes.c_double * 7),
                ("lo", ctypes.c_double * 10), ("hi", ctypes.c_double * 10),
                ("check_tol", ctypes.c_double),
                ("sample_period", ctypes.c_int32),
                ("sample_offset", ctypes.c_int32)]


class MgEnsCase(ctypes.Structure):
    """``mg_ens_case`` (include/mgdg.h): one case of a batched ensemble."""
    _fields_ = [("mesh", MgMesh), ("gas", MgGas), ("q", ctypes.c_void_p),
                ("t_seed", ctypes.c_void_p), ("temperature", ctypes.c_void_p),
                ("fnet", ctypes.c_void_p), ("grad", ctypes.c_void_p),
                ("rhs", ctypes.c_void_p)]


class MgEnsTuning(ctypes.Structure):
    _fields_ = [("fused", ctypes.c_int32), ("persistent", ctypes.c_int32),
                ("wg_per_cu", ctypes.c_int32 * 5)]


class MgEnsPlan(ctypes.Structure):
    _I5 = ctypes.c_int64 * 5
    _fields_ = [("dev", ctypes.c_void_p), ("ncases", ctypes.c_int32),
                ("fused", ctypes.c_int32), ("niter", ctypes.c_int32),
                ("cus", ctypes.c_int32), ("surrogate", ctypes.c_int32),
                ("persistent", ctypes.c_int32),
                ("grad_begin", _I5), ("grad_count", _I5), ("div_begin", _I5),
                ("div_count", _I5),
                ("grad_seg_begin", ctypes.c_int32 * 5),
                ("grad_seg_count", ctypes.c_int32 * 5),
                ("div_seg_begin", ctypes.c_int32 * 5),
                ("div_seg_count", ctypes.c_int32 * 5),
                ("pt_count", ctypes.c_int64),
                ("src_count", ctypes.c_int64),
                ("grid_grad", ctypes.c_int32 * 5), ("grid_div", ctypes.c_int32 * 5),
                ("grid_pt", ctypes.c_int32), ("grid_src", ctypes.c_int32),
                ("off_grad_tab", ctypes.c_int64), ("off_div_tab", ctypes.c_int64),
                ("off_pt_tab", ctypes.c_int64), ("off_src_tab", ctypes.c_int64),
                ("off_grad_items", ctypes.c_int64), ("off_div_items", ctypes.c_int64),
                ("off_pt_items", ctypes.c_int64), ("off_src_items", ctypes.c_int64)]


# exported symbol -> (restype, argtypes)
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double


class MgWall(ctypes.Structure):
    """``mg_wall`` (include/mgdg.h): Mengaldo boundary data."""
    _fields_ = [("q_bc", ctypes.c_void_p), ("t_bc", ctypes.c_void_p),
                ("t_state_bc", ctypes.c_void_p), ("flags", ctypes.c_void_p),
                ("vol_index", ctypes.c_void_p), ("normals", ctypes.c_void_p),
                ("work", ctypes.c_void_p)]


_W = ctypes.POINTER(MgWall)
SIGNATURES = {
    "mg_euler_rhs": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                            _P, _P, _P, _P, _I32, _P, _P]),
    "mg_euler_rhs_mixed": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                                  _P, _P, _P, _P, _I32, _P, _P]),
    "mg_ns_gradient": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                              _P, _P, _P, _P, _W, _P, _P]),
    "mg_ns_rhs": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                         _P, _P, _P, _P, _P, _W, _I32, _I32, _P, _P]),
    "mg_ns_flux": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                          _P, _P, _P, _P, _W, _I32, _P, _P, _P, _P]),
    "mg_ns_div": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                         _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P]),
    "mg_ns_flux_mixed": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                                _P, _P, _P, _P, _W, _I32, _P, _P, _P, _P]),
    "mg_ns_div_mixed": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                               _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P]),
    "mg_mixture_temperature": (_I32, [ctypes.POINTER(MgGas), _I32, _I64, _P,
                                      _P, _I32, _P, _P]),
    "mg_mixture_temperature_tol": (_I32, [ctypes.POINTER(MgGas), _I32, _I64, _P, _P,
                                          _D, _I32, _P, _P, _P]),
    "mg_species_source": (_I32, [ctypes.POINTER(MgGas), _I32, _I64, _P, _P,
                                 _P, _P]),
    "mg_species_source_add": (_I32, [ctypes.POINTER(MgGas), _I32, _I64, _P, _P,
                                     _P, _P]),
    "mg_local_timestep": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                                 _P, _P, _P, _I32, _P, _P]),
    "mg_rk_axpy": (_I32, [_I64, _P, _P, _D, _P, _P, _D, _P]),
    "mg_euler_rhs_quad": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                                 ctypes.POINTER(MgQuad), _P, _P, _P, _P, _I32,
                                 _I32, _P, _P]),
    "mg_ns_gradient_quad": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                                   ctypes.POINTER(MgQuad), _P, _P, _P, _P, _I32,
                                   _P, _P]),
    "mg_ns_rhs_quad": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                              ctypes.POINTER(MgQuad), _P, _P, _P, _P, _P, _I32,
                              _I32, _I32, _P, _P]),
    "mg_euler_rhs_esdg": (_I32, [ctypes.POINTER(MgMesh), ctypes.POINTER(MgGas),
                                 ctypes.POINTER(MgQuad), ctypes.POINTER(MgEsdg),
                                 _P, _P, _I32, _P, _P, _P]),
    "mg_euler_rhs_esdg_mixture": (_I32, [ctypes.POINTER(MgMesh),
                                         ctypes.POINTER(MgGas),
                                         ctypes.POINTER(MgQuad),
                                         ctypes.POINTER(MgEsdg), _P, _P, _P, _P,
                                         _I32, _I32, _I32, _P, _P, _P]),
    "mg_euler_rhs_esdg_noalloc": (_I32, [ctypes.POINTER(MgMesh),
                                         ctypes.POINTER(MgGas),
                                         ctypes.POINTER(MgQuad),
                                         ctypes.POINTER(MgEsdg),
                                         _P, _P, _I32, _P, _P]),
    "mg_bound_limiter": (_I32, [_I32, _I64, _I32, _P, _P, _P, _D, _D, _I32,
                                _P]),
    "mg_load_mechanism": (_I32, [ctypes.POINTER(ctypes.c_double), _I64]),
    "mg_load_transport": (_I32, [ctypes.POINTER(ctypes.c_double), _I64]),
    "mg_transport_vars": (_I32, [ctypes.POINTER(MgGas), _I32, _I64, _P, _P, _P,
                                 _P, _P, _P, _P]),
    "mg_reactor_trajectories": (_I32, [_I64, _P, _P, _P, _D, _I64, _D, _I32, _P,
                                       _P, _P, _P]),
    "mg_chem_substep": (_I32, [_I64, _P, _D, _I32, _P, _P]),
    "mg_chem_features": (_I32, [_I64, _P, _P, ctypes.POINTER(MgChemNorm), _P, _P,
                                _P]),
    "mg_mlp_param_count": (_I64, []),
    "mg_mlp_bf16_count": (_I64, []),
    "mg_mlp_workspace_bytes": (_I64, [_I64]),
    "mg_mlp_forward": (_I32, [_I64, _P, _P, _P, _P, _P, _P]),
    "mg_mlp_backward": (_I32, [_I64, _P, _P, _P, _P, _P, _P, _P]),
    "mg_mlp_adam": (_I32, [_P, _P, _P, _P, _I64, _I32, ctypes.c_float,
                           ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P,
                           _P]),
    "mg_mlp_sync_bf16": (_I32, [_P, _P, _P]),
    "mg_chem_surrogate_source": (_I32, [ctypes.POINTER(MgGas), _I32, _I32, _I64,
                                        _P, _P, _D, _P, _P,
                                        ctypes.POINTER(MgChemNorm), _I32, _P, _P,
                                        _P]),
    "mg_ensemble_plan_bytes": (_I64, [_I32, ctypes.POINTER(MgEnsCase)]),
    "mg_ensemble_prepare": (_I32, [_I32, ctypes.POINTER(MgEnsCase), _I32, _D, _P, _P,
                                   ctypes.POINTER(MgChemNorm), _P,
                                   ctypes.POINTER(MgEnsTuning), _P, _I64,
                                   ctypes.POINTER(MgEnsPlan), _P]),
    "mg_ensemble_ns_rhs": (_I32, [ctypes.POINTER(MgEnsPlan), _P]),
    "mg_mesh_workspace_bytes": (_I64, [ctypes.POINTER(MgMesh)]),
    "mg_mesh_prepare": (_I32, [ctypes.POINTER(MgMesh), _P, _P]),
    "mg_last_error": (ctypes.c_char_p, []),
    "mg_abi_version": (_I32, []),
}

_lib = None


class MgError(RuntimeError):
    pass


def load():
    """Load and type the library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m mirgecom_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        msg = load().mg_last_error().decode()
        raise MgError(f"libmgdg error {status}: {msg}")


_mech_loaded = False
_tran_loaded = False


def ensure_mechanism():
    global _mech_loaded
    if _mech_loaded:
        return
    from mirgecom_amd.mechanisms import packed_mechanism
    arr = np.ascontiguousarray(packed_mechanism())
    check(load().mg_load_mechanism(
        arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr.size))
    _mech_loaded = True


def ensure_transport():
    """Upload the uiuc_7sp transport fits (once per process)."""
    global _tran_loaded
    if _tran_loaded:
        return
    from mirgecom_amd.transport_fits import packed_transport
    arr = np.ascontiguousarray(packed_transport())
    check(load().mg_load_transport(
        arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr.size))
    _tran_loaded = True


# Optional per-entry-point timing: when ``timing`` is a dict, ``call`` records
# a pair of events on the current stream around every launch and appends it
# to ``timing[name]`` (used by the benchmark; off by default).
timing = None


def call(name, *args):
    """Call entry point ``name`` and raise ``MgError`` on a nonzero status."""
    fn = getattr(load(), name)
    if timing is None:
        check(fn(*args))
        return
    import torch
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    check(fn(*args))
    stop.record()
    timing.setdefault(name, []).append((start, stop))


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
