"""Restart files: call surface of mirgecom/restart.py:54-78.

``write_restart_file(actx, restart_data, filename, comm=None)`` and
``read_restart_data(actx, filename)`` keep the reference's signatures and
semantics (a nested dictionary of simulation data written by rank 0 into a
directory it creates; ``PathError`` when that path is a file), and the round
trip is exact.  The container is not a pickle: the data tree is stored as
JSON next to its arrays in one uncompressed ``.npz`` (zip) archive, so that
loading a restart file executes nothing from it.  Device tensors are copied
to the host on write and come back on the current GPU (or the CPU when none
is present) on read; ``ConservedVars`` and box meshes keep their types.
"""

from __future__ import annotations

import dataclasses
import io
import json
import os
import zipfile

import numpy as np


class PathError(RuntimeError):
    """mirgecom/restart.py:48-51"""


_META = "__restart_tree__.json"


def _encode(obj, arrays):
    from mirgecom_amd.discretization import BoxMesh
    from mirgecom_amd.fluid import ConservedVars

    def ref(a):
        arrays.append(np.ascontiguousarray(a))
        return len(arrays) - 1

    if obj is None:
        return {"k": "none"}
    if isinstance(obj, (bool, np.bool_)):
        return {"k": "bool", "v": bool(obj)}
    if isinstance(obj, (int, np.integer)):
        return {"k": "int", "v": int(obj)}
    if isinstance(obj, (float, np.floating)):
        return {"k": "float", "v": float(obj).hex()}
    if isinstance(obj, str):
        return {"k": "str", "v": obj}
    if isinstance(obj, ConservedVars):
        return {"k": "cv", "dim": obj.dim, "q": _encode(obj.q, arrays)}
    if isinstance(obj, BoxMesh):
        return {"k": "boxmesh", "fields": {
            f.name: _encode(getattr(obj, f.name), arrays)
            for f in dataclasses.fields(obj)}}
    try:
        import torch
        if isinstance(obj, torch.Tensor):
            return {"k": "tensor", "cuda": obj.device.type == "cuda",
                    "a": ref(obj.detach().cpu().numpy())}
    except ImportError:
        pass
    if isinstance(obj, np.ndarray):
        if obj.dtype == object:
            return {"k": "objarray", "shape": list(obj.shape),
                    "items": [_encode(x, arrays) for x in obj.ravel()]}
        return {"k": "ndarray", "a": ref(obj)}
    if isinstance(obj, (list, tuple)):
        return {"k": "list" if isinstance(obj, list) else "tuple",
                "items": [_encode(x, arrays) for x in obj]}
    if isinstance(obj, dict):
        return {"k": "dict", "items": [[_encode(key, arrays), _encode(v, arrays)]
                                       for key, v in obj.items()]}
    raise TypeError(f"cannot write {type(obj).__name__} to a restart file")


def _decode(node, arrays, device):
    from mirgecom_amd.discretization import BoxMesh
    from mirgecom_amd.fluid import ConservedVars

    k = node["k"]
    if k == "none":
        return None
    if k in ("bool", "int", "str"):
        return node["v"]
    if k == "float":
        return float.fromhex(node["v"])
    if k == "ndarray":
        return arrays[node["a"]]
    if k == "tensor":
        import torch
        t = torch.from_numpy(arrays[node["a"]])
        return t.to(device) if node["cuda"] and device is not None else t
    if k == "cv":
        return ConservedVars(node["dim"], _decode(node["q"], arrays, device))
    if k == "boxmesh":
        return BoxMesh(**{name: _decode(v, arrays, device)
                          for name, v in node["fields"].items()})
    if k == "objarray":
        out = np.empty(len(node["items"]), dtype=object)
        for i, x in enumerate(node["items"]):
            out[i] = _decode(x, arrays, device)
        return out.reshape(node["shape"])
    if k in ("list", "tuple"):
        items = [_decode(x, arrays, device) for x in node["items"]]
        return items if k == "list" else tuple(items)
    if k == "dict":
        return {_decode(key, arrays, device): _decode(v, arrays, device)
                for key, v in node["items"]}
    raise ValueError(f"unknown restart entry kind {k!r}")


def write_restart_file(actx, restart_data, filename, comm=None):
    """mirgecom/restart.py:62-78: write ``restart_data`` (a dictionary) to
    ``filename``; rank 0 creates the directory.  ``actx`` is accepted for
    the reference's signature and not used."""
    rank = comm.Get_rank() if comm is not None else 0
    if rank == 0:
        rst_dir = os.path.dirname(filename)
        if rst_dir:
            if os.path.exists(rst_dir) and not os.path.isdir(rst_dir):
                raise PathError(f"{rst_dir} exists and is not a directory.")
            os.makedirs(rst_dir, exist_ok=True)
    if comm is not None:
        comm.barrier()
    arrays = []
    tree = _encode(restart_data, arrays)
    with zipfile.ZipFile(filename, "w", compression=zipfile.ZIP_STORED) as zf:
        zf.writestr(_META, json.dumps(tree))
        for i, a in enumerate(arrays):
            buf = io.BytesIO()
            np.save(buf, a, allow_pickle=False)
            zf.writestr(f"a{i}.npy", buf.getvalue())


def read_restart_data(actx, filename):
    """mirgecom/restart.py:55-59: the dictionary written by
    ``write_restart_file``."""
    device = None
    try:
        import torch
        if torch.cuda.is_available():
            device = torch.device("cuda", torch.cuda.current_device())
    except ImportError:
        pass
    with zipfile.ZipFile(filename, "r") as zf:
        tree = json.loads(zf.read(_META))
        names = set(zf.namelist())
        arrays = []
        while f"a{len(arrays)}.npy" in names:
            arrays.append(np.load(io.BytesIO(zf.read(f"a{len(arrays)}.npy")),
                                  allow_pickle=False))
    return _decode(tree, arrays, device)


__all__ = ["PathError", "read_restart_data", "write_restart_file"]
