"""Build the HIP library ``libmgdg.so`` in-tree for gfx950.

``python -m mirgecom_amd.build`` compiles every ``csrc/*.hip`` file to an
object with hipcc (in parallel: each file is its own device module) and links
them into ``libmgdg.so``.  The library exports the C ABI declared in
``include/mgdg.h``; Python binds it with ctypes (``mirgecom_amd._lib``).
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmgdg.so")
OBJ = os.path.join(HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions (v_pk_add/mul/fma_f32) stay off in the device
    # code: hipcc -O3 otherwise forms them from ordinary float code.  The
    # host pass does not know the feature and warns; the device pass applies
    # it (a float2 multiply-add compiles to one v_pk_fma_f32 without it and
    # to none with it).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # MFMA accumulators in ordinary VGPRs: AGPRs would be allocated on top
    # of the kernels' VGPR peak (unified register file) and cost a wave per
    # SIMD in the mixed-precision kernels
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wno-unused-command-line-argument",
]


def sources():
    return [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))
            if f.endswith(".hip")]


def _jobs():
    env = os.environ.get("MAX_JOBS")
    n = int(env) if env else (os.cpu_count() or 4)
    return max(1, min(n, 16))


def build(verbose: bool = True) -> str:
    """Compile the library (always; the result replaces any previous one)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, *inc, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    srcs = sources()
    # largest files first: they bound the wall time
    srcs.sort(key=lambda s: -os.path.getsize(s))
    with ThreadPoolExecutor(_jobs()) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
