"""Build the HIP library ``libmgdg.so`` in-tree for gfx950.

``python -m mirgecom_amd.build`` compiles ``csrc/mgdg.hip`` with hipcc.  The
library exports the C ABI declared in ``include/mgdg.h``; Python binds it
with ctypes (``mirgecom_amd._lib``).
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmgdg.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (they are generated from ordinary
    # float code otherwise)
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


def build(verbose: bool = True) -> str:
    """Compile the library (always; the result replaces any previous one)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"),
           *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
