"""Oracle of the chemistry-surrogate network (NumPy FP64; test
infrastructure only).

The surrogate has no counterpart in the reference; this is the CPU
statement of the network that mirgecom_amd/csrc/mg_chem.hip runs in BF16 on
the matrix cores, used to check its forward pass and its gradients:

* layers 32 -> 256 -> 256 -> 256 -> 32 (10 and 7 real inputs / outputs),
  ``z_l = a_{l-1} W_l + b_l``, ``a_l = gelu(z_l)`` with the tanh form
  ``gelu(z) = z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) / 2`` on the
  hidden layers;
* loss ``L = sum_{k<7} (out_k - target_k)^2 / (7 B)``;
* reverse mode: ``dz_4 = 2 (out - target) / (7 B)``,
  ``dW_l = a_{l-1}^T dz_l``, ``db_l = sum_b dz_l``,
  ``dz_{l-1} = (dz_l W_l^T) gelu'(z_{l-1})``.

With ``bf16=True`` the inputs, weights, hidden activations and backward
signals are rounded to bfloat16 (round to nearest, ties to even) where the
kernels store them in BF16, so the comparison isolates the FP32
accumulation.  Pinned by central finite differences of the FP64 loss
(tests/test_oracle_mlp.py).
"""

from __future__ import annotations

import math

import numpy as np

N_OUT = 7


def to_bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


C0 = math.sqrt(2.0 / math.pi)
C1 = 0.044715


def gelu(z):
    """GELU, tanh form: z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) / 2."""
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + C1 * z**3)))


def gelu_grad(z):
    t = np.tanh(C0 * (z + C1 * z**3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C0 * (1.0 + 3.0 * C1 * z * z)


def _r(a, bf16):
    return to_bf16(a) if bf16 else np.asarray(a, dtype=np.float64)


def forward(layers, x, bf16=False):
    """``layers = [(W [K][N], b [N]) x 4]``; returns ``(out, cache)``."""
    a = _r(x, bf16)
    cache = {"a": [a], "z": []}
    for li, (w, b) in enumerate(layers):
        z = a @ _r(w, bf16) + b
        if li < len(layers) - 1:
            if bf16:
                z = np.asarray(z, dtype=np.float32).astype(np.float64)
            cache["z"].append(z)
            a = _r(gelu(z), bf16)
            cache["a"].append(a)
        else:
            out = z
    return out, cache


def loss_and_grad(layers, x, target, bf16=False):
    """``(L, [(dW, db) x 4])`` of a batch."""
    out, cache = forward(layers, x, bf16)
    bsz = x.shape[0]
    diff = out - target
    diff[:, N_OUT:] = 0.0
    loss = float(np.sum(diff[:, :N_OUT] ** 2) / (N_OUT * bsz))
    dz = _r(2.0 * diff / (N_OUT * bsz), bf16)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        a_in = cache["a"][li]
        grads[li] = (a_in.T @ dz, dz.sum(axis=0))
        if li == 0:
            break
        w = _r(layers[li][0], bf16)
        dz = _r((dz @ w.T) * gelu_grad(cache["z"][li - 1]), bf16)
    return loss, grads
