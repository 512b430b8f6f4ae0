"""Pins of the surrogate-network oracle (oracle/mlp.py): its reverse-mode
gradient against central finite differences of the FP64 loss, and its
bfloat16 rounding against exactly representable cases."""

import numpy as np
import pytest

from oracle import mlp


def _layers(rng, sizes):
    out = []
    for kin, nout in zip(sizes[:-1], sizes[1:]):
        out.append((rng.normal(0, 1 / np.sqrt(kin), (kin, nout)),
                    rng.normal(0, 0.1, nout)))
    return out


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    layers = _layers(rng, [32, 12, 12, 12, 16])
    x = rng.normal(size=(20, 32))
    tgt = rng.normal(size=(20, 16))
    loss, grads = mlp.loss_and_grad(layers, x, tgt)
    eps = 1e-6
    for li, (w, b) in enumerate(layers):
        for arr, g in ((w, grads[li][0]), (b, grads[li][1])):
            for _ in range(6):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + eps
                lp, _ = mlp.loss_and_grad(layers, x, tgt)
                arr[idx] = old - eps
                lm, _ = mlp.loss_and_grad(layers, x, tgt)
                arr[idx] = old
                fd = (lp - lm) / (2 * eps)
                assert abs(fd - g[idx]) <= 1e-7 + 1e-6 * abs(fd), (li, idx)
    # padded outputs (7..15) carry no loss
    assert np.all(grads[-1][0][:, 7:] == 0) and np.all(grads[-1][1][7:] == 0)


def test_gelu_derivative():
    z = np.linspace(-6, 6, 101)
    fd = (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6
    assert np.abs(fd - mlp.gelu_grad(z)).max() < 1e-8
    assert mlp.gelu(np.array(0.0)) == 0.0
    assert abs(mlp.gelu(np.array(10.0)) - 10.0) < 1e-12


def test_bf16_rounding():
    # 8 significant bits (7 stored): 1 + 2^-7 is representable, 1 + 2^-8 is
    # a tie
    # (rounds to the even 1.0), 1 + 3 * 2^-8 ties up to 1 + 2^-6
    assert mlp.to_bf16(1.0 + 2.0**-7) == 1.0 + 2.0**-7
    assert mlp.to_bf16(1.0 + 2.0**-8) == 1.0
    assert mlp.to_bf16(1.0 + 3 * 2.0**-8) == 1.0 + 2.0**-6
    assert mlp.to_bf16(1.0 + 2.0**-8 + 2.0**-12) == 1.0 + 2.0**-7
    assert mlp.to_bf16(-2.5) == -2.5
    r = np.random.default_rng(1).normal(size=1000) * 100
    rel = np.abs(mlp.to_bf16(r) - r) / np.abs(r)
    assert rel.max() <= 2.0**-8 * (1 + 1e-6)
