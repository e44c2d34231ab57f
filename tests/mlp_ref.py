"""numpy reference of the surrogate MLP (test infrastructure).

`forward64` / `backward64`: the network in float64 with no rounding, the
mathematical function the kernels approximate; `forward_bf16`: the same with
the kernels' bfloat16 rounding points (inputs, weights and hidden
activations rounded to bfloat16, sums and biases in float64)."""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(x):
    """float32 -> bfloat16 -> float32, round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward64(layers, x):
    h = np.asarray(x, dtype=np.float64)
    zs, hs = [], [h]
    for l, (W, b) in enumerate(layers):
        z = h @ np.asarray(W, np.float64).T + np.asarray(b, np.float64)
        zs.append(z)
        if l + 1 < len(layers):
            h = gelu(z)
            hs.append(h)
    return zs[-1], zs, hs


def forward_bf16(layers, x):
    h = bf16(x).astype(np.float64)
    for l, (W, b) in enumerate(layers):
        z = h @ bf16(W).astype(np.float64).T + np.asarray(b, np.float64)
        if l + 1 < len(layers):
            h = bf16(gelu(z).astype(np.float32)).astype(np.float64)
    return z


def poisson_loss64(layers, x, y):
    f, _, _ = forward64(layers, x)
    return float(np.mean(np.exp(f) - y * f))


def backward64(layers, x, dout):
    """(weight gradients [(dW, db)], input gradient) of sum(dout * f(x))."""
    f, zs, hs = forward64(layers, x)
    dz = np.asarray(dout, np.float64)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W = np.asarray(layers[l][0], np.float64)
        grads[l] = (dz.T @ hs[l], dz.sum(axis=0))
        dh = dz @ W
        if l > 0:
            dz = dh * gelu_grad(zs[l - 1])
    return grads, dh


def loss_grad64(layers, x, y):
    f, _, _ = forward64(layers, x)
    dout = (np.exp(f) - y) / f.size
    grads, _ = backward64(layers, x, dout)
    return float(np.mean(np.exp(f) - y * f)), np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])


def random_layers(widths, seed):
    rng = np.random.default_rng(seed)
    return [(rng.normal(0, np.sqrt(2.0 / k), (n, k)).astype(np.float32), rng.normal(0, 0.1, n).astype(np.float32))
            for k, n in zip(widths[:-1], widths[1:])]


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers]).astype(np.float32)
