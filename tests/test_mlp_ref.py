"""Pins the numpy reference of the surrogate MLP (CPU): its float64 backward
against central finite differences of its float64 forward, for weights and
inputs, and its bfloat16 rounding against exact cases."""
import numpy as np

from tests.mlp_ref import backward64, bf16, flatten, forward64, loss_grad64, poisson_loss64, random_layers


def unflatten(p, widths):
    out, o = [], 0
    for k, n in zip(widths[:-1], widths[1:]):
        W = p[o:o + n * k].reshape(n, k)
        o += n * k
        out.append((W, p[o:o + n]))
        o += n
    return out


def test_backward_matches_finite_differences():
    widths = [6, 9, 7, 5]
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in random_layers(widths, 3)]
    rng = np.random.default_rng(4)
    x = rng.normal(size=(11, 6))
    y = rng.poisson(2.0, size=(11, 5)).astype(np.float64)
    _, g = loss_grad64(layers, x, y)
    p = np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (poisson_loss64(unflatten(p + e, widths), x, y) - poisson_loss64(unflatten(p - e, widths), x, y)) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-9)
    dout = rng.normal(size=(11, 5))
    _, dx = backward64(layers, x, dout)
    fdx = np.zeros_like(x)
    for r in range(x.shape[0]):
        for c in range(x.shape[1]):
            e = np.zeros_like(x)
            e[r, c] = h
            fdx[r, c] = (np.sum(dout * forward64(layers, x + e)[0]) - np.sum(dout * forward64(layers, x - e)[0])) / (2 * h)
    assert np.allclose(dx, fdx, rtol=1e-6, atol=1e-9)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e38, 1e-40], np.float32)
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.0078125   # ties to even
    assert r[4] == -2.5 and np.isfinite(r[5]) and r[6] >= 0
    assert flatten(random_layers([3, 4, 2], 1)).size == 3 * 4 + 4 + 4 * 2 + 2
