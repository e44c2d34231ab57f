"""The surrogate MLP on the GPU (bfloat16 MFMA, float32 accumulation) against
the numpy reference: forward with the kernels' rounding points; the backward
(weights and inputs) against central finite differences of the float64
network at the kernels' operating point (bfloat16 weights and inputs) along
the exact gradient (relative error < 1e-2) and along random directions, and
against the exact float64 gradient in direction and length (element-wise,
bfloat16 hidden activations and gradients leave a few per cent on weight
gradients summed over a batch); training on data from a known mean
function."""
import numpy as np
import pytest

from arbo_gpu_fit_amd.surrogate import Surrogate
from tests.mlp_ref import (backward64, bf16, flatten, forward64, forward_bf16, loss_grad64, poisson_loss64,
                           random_layers)

pytestmark = pytest.mark.gpu


def _net(widths, seed, hip_lib):
    s = Surrogate(widths, max_batch=512, seed=seed, lib=hip_lib)
    layers = random_layers(widths, seed)
    s.set_params(flatten(layers))
    return s, layers


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_forward_matches_reference(hip_lib):
    widths = [6, 64, 96, 40]
    print("[sur] forward", flush=True)
    s, layers = _net(widths, 1, hip_lib)
    x = np.random.default_rng(2).normal(size=(77, 6)).astype(np.float32)
    out = s.forward(x)
    assert rel(out, forward_bf16(layers, x)) < 2e-3
    assert rel(out, forward64(layers, x)[0]) < 2e-2
    # the full-size network of the fit
    big = [6, 256, 256, 256, 256, 260]
    s2, l2 = _net(big, 3, hip_lib)
    x2 = np.random.default_rng(4).normal(size=(130, 6)).astype(np.float32)
    assert rel(s2.forward(x2), forward_bf16(l2, x2)) < 2e-3
    s.close()
    s2.close()


def test_backward_against_finite_differences(hip_lib):
    widths = [6, 64, 64, 30]
    print("[sur] backward", flush=True)
    s, layers = _net(widths, 5, hip_lib)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(96, 6)).astype(np.float32)
    f64 = forward64(layers, x)[0]
    y = rng.poisson(np.exp(np.clip(f64, -5, 5))).astype(np.float32)
    loss, g = s.loss_grad(x, y)
    # the float64 network at the point the kernels evaluate: weights and
    # inputs rounded to bfloat16 (the kernels' operands), everything else exact
    layers = [(bf16(W).astype(np.float64), b.astype(np.float64)) for W, b in layers]
    x = bf16(x)
    l64, g64 = loss_grad64(layers, x, y)
    assert abs(loss - l64) < 1e-2 * abs(l64)
    # element-wise the bfloat16 operands leave a few per cent on weight
    # gradients (sums over the batch of mixed-sign terms): same direction,
    # same length
    cos = float(g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64)))
    print(f"[sur] weight gradient: norm-wise rel err {rel(g, g64):.4f}, cosine {cos:.5f}", flush=True)
    assert cos > 0.995 and abs(np.linalg.norm(g) / np.linalg.norm(g64) - 1) < 0.02
    # central differences of the float64 network along the exact gradient
    # (relative error < 1e-2) and along random directions
    p = np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])

    def unflat(q):
        out, o = [], 0
        for k, n in zip(widths[:-1], widths[1:]):
            out.append((q[o:o + n * k].reshape(n, k), q[o + n * k:o + n * k + n]))
            o += n * k + n
        return out

    h = 1e-4
    for d in range(4):
        u = g64.astype(np.float64) if d == 0 else rng.normal(size=p.size)
        u /= np.linalg.norm(u)
        fd = (poisson_loss64(unflat(p + h * u), x, y) - poisson_loss64(unflat(p - h * u), x, y)) / (2 * h)
        err = abs(g @ u - fd)
        print(f"[sur] direction {d}: gpu {g @ u:.6e} fd {fd:.6e}", flush=True)
        if d == 0:
            assert err < 1e-2 * abs(fd)
        else:
            assert err < 0.1 * abs(fd) + 0.02 * np.linalg.norm(g64) / np.sqrt(p.size) * 5
    # gradient with respect to the inputs
    dout = rng.normal(size=(96, 30)).astype(np.float32)
    out, dx = s.input_grad(x, dout)
    _, dx64 = backward64(layers, x, dout)
    u = (dx64 / np.linalg.norm(dx64)).ravel()
    xs = x.astype(np.float64)
    fdx = (np.sum(dout * forward64(layers, xs + h * u.reshape(x.shape))[0])
           - np.sum(dout * forward64(layers, xs - h * u.reshape(x.shape))[0])) / (2 * h)
    print(f"[sur] input gradient: norm-wise rel err {rel(dx, dx64):.4f}; along it gpu {dx.ravel() @ u:.6e} fd {fdx:.6e}",
          flush=True)
    assert abs(dx.ravel() @ u - fdx) < 1e-2 * abs(fdx)
    assert float(dx.ravel() @ dx64.ravel() / (np.linalg.norm(dx) * np.linalg.norm(dx64))) > 0.995
    assert rel(out, forward64(layers, x)[0]) < 1e-2
    s.close()


def test_training_reaches_the_true_mean(hip_lib):
    widths = [6, 64, 64, 20]
    print("[sur] training", flush=True)
    truth = random_layers(widths, 11)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, size=(4096, 6)).astype(np.float32)
    mu = np.exp(np.clip(forward64(truth, x)[0] * 0.5 + 2.0, -3, 6))
    y = rng.poisson(mu).astype(np.float32)
    floor = float(np.mean(mu - y * np.log(mu)))        # loss of the true mean function
    s = Surrogate(widths, max_batch=512, seed=13, lib=hip_lib)
    p = s.get_params()
    p[-20:] = np.log(y.mean(axis=0) + 1e-3)             # output bias at the data's log-mean
    s.set_params(p)
    hist = s.train(x, y, epochs=60, batch=256, lr=2e-3, seed=3)
    assert np.all(np.isfinite(hist))
    start = float(np.mean(np.exp(p[-20:]) - y * p[-20:]))
    assert hist[-1] < floor + 0.15 * (start - floor), (start, hist[-1], floor)
    s.close()
