"""The surrogate MLP on the GPU (bfloat16 MFMA with each operand split into
two bfloat16 halves, float32 accumulation) against the float64 numpy
network with the same float32 parameters: forward, and the backward
(weights and inputs) against the exact float64 gradient element by element
and against central finite differences along the gradient and along random
directions, relative error < 1e-2 each; training on data from a known mean
function."""
import numpy as np
import pytest

from arbo_gpu_fit_amd.surrogate import Surrogate
from tests.mlp_ref import backward64, flatten, forward64, loss_grad64, poisson_loss64, random_layers

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
    assert rel(out, forward64(layers, x)[0]) < 1e-4
    # the full-size network of the fit
    big = [6, 256, 256, 256, 256, 260]
    s2, l2 = _net(big, 3, hip_lib)
    x2 = np.random.default_rng(4).normal(size=(130, 6)).astype(np.float32)
    out2 = s2.forward(x2)
    print(f"[sur] 6-256^4-260 forward: relative error {rel(out2, forward64(l2, x2)[0]):.2e}", flush=True)
    assert rel(out2, forward64(l2, x2)[0]) < 1e-4
    s.close()
    s2.close()


@pytest.mark.parametrize("widths", [[6, 64, 64, 30], [6, 256, 256, 256, 256, 260]])
def test_backward_against_finite_differences(hip_lib, widths):
    print(f"[sur] backward {widths}", flush=True)
    s, layers = _net(widths, 5, hip_lib)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(96, 6)).astype(np.float32)
    f64 = forward64(layers, x)[0]
    y = rng.poisson(np.exp(np.clip(f64, -5, 5))).astype(np.float32)
    loss, g = s.loss_grad(x, y)
    l64, g64 = loss_grad64(layers, x, y)
    assert abs(loss - l64) < 1e-4 * abs(l64)
    print(f"[sur] weight gradient: norm-wise relative error {rel(g, g64):.2e}", flush=True)
    assert rel(g, g64) < 1e-2
    p = np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers]).astype(np.float64)

    def unflat(q):
        out, o = [], 0
        for k, n in zip(widths[:-1], widths[1:]):
            out.append((q[o:o + n * k].reshape(n, k), q[o + n * k:o + n * k + n]))
            o += n * k + n
        return out

    # central differences of the float64 network along the exact gradient and
    # four random directions: relative error < 1e-2 in every one
    h = 1e-4
    for d in range(5):
        u = g64.astype(np.float64) if d == 0 else rng.normal(size=p.size)
        u /= np.linalg.norm(u)
        fd = (poisson_loss64(unflat(p + h * u), x, y) - poisson_loss64(unflat(p - h * u), x, y)) / (2 * h)
        err = abs(g @ u - fd) / abs(fd)
        print(f"[sur] direction {d}: gpu {g @ u:.6e} fd {fd:.6e} relative error {err:.2e}", flush=True)
        assert err < 1e-2
    # gradient with respect to the inputs
    dout = rng.normal(size=(96, widths[-1])).astype(np.float32)
    out, dx = s.input_grad(x, dout)
    _, dx64 = backward64(layers, x, dout)
    print(f"[sur] input gradient: norm-wise relative error {rel(dx, dx64):.2e}", flush=True)
    assert rel(dx, dx64) < 1e-2
    xs = x.astype(np.float64)
    for d in range(3):
        u = (dx64 / np.linalg.norm(dx64)).ravel() if d == 0 else rng.normal(size=x.size)
        u /= np.linalg.norm(u)
        fdx = (np.sum(dout * forward64(layers, xs + h * u.reshape(x.shape))[0])
               - np.sum(dout * forward64(layers, xs - h * u.reshape(x.shape))[0])) / (2 * h)
        assert abs(dx.ravel() @ u - fdx) < 1e-2 * abs(fdx), d
    assert rel(out, forward64(layers, x)[0]) < 1e-4
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
