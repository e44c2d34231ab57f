"""ESS and R-hat (CPU) against closed forms: AR(1) chains x_t = a x_{t-1} + e_t
have integrated autocorrelation time (1 + a) / (1 - a); independent chains
with the same law have R-hat 1, chains with shifted means do not."""
import numpy as np

from arbo_gpu_fit_amd.diagnostics import circular_unwrap, ess, rhat


def ar1(a, n, m, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, m))
    e = rng.normal(size=(n, m)) * np.sqrt(1 - a * a)
    x[0] = rng.normal(size=m)
    for t in range(1, n):
        x[t] = a * x[t - 1] + e[t]
    return x


def test_ess_of_ar1_chains():
    for a in (0.0, 0.5, 0.9):
        x = ar1(a, 4000, 16, seed=int(a * 10))
        want = x.size * (1 - a) / (1 + a)
        assert abs(ess(x) / want - 1) < 0.15, (a, ess(x), want)


def test_rhat():
    x = ar1(0.5, 2000, 8, seed=3)
    assert rhat(x) < 1.01
    y = x + np.arange(8)[None, :] * 0.5
    assert rhat(y) > 1.1


def test_circular_unwrap_removes_the_wrap():
    rng = np.random.default_rng(2)
    s = (rng.normal(0.0, 0.2, size=(500, 4))) % (2 * np.pi)        # centred on the wrap point
    u = circular_unwrap(s, 0.0, 2 * np.pi)
    assert abs(u.mean()) < 0.05 and abs(u.std() - 0.2) < 0.02
