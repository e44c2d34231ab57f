"""Pins the oracle's randomness and deterministic math (CPU).

Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11) and
64 vectors written by rocRAND's open-source philox4x32_10 engine
(tests/golden/philox_rocrand.json, made by tests/golden/make_golden.py).
exp/log/sin/log-factorial: against numpy/scipy to a few ulp.
Poisson: moments and chi-square against scipy.stats.poisson on both branches.
Stochastic rounding: exact on integers, unbiased on fractions.
The workload generators' numpy Philox is pinned the same way.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest
from scipy import special, stats

from arbo_gpu_fit_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))

KATS = [  # Random123 kat_vectors, philox4x32 10 rounds
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def oracle_philox(lib, ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.oracle_philox(c, k, o)
    return tuple(o)


def test_philox_known_answers(oracle_lib):
    for ctr, key, out in KATS:
        assert oracle_philox(oracle_lib, ctr, key) == out


def test_philox_matches_rocrand(oracle_lib):
    vecs = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    assert len(vecs) == 64
    for v in vecs:
        assert oracle_philox(oracle_lib, v["ctr"], v["key"]) == tuple(v["out"]), v


def test_generator_philox_matches(oracle_lib):
    vecs = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    for v in vecs[:16] + [{"ctr": c, "key": k, "out": o} for c, k, o in KATS]:
        r = configs.philox4x32(*v["ctr"], *v["key"])
        assert tuple(int(x) for x in r) == tuple(v["out"])


def ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_exp_log_sin_accuracy(oracle_lib):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-1, 1, 2000), [-744.0, 0.0, 1e-300, 709.0]])
    e = np.array([oracle_lib.oracle_exp(float(x)) for x in xs])
    ref = np.exp(xs)
    ok = ref > 1e-300
    assert ulps(e[ok], ref[ok]).max() <= 4
    ys = np.concatenate([np.exp(rng.uniform(-700, 700, 2000)), rng.uniform(0.5, 2.0, 2000), [1.0, 2.0, 5e-324]])
    lg = np.array([oracle_lib.oracle_log(float(y)) for y in ys])
    lref = np.log(ys)
    assert np.all(np.abs(lg - lref) <= 4 * np.spacing(np.maximum(np.abs(lref), 1e-300)) + 1e-300)
    assert oracle_lib.oracle_log(1.0) == 0.0
    assert oracle_lib.oracle_log(-1.0) == -math.inf
    zs = rng.uniform(-20, 20, 4000)
    sn = np.array([oracle_lib.oracle_sin(float(z)) for z in zs])
    assert np.abs(sn - np.sin(zs)).max() < 2e-15


def test_logfact(oracle_lib):
    ks = np.concatenate([np.arange(0, 40), np.array([100, 1000, 12345, 1e6])])
    got = np.array([oracle_lib.oracle_logfact(float(k)) for k in ks])
    ref = special.gammaln(ks + 1.0)
    assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(1.0, np.abs(ref)))


def poisson_draws(lib, mu, n, seed=5):
    return np.array([lib.oracle_poisson(mu, seed, i, 7, 22) for i in range(n)], np.int64)


@pytest.mark.parametrize("mu", [0.02, 0.7, 3.0, 9.99, 10.0, 25.0, 300.0, 4000.0])
def test_poisson_distribution(oracle_lib, mu):
    n = 20000
    x = poisson_draws(oracle_lib, mu, n)
    # moments within 5 standard errors
    assert abs(x.mean() - mu) < 5 * math.sqrt(mu / n)
    var_se = math.sqrt((mu + 2 * mu * mu) / n)   # se of the sample variance
    assert abs(x.var() - mu) < 5 * var_se + 1e-9
    # chi-square over bins with expected count >= 20
    lo, hi = int(stats.poisson.ppf(1e-4, mu)), int(stats.poisson.ppf(1 - 1e-4, mu))
    edges = np.arange(lo, hi + 2)
    obs = np.histogram(x, bins=edges)[0].astype(float)
    exp = n * (stats.poisson.cdf(edges[1:] - 1, mu) - stats.poisson.cdf(edges[:-1] - 1, mu))
    keep = exp >= 20
    # pool the rest into one cell
    o = np.append(obs[keep], n - obs[keep].sum())
    e = np.append(exp[keep], n - exp[keep].sum())
    chi = ((o - e) ** 2 / e).sum()
    assert stats.chi2.sf(chi, len(o) - 1) > 1e-4


def test_poisson_edge_cases(oracle_lib):
    assert oracle_lib.oracle_poisson(0.0, 1, 0, 0, 0) == 0
    assert oracle_lib.oracle_poisson(-3.0, 1, 0, 0, 0) == 0
    assert oracle_lib.oracle_poisson(float("nan"), 1, 0, 0, 0) == 0
    # a fixed stream is a fixed value
    assert oracle_lib.oracle_poisson(55.5, 9, 1, 2, 3) == oracle_lib.oracle_poisson(55.5, 9, 1, 2, 3)


def test_stochastic_rounding(oracle_lib):
    rng = np.random.default_rng(1)
    rs = rng.integers(0, 2**32, 40000, dtype=np.uint64)
    for x in (0.0, 1.0, 7.0, 365.0):
        assert all(oracle_lib.oracle_sround(x, int(r)) == int(x) for r in rs[:200])
    for x in (0.25, 4.5, 9.9):
        v = np.array([oracle_lib.oracle_sround(x, int(r)) for r in rs])
        assert set(np.unique(v)) <= {math.floor(x), math.ceil(x)}
        frac = x - math.floor(x)
        p = (v == math.ceil(x)).mean()
        assert abs(p - frac) < 5 * math.sqrt(frac * (1 - frac) / rs.size)
    # u = 1 (word 0xFFFFFFFF) rounds up only a fraction of 1
    assert oracle_lib.oracle_sround(2.999999, 0xFFFFFFFF) == 2
    assert oracle_lib.oracle_sround(2.5, 0) == 3
