"""Delayed-acceptance MCMC with the GPU surrogate.

* Exactness: the second stage is a known Gaussian target and the surrogate a
  network that knows nothing of it; the chains must still sample that exact
  Gaussian (the first stage only filters proposals), with Langevin sub-steps
  and with the screened adaptive-Metropolis proposal.
* Parity: the same delayed-acceptance run over the GPU engine and over the
  CPU oracle's engine gives identical chains (bit-exact simulations).
"""
import math

import numpy as np
import pytest

from arbo_gpu_fit_amd import abi, configs, mcmc
from arbo_gpu_fit_amd.diagnostics import ess
from arbo_gpu_fit_amd.engine import Engine
from arbo_gpu_fit_amd.surrogate import Surrogate
from tests.test_mcmc import _tiny_driver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [mcmc.DA_LANGEVIN, mcmc.DA_SCREEN])
def test_delayed_acceptance_keeps_the_target(hip_lib, mode):
    lib = mcmc.bind(hip_lib)
    W = 4
    priors = [(0, 0.0, 1.0), (0, 0.0, 1.0), (0, -3.0, -1.0), (0, 0.0, 1.0), (1, 0.0, 2 * math.pi), (0, 0.01, 1.0)]
    mean = np.array([0.5, 0.4, -2.0, 0.6, math.pi, 0.5])
    sd = np.array([0.06, 0.05, 0.15, 0.07, 0.3, 0.08])
    rng = np.random.default_rng(3)
    A = rng.normal(size=(6, 6)) * 0.3 + np.eye(6)
    corr = A @ A.T
    corr = corr / np.sqrt(np.outer(np.diag(corr), np.diag(corr)))
    cov = corr * np.outer(sd, sd)
    prec = np.linalg.inv(cov)

    def target(th):
        d = th - mean
        return -0.5 * np.einsum("ni,ij,nj->n", d, prec, d)

    print(f"[da] exactness mode={mode}", flush=True)
    sur = Surrogate([6, 32, 32, W], max_batch=256, seed=5, lib=hip_lib)
    p = sur.get_params()
    p *= 0.2                                   # a weakly informative, unrelated surrogate
    p[-W:] = math.log(20.0)
    sur.set_params(p)
    cfg = mcmc.McmcConfig(n_chains=64, n_weeks=W, priors=priors, adapt_after=300, seed=8)
    m = mcmc.Mcmc(cfg, 1, 1, [10], configs.base_params(), np.array([3, 9, 7, 12], np.uint32),
                  theta0=np.tile(mean, (64, 1)), lib=lib)
    m.run_da_target(target, sur, 2500, sub_steps=3, step0=0.5, mode=mode)
    th, _, acc = m.history()
    post = th[300:]
    samples = post.reshape(-1, 6)
    n_eff = np.array([ess(post[:, :, i]) for i in range(6)])
    se = sd / np.sqrt(n_eff)
    print(f"[da] ess {n_eff.round()} stage-2 acceptance {acc[300:].mean():.3f} {m.da_counters()}", flush=True)
    assert np.all(np.abs(samples.mean(axis=0) - mean) < 5 * se)
    # the spread too, to the precision the effective sample size allows
    assert np.all(np.abs(samples.std(axis=0) / sd - 1) < 5 / np.sqrt(2 * n_eff))
    c = m.da_counters()
    if mode == mcmc.DA_LANGEVIN:
        assert abs(c["surrogate_accepted"] / c["surrogate_steps"] - 0.574) < 0.15
    else:   # the screen passes some proposals and stops others
        assert 0.05 < c["surrogate_accepted"] / c["surrogate_steps"] < 0.95
    m.close()
    sur.close()


@pytest.mark.parametrize("mode", [mcmc.DA_LANGEVIN, mcmc.DA_SCREEN])
def test_delayed_acceptance_equal_over_gpu_and_oracle(hip_lib, oracle_lib, mode):
    lib = mcmc.bind(hip_lib)
    print("[da] gpu vs oracle", flush=True)
    sur = Surrogate([6, 16, 6], max_batch=64, seed=2, lib=hip_lib)
    p = sur.get_params()
    p[-6:] = math.log(30.0)
    sur.set_params(p)
    histories = []
    for use_gpu in (False, True):
        m, orc = _tiny_driver(lib, oracle_lib, seed=33, n_chains=8, n_particles=1)
        if use_gpu:
            eng = Engine(hip_lib)
            m.run_da(eng, sur, 3, sub_steps=2, mode=mode)
            eng.close()
        else:
            m.run_da(orc, sur, 3, sub_steps=2, api=mcmc.engine_api(oracle_lib), mode=mode)
        histories.append((m.history(), m.counters(), m.da_counters()))
        m.close()
        orc.close()
    (ha, ca, da), (hb, cb, db) = histories
    for x, y in zip(ha, hb):
        assert np.array_equal(x, y)
    assert ca["simulations"] == cb["simulations"] and da["surrogate_accepted"] == db["surrogate_accepted"]
    sur.close()
