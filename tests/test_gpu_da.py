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
    priors = [(0, 0.0, 1.0), (0, 0.0, 1.0), (0, -3.0, -1.0), (0, 0.0, 1.0), (1, 0.0, 2 * math.pi), (0, 0.01, 1.0),
              (mcmc.PRIOR_FIXED, 0.1, 0.0)]
    mean = np.array([0.5, 0.4, -2.0, 0.6, math.pi, 0.5])
    sd = np.array([0.06, 0.05, 0.15, 0.07, 0.3, 0.08])
    rng = np.random.default_rng(3)
    A = rng.normal(size=(6, 6)) * 0.3 + np.eye(6)
    corr = A @ A.T
    corr = corr / np.sqrt(np.outer(np.diag(corr), np.diag(corr)))
    cov = corr * np.outer(sd, sd)
    prec = np.linalg.inv(cov)

    def target(th):
        d = th[:, :6] - mean
        return -0.5 * np.einsum("ni,ij,nj->n", d, prec, d)

    print(f"[da] exactness mode={mode}", flush=True)
    sur = Surrogate([6, 32, 32, W], max_batch=256, seed=5, lib=hip_lib)
    p = sur.get_params()
    p *= 0.2                                   # a weakly informative, unrelated surrogate
    p[-W:] = math.log(20.0)
    sur.set_params(p)
    cfg = mcmc.McmcConfig(n_chains=64, n_weeks=W, priors=priors, adapt_after=300, seed=8)
    m = mcmc.Mcmc(cfg, 1, 1, [10], configs.base_params(), np.array([3, 9, 7, 12], np.uint32),
                  theta0=np.tile(np.append(mean, 0.1), (64, 1)), lib=lib)
    m.run_da_target(target, sur, 2500, sub_steps=3, step0=0.5, mode=mode)
    th, _, acc = m.history()
    assert np.all(th[:, :, 6] == 0.1)
    post = th[300:, :, :6]
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


def _curve(th, W):
    """Log weekly mean before reporting of a smooth stand-in for the simulator:
    level from the two transmission rates and the long-distance fraction, a
    seasonal wave of amplitude A and phase phi."""
    th = np.atleast_2d(th)
    w = np.arange(W)
    level = math.log(2000.0) + 2.5 * (th[:, 0:1] - 0.3) + 1.5 * (th[:, 1:2] - 0.3) + 0.3 * (th[:, 2:3] + 2.0)
    return level + th[:, 3:4] * np.sin(2 * math.pi * w / W + th[:, 4:5])


def test_trained_surrogate_saves_second_stage_evaluations(hip_lib):
    """What the learned stage buys: on a second stage whose cost is its number
    of evaluations, screened delayed acceptance with a surrogate trained on
    prior draws (the MFMA kernels) reaches the same posterior as plain
    adaptive Metropolis with at least twice the effective samples per
    evaluation."""
    from scipy import stats

    lib = mcmc.bind(hip_lib)
    W, r, skip = 26, 1.0 / configs.MCMC_DISPERSION, 1
    priors = configs.mcmc_priors()
    truth = configs.MCMC_TRUTH
    rng = np.random.default_rng(21)
    mu_true = truth[5] * np.exp(_curve(truth, W)[0])
    data = stats.nbinom.rvs(r, r / (r + mu_true), random_state=rng).astype(np.uint32)
    counts = {"plain": 0, "screened": 0}

    def loglik_for(name):
        def f(th):
            counts[name] += len(th)
            mu = th[:, 5:6] * np.exp(_curve(th, W))
            ll = stats.nbinom.logpmf(data[None, skip:], r, r / (r + mu[:, skip:]))
            return np.maximum(ll, -250.0).sum(axis=1)
        return f

    print("[da] surrogate training for the evaluation test", flush=True)
    from arbo_gpu_fit_amd.fit import prior_sample
    from arbo_gpu_fit_amd.surrogate import features
    th_train = prior_sample(4096, seed=5, priors=priors)
    sur = Surrogate([6, 64, 64, W], max_batch=512, seed=9, lib=hip_lib)
    p = sur.get_params()
    y = np.exp(_curve(th_train, W)).astype(np.float32)
    p[-W:] = np.log(y.mean(axis=0))
    sur.set_params(p)
    hist = sur.train(features(th_train, priors), y, epochs=300, batch=256, lr=2e-3, seed=3)
    pred = np.exp(sur.forward(features(th_train[:512], priors)))
    rel = np.abs(pred / y[:512] - 1).mean()
    print(f"[da] surrogate loss {hist[0]:.4g} -> {hist[-1]:.4g}, mean relative error {rel:.3f}", flush=True)
    assert rel < 0.05

    n_it, burn, chains = 800, 200, 32
    theta0 = np.tile(truth, (chains, 1))
    theta0[:, :6] += rng.normal(scale=0.002, size=(chains, 6))
    out = {}
    for name in ("plain", "screened"):
        cfg = mcmc.McmcConfig(n_chains=chains, n_weeks=W, priors=priors, adapt_after=1000, skip_weeks=skip,
                              seed=17)
        m = mcmc.Mcmc(cfg, 1, 1, [10], configs.base_params(), data, theta0=theta0, lib=lib)
        if name == "plain":
            m.run_target(loglik_for(name), n_it)
        else:
            m.run_da_target(loglik_for(name), sur, n_it, mode=mcmc.DA_SCREEN)
        th = m.history()[0][burn:, :, :6]
        th[:, :, 4] = np.mod(th[:, :, 4] - truth[4] + math.pi, 2 * math.pi) + truth[4] - math.pi
        e = np.array([ess(th[:, :, i]) for i in range(6)])
        out[name] = (th.reshape(-1, 6).mean(axis=0), th.reshape(-1, 6).std(axis=0), e, counts[name])
        print(f"[da] {name}: evaluations {counts[name]}, min ESS {e.min():.0f}, "
              f"ESS per 1k evaluations {1000 * e.min() / counts[name]:.2f}", flush=True)
        m.close()
    (ma, sa, ea, na), (mb, sb, eb, nb) = out["plain"], out["screened"]
    # the same posterior, to the precision the effective sample sizes allow
    se = np.sqrt(sa ** 2 / ea + sb ** 2 / eb)
    assert np.all(np.abs(ma - mb) < 5 * se)
    # and the learned stage pays
    assert eb.min() / nb > 2.0 * ea.min() / na
    sur.close()
