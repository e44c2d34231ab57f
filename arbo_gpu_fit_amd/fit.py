"""The fitting workflow around the engine: training data for the surrogate
from simulator batches, surrogate training, and the plain and
delayed-acceptance MCMC fits with their efficiency summaries.

All heavy work runs in the C++/HIP library: simulations on the engine,
the MLP's matrix products on the MFMA units, the chains in the C++ driver;
this module sequences them and summarises the output."""
from __future__ import annotations

import dataclasses
import math
import sys
import time

import numpy as np

from . import abi, configs, mcmc
from .diagnostics import circular_unwrap, ess, rhat
from .engine import Engine, SimSpec
from .surrogate import Surrogate, features

SURROGATE_WIDTHS = [6, 256, 256, 256, 256, configs.MCMC_WEEKS]


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def prior_sample(n: int, seed: int, priors=None) -> np.ndarray:
    """n parameter vectors from the prior (numpy PCG64 with `seed`)."""
    priors = priors or configs.mcmc_priors()
    rng = np.random.Generator(np.random.PCG64(seed))
    th = np.zeros((n, mcmc.D))
    for i, (k, a, b) in enumerate(priors):
        if k == mcmc.PRIOR_FIXED:
            th[:, i] = a
        elif k == mcmc.PRIOR_BETA:
            th[:, i] = rng.beta(a, b, n)
        elif k == mcmc.PRIOR_TRUNC_NORMAL:   # N(a, b^2) on (0, TRUNC_HI] by rejection
            v = rng.normal(a, b, 4 * n + 64)
            v = v[(v > 0.0) & (v <= mcmc.TRUNC_HI)]
            while v.size < n:
                w = rng.normal(a, b, 4 * n + 64)
                v = np.concatenate([v, w[(w > 0.0) & (w <= mcmc.TRUNC_HI)]])
            th[:, i] = v[:n]
        else:
            th[:, i] = rng.uniform(a, b, n)
    return th


def model_spec(lib, theta, seed: int, g: int, humans, base, param_set: int = mcmc.FIT_SEASONAL,
               climate=None) -> SimSpec:
    """The simulation of one parameter vector: the seasonal set's five
    parameters in ArboParams, or the climate set's five coefficients in the
    climate (MCMC::copyParams, MCMC.cpp:79-86)."""
    if param_set == mcmc.FIT_CLIMATE:
        t = [float(x) for x in theta]
        cl = dataclasses.replace(climate, scale_eip=t[0], scale_mle=t[1], scale_cc=t[2], scale_humidity=t[3],
                                 min_mos_to_human=t[4])
        return SimSpec(g, g, humans, base, seed, climate=cl)
    return SimSpec(g, g, humans, mcmc.sim_params(lib, theta, base), seed, climate=climate)


def simulate_weekly(engine: Engine, thetas: np.ndarray, seed: int, batch: int = 256,
                    n_weeks: int = configs.MCMC_WEEKS, param_set: int = mcmc.FIT_SEASONAL, climate=None,
                    shape=None, lib=None) -> np.ndarray:
    """Weekly incidence [n, n_weeks] of one simulation per vector of the MCMC
    model (1e6 humans, 32 x 32 communities, or shape = (grid, humans)),
    `batch` simulations per engine batch; simulation i has Philox seed
    (i << 32) | seed.  lib: the library of the parameter mapping and weekly
    binning (default: the engine's)."""
    lib = mcmc.bind(lib if lib is not None else engine.lib)
    g, humans = shape if shape is not None else configs.mcmc_shape()
    base = configs.base_params()
    days = 7 * n_weeks - 1
    out = np.zeros((len(thetas), n_weeks), np.uint32)
    t0 = time.perf_counter()
    for b0 in range(0, len(thetas), batch):
        th = thetas[b0:b0 + batch]
        specs = [model_spec(lib, t, ((b0 + i) << 32) | (seed & 0xFFFFFFFF), g, humans, base, param_set, climate)
                 for i, t in enumerate(th)]
        engine.init(specs, days)
        engine.step(days)
        for i in range(len(th)):
            out[b0 + i] = mcmc.weekly_incidence(lib, engine.series(i)["n_incidence"], n_weeks)
        log(f"[fit] simulated {b0 + len(th)}/{len(thetas)} runs, {time.perf_counter() - t0:.1f} s")
    return out


def train_surrogate(thetas: np.ndarray, weekly: np.ndarray, epochs: int = 200, batch: int = 512,
                    lr: float = 1e-3, seed: int = 1, lib=None, stage: int = 10,
                    priors=None) -> tuple[Surrogate, np.ndarray]:
    """The MLP (6 -> 256^4 -> n_weeks, GELU) fitted to the weekly incidence
    by the Poisson loss (Adam), its output bias started at the data's
    log-mean per week.  The learning rate follows a cosine decay from `lr` to
    lr / 100, changed every `stage` epochs (the optimiser state carries over
    between stages)."""
    widths = list(SURROGATE_WIDTHS[:-1]) + [weekly.shape[1]]
    sur = Surrogate(widths, max_batch=max(batch, 4096), seed=seed, lib=lib)
    x = features(thetas, priors or configs.mcmc_priors()).astype(np.float32)
    y = weekly.astype(np.float32)
    p = sur.get_params()
    p[-weekly.shape[1]:] = np.log(y.mean(axis=0) + 1e-2)
    sur.set_params(p)
    hist = []
    for e0 in range(0, epochs, stage):
        n = min(stage, epochs - e0)
        frac = (e0 + 0.5 * n) / epochs
        rate = lr * (0.01 + 0.99 * 0.5 * (1.0 + math.cos(math.pi * frac)))
        hist.append(sur.train(x, y, epochs=n, batch=batch, lr=rate, seed=seed + e0))
    return sur, np.concatenate(hist)


def summarise(m: mcmc.Mcmc, wall: float, burn: int, names=mcmc.PARAM_NAMES) -> dict:
    """ESS, R-hat, acceptance and posterior moments over the iterations after
    `burn`, over the fitted parameters (fixed ones left out)."""
    th, ll, acc = m.history()
    post = th[burn:].copy()
    for i, (k, a, b) in enumerate(m.cfg.priors):   # circular parameters unwrapped before ESS and R-hat
        if k == mcmc.PRIOR_CIRCULAR:
            post[:, :, i] = circular_unwrap(post[:, :, i], a, b)
    fitted = [i for i, (k, _, _) in enumerate(m.cfg.priors) if k != mcmc.PRIOR_FIXED]
    names = [names[i] for i in fitted]
    e = [ess(post[:, :, i]) for i in fitted]
    r = [rhat(post[:, :, i]) for i in fitted]
    n = th.shape[0]
    sampling_wall = wall * (n - burn) / n
    return {
        "iterations": n,
        "burn_in": burn,
        "wall_s": wall,
        "s_per_iteration": wall / n,
        "ess_min": min(e),
        "ess": dict(zip(names, e)),
        "ess_per_hour": min(e) / (sampling_wall / 3600.0),
        "rhat": dict(zip(names, r)),
        "rhat_max": max(r),
        "acceptance": float(acc[burn:].mean()),
        "posterior_mean": dict(zip(names, th[burn:].reshape(-1, mcmc.D)[:, fitted].mean(axis=0).tolist())),
        "posterior_sd": dict(zip(names, th[burn:].reshape(-1, mcmc.D)[:, fitted].std(axis=0).tolist())),
        "counters": m.counters(),
    }


def run_fit(engine: Engine, data: np.ndarray, iters: int, chains: int = 64, seed: int = 1, sur: Surrogate | None = None,
            sub_steps: int = 10, theta0=None, burn: int | None = None, chunk: int = 10,
            da_mode: int = mcmc.DA_LANGEVIN, param_set: int = mcmc.FIT_SEASONAL, climate=None,
            shape=None, api=None, lib=None, n_particles: int = 1, n_blocks: int = 0,
            priors=None) -> tuple[mcmc.Mcmc, dict]:
    """A plain (sur None) or delayed-acceptance fit of `data` (first stage
    `da_mode`); returns the driver and its summary over the iterations after
    `burn`.  param_set FIT_CLIMATE fits the reference's own parameters
    (climate coefficients, configs.climate_priors()) under `climate`.  api:
    mcmc.engine_api of the engine's library when it is not the driver's
    library `lib` (default: the engine's) (plain fits).  n_particles
    simulations per proposal (the reference's C_MAXSIMRUN is 30,
    constant.h:8) and n_blocks pseudo-marginal seed blocks (0: fresh seeds
    for every proposal)."""
    g, humans = shape if shape is not None else configs.mcmc_shape()
    fit_climate = param_set == mcmc.FIT_CLIMATE
    if priors is None:
        priors = configs.climate_priors() if fit_climate else configs.mcmc_priors()
    names = mcmc.CLIMATE_PARAM_NAMES if fit_climate else mcmc.PARAM_NAMES
    cfg = mcmc.McmcConfig(n_chains=chains, n_weeks=len(data), priors=priors, adapt_after=100 * mcmc.D, seed=seed,
                          param_set=param_set, n_particles=n_particles, n_blocks=n_blocks)
    m = mcmc.Mcmc(cfg, g, g, humans, configs.base_params(), data, theta0=theta0,
                  lib=lib if lib is not None else engine.lib, climate=climate)
    t0 = time.perf_counter()
    done = 0
    while done < iters:
        n = min(chunk, iters - done)
        if sur is None:
            m.run(engine, n, api=api)
        else:
            m.run_da(engine, sur, n, sub_steps=sub_steps, mode=da_mode)
        done += n
        kind = "plain" if sur is None else ("screened" if da_mode == mcmc.DA_SCREEN else "delayed-acceptance")
        log(f"[fit] {kind} {done}/{iters} iterations, "
            f"{time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    out = summarise(m, wall, iters // 4 if burn is None else burn, names)
    if sur is not None:
        out["delayed_acceptance"] = m.da_counters()
    return m, out
