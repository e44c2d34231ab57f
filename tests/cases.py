"""Parity cases shared by the CPU oracle tests and the GPU parity tests.

Each case is a list of simulations and a number of days.  Together they cover
the 10x10-lattice workload (1e5 humans, 365 days, in full), the reference's own parameter set
(von Neumann lattice, unit weights, one serotype, bi-Weibull humans,
Weibull(4, 23 d) mosquitoes, external introductions), isolated and ragged
communities (1 human; 8,192 humans and slots, one wave's bitmap; a community
with no mosquito slots), communities up to the maximum of 65,535 humans and
slots, large Poisson means (the PTRS branch), Beta(2,5) immunity,
both event-count laws with and without a staying fraction, two
simulations of the 256-simulation batch, climate-driven days (per-day
transmission, EIP and mosquito lifetimes, the logistic emergence target),
degree-heterogeneous human and mosquito community sizes, and bursts of
hundreds of human deaths and onsets per community and day.
"""
from __future__ import annotations

import numpy as np

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import SimSpec


def case_lattice10():
    specs, T = configs.lattice10()
    return specs, T


def case_tiny_isolated():
    p = configs.base_params(n_serotypes=1, beta_hm=1.5, beta_mh=1.5, include_self=0, iip_min=1, iip_max=2,
                            recovery=2.0, eip=3.0, seed_frac_h=0.1, seed_frac_m=0.1, external_infection=2.0)
    return [SimSpec(1, 1, np.array([50], np.uint32), p, 11)], 120


def case_reference_params():
    p = configs.reference_params(seed_frac_h=0.02, seed_frac_m=0.02, beta_hm=0.4, beta_mh=0.4)
    return [SimSpec(4, 4, np.full(16, 200, np.uint32), p, 12)], 200


def case_ragged():
    humans = np.array([1, 7, 8192, 13, 64, 65, 255, 256, 257, 1000, 2, 3, 4, 5, 999], np.uint32)
    slots = np.array([0, 16, 8192, 30, 100, 200, 511, 512, 513, 3000, 1, 8, 9, 7, 2500], np.uint32)
    p = configs.base_params(n_serotypes=3, beta_hm=0.8, beta_mh=0.8, seed_frac_h=0.02, seed_frac_m=0.02,
                            long_distance=0.2)
    return [SimSpec(5, 3, humans, p, 13, mosq_slots=slots)], 60


def case_beta25():
    p = configs.base_params(immunity_mode=1)
    return [SimSpec(12, 12, np.full(144, 500, np.uint32), p, 14)], 120


def case_high_lambda():
    p = configs.base_params(beta_hm=0.6, beta_mh=0.6, seed_frac_h=0.05, seed_frac_m=0.05, long_distance=0.05)
    return [SimSpec(3, 3, np.full(9, 2000, np.uint32), p, 15)], 60


def case_stoch_round_moore():
    """The reference's stochastic-rounding counts on a Moore lattice with self
    entries and long-distance events."""
    p = configs.base_params(count_law=abi.COUNT_STOCH_ROUND, long_distance=0.05, seed_frac_h=0.01,
                            seed_frac_m=0.01, beta_hm=0.5, beta_mh=0.5)
    return [SimSpec(6, 5, np.full(30, 700, np.uint32), p, 16)], 90


def case_poisson_stay():
    """Poisson counts with a staying fraction on a lattice without self
    entries, and one with them (staying and dispersed events share the
    source's column)."""
    a = configs.base_params(no_distance=0.3, include_self=0, long_distance=0.02, seed_frac_h=0.01, seed_frac_m=0.01)
    b = configs.base_params(no_distance=0.6, include_self=1, neighbourhood=4, dispersal_scale=0.0,
                            count_law=abi.COUNT_STOCH_ROUND, seed_frac_h=0.01, seed_frac_m=0.01)
    return [SimSpec(5, 5, np.full(25, 600, np.uint32), a, 17), SimSpec(4, 6, np.full(24, 400, np.uint32), b, 18)], 90


def case_climate():
    """The reference's parameter set under a seasonal climate, and the
    configured model under a faster-varying one beside it."""
    T = 200
    a = configs.reference_params(seed_frac_h=0.02, seed_frac_m=0.02)
    ca = configs.seasonal_climate(T + 1, biting_rate=0.5)
    b = configs.base_params(seed_frac_h=0.01, seed_frac_m=0.01)
    cb = configs.seasonal_climate(T + 1, t_amp=6.0, h_amp=0.2, p_amp=0.6, phase=1.0, biting_rate=0.8,
                                  mos_net_growth=0.3, scale_mle=3.0)
    cb.humidity = np.where(np.arange(T + 1) % 3 == 0, cb.humidity, cb.humidity * 0.5).astype(np.float32)
    return [SimSpec(5, 5, np.full(25, 300, np.uint32), a, 31, climate=ca),
            SimSpec(4, 4, np.full(16, 500, np.uint32), b, 32, climate=cb)], T


def case_hetero():
    """Humans drawn over the communities by degree^1.5 and mosquito shares by
    degree^2.5 (network.cpp:138-159), von Neumann lattice."""
    from arbo_gpu_fit_amd.data import community_sizes, mosquito_shares
    humans = community_sizes(6, 5, 15000, 1.5, seed=7, neighbourhood=4)
    p = configs.base_params(neighbourhood=4, seed_frac_h=0.01, seed_frac_m=0.01, beta_hm=0.4, beta_mh=0.4)
    return [SimSpec(6, 5, humans, p, 19, mosq_share=mosquito_shares(6, 5, 2.5, neighbourhood=4))], 90


def case_burst():
    """Days with hundreds of human deaths and E->I onsets in one community:
    8,192 humans with exponential lifetimes of mean 20 days (about 400 deaths
    a day) and a synchronised epidemic (half the mosquitoes infectious at
    t = 0, a fixed 5-day incubation, so thousands of onsets on day 5), which
    fills the demography kernel's queue of deferred human changes several
    times in a day, with one wave per community and with several."""
    humans = np.array([8192, 8192, 5000, 3001], np.uint32)
    slots = np.array([8192, 4096, 8192, 2000], np.uint32)
    p = configs.base_params(n_serotypes=2, h_scale=20.0, h_shape=1.0, seed_frac_h=0.0, seed_frac_m=0.5,
                            beta_hm=0.9, beta_mh=0.9, iip_min=5, iip_max=5, m_ratio=0.5, season_amp=0.0)
    return [SimSpec(2, 2, humans, p, 31, mosq_slots=slots)], 40


def case_large():
    """Communities above one wave's 8,192 bitmap slots: up to the maximum of
    65,535 humans and 65,535 mosquito slots per community, which the day
    kernels handle with eight waves per community (demography) and eight
    bitmap passes (events)."""
    humans = np.array([65535, 20000, 9000, 300], np.uint32)
    slots = np.array([40000, 65535, 8193, 600], np.uint32)
    p = configs.base_params(n_serotypes=2, beta_hm=0.5, beta_mh=0.5, seed_frac_h=0.005, seed_frac_m=0.005,
                            long_distance=0.05)
    return [SimSpec(2, 2, humans, p, 47, mosq_slots=slots)], 30


def case_batch_pair():
    sims = [configs.batch_sim(i) for i in range(24)]
    sims.sort(key=lambda s: int(s.humans.sum()))
    return sims[:2], 40


CASES = {
    "lattice10": case_lattice10,
    "tiny_isolated": case_tiny_isolated,
    "reference_params": case_reference_params,
    "ragged": case_ragged,
    "beta25": case_beta25,
    "high_lambda": case_high_lambda,
    "stoch_round_moore": case_stoch_round_moore,
    "poisson_stay": case_poisson_stay,
    "batch_pair": case_batch_pair,
    "climate": case_climate,
    "hetero": case_hetero,
    "burst": case_burst,
    "large": case_large,
}

# days at which the full census is compared (plus the last day)
CHECK_DAYS = (0, 1, 7)
