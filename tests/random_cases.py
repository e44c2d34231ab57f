"""Seeded random simulations for the parity sweep: each draws its lattice,
community sizes and every switch of the model at once (count law,
staying fraction, neighbourhood, self entries, dispersal weights,
serotypes, immunity mode, introductions, seasonality or a climate,
mosquito shares, lifespans), so combinations no hand-written case covers
are compared bit for bit as well.  Draws use numpy's PCG64 with the case
seed; the same seed gives the same simulations everywhere."""
from __future__ import annotations

import math

import numpy as np

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import ClimateSpec, SimSpec, climate_day
from tests.oracle_lib import load_oracle

N_RANDOM = 12


def random_sim(rng: np.random.Generator, seed: int, days: int) -> SimSpec:
    W, H = int(rng.integers(1, 8)), int(rng.integers(1, 8))
    C = W * H
    humans = rng.integers(1, 900, C).astype(np.uint32)
    if rng.random() < 0.2:
        humans[int(rng.integers(C))] = 1   # a one-human community
    nb = int(rng.choice([4, 8]))
    p = configs.base_params(
        n_serotypes=int(rng.integers(1, 5)),
        beta_hm=float(rng.uniform(0.05, 0.8)), beta_mh=float(rng.uniform(0.05, 0.8)),
        long_distance=float(rng.choice([0.0, 0.01, 0.2])),
        external_infection=float(rng.choice([0.0, 0.0, 3.0])),
        neighbourhood=nb, include_self=int(rng.integers(0, 2)),
        dispersal_scale=float(rng.choice([0.0, 1.0, 2.5])),
        iip_min=int(rng.integers(2, 6)), recovery=float(rng.uniform(2.0, 7.0)), eip=float(rng.uniform(4.0, 12.0)),
        m_scale=float(rng.uniform(6.0, 25.0)), m_shape=float(rng.choice([1.0, 2.0, 4.0])),
        m_ratio=float(rng.uniform(0.5, 3.0)), season_amp=float(rng.uniform(0.0, 0.9)),
        season_phase=float(rng.uniform(0.0, 6.28)),
        seed_frac_h=float(rng.choice([0.0, 0.005, 0.03])), seed_frac_m=float(rng.choice([0.005, 0.03])),
        immunity_mode=int(rng.integers(0, 2)), sero_prev=float(rng.uniform(0.0, 0.4)),
        no_distance=float(rng.choice([0.0, 0.0, 0.5, 0.75])),
        count_law=int(rng.choice([abi.COUNT_POISSON, abi.COUNT_STOCH_ROUND])))
    p.iip_max = p.iip_min + int(rng.integers(0, 4))
    climate = None
    if rng.random() < 0.35:
        climate = configs.seasonal_climate(days + 1, t_mean=float(rng.uniform(20.0, 30.0)),
                                           t_amp=float(rng.uniform(0.0, 6.0)), h_mean=float(rng.uniform(0.4, 0.8)),
                                           h_amp=float(rng.uniform(0.0, 0.15)), p_mean=float(rng.uniform(0.0, 1.0)),
                                           p_amp=float(rng.uniform(0.0, 0.9)) * 0.5,
                                           biting_rate=float(rng.uniform(0.3, 1.2)),
                                           mos_net_growth=float(rng.uniform(0.0, 0.6)),
                                           min_mos_to_human=float(rng.uniform(0.5, 2.0)))
        if p.m_shape == 1.0:
            # exponential lifetimes under a cold climate (scale above about 120
            # days) outlast the 2,047-day lifetime table, which the engine
            # rejects: such a draw gets shape 4 (no extra random draw, so the
            # other cases are unchanged)
            smax = max(climate_day(climate, r, load_oracle())[3] for r in range(1, days + 1))
            if -math.expm1(-2046.0 / smax) < 1.0 - 2.0 ** -24.5:
                p.m_shape = 4.0
    share = None
    if C > 1 and rng.random() < 0.3:
        w = rng.uniform(0.2, 1.0, C)
        share = (w / w.sum()).astype(np.float32)
        share = (share / share.sum(dtype=np.float64)).astype(np.float32)
    # now and then one community of 9,000-11,000 humans (without a climate, whose
    # mosquito capacity is unbounded here): above one wave's 8,192 slots when
    # its mosquito ratio and seasonality ask for it
    big_u, big_c, big_n = rng.random(), int(rng.integers(C)), int(rng.integers(9000, 11001))
    if climate is None and share is None and big_u < 0.2:
        humans[big_c] = big_n
    return SimSpec(W, H, humans, p, seed, climate=climate, mosq_share=share)


def random_case(k: int):
    """Case k: one to three simulations, 40 to 120 days."""
    rng = np.random.Generator(np.random.PCG64(1000 + k))
    days = int(rng.integers(40, 121))
    n = int(rng.integers(1, 4))
    return [random_sim(rng, (k << 8) | i, days) for i in range(n)], days
