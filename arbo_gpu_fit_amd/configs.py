"""Workload generators: the five configured workloads of the dengue IBM.

  0  plain single simulation, 10x10 lattice, 1,000 humans per community, 365 days
  1  plain single simulation, 100x100 lattice, 1,000 humans per community,
     3,650 days, Beta(2,5) initial immunity per community and serotype
  2  256 simulations in one batch: communities log-uniform 400..10,000
     (squared like SimulationParam::squareLattice), humans per community
     U{200..2000}, beta_hm, beta_mh ~ U[0.05,0.6], long-distance ~ LogU[1e-3,1e-1],
     seasonal amplitude ~ U[0,0.9], phase ~ U[0,2pi), 1,825 days
  3  MCMC model: 1e6 humans in 1,000 -> 32x32 communities, 1,825 days
  4  as 3, for the surrogate's training runs

Common parameters (SimulationParam.cpp:11-47 with the configured values):
human lifespan Weibull with mean 70 y and shape 6, mosquito lifespan
exponential with mean 10 d, 4 serotypes, intrinsic incubation U{4..7} d,
human recovery 5 d, extrinsic incubation 10 d, beta_hm = beta_mh = 0.3,
long-distance fraction 0.01, Moore-8 + self dispersal with weights exp(-d),
mosquito density 2 x (1 + A sin(2 pi t/365 + phi)) per human, 0.1% of humans
and mosquitoes infectious at t=0.

Random draws of the generators use Philox4x32-10 keyed by the workload seed
with counter (simulation, field, element, 0).
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .engine import ClimateSpec, SimSpec

# --- Philox4x32-10 over numpy uint64 lanes ---------------------------------

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox4x32-10; counters are integer arrays (or scalars)."""
    x = [np.asarray(v, dtype=np.uint64) & _MASK for v in (c0, c1, c2, c3)]
    x = np.broadcast_arrays(*x)
    x = [a.copy() for a in x]
    ka, kb = k0 & _MASK, k1 & _MASK
    for _ in range(10):
        p0 = x[0] * np.uint64(_M0)
        p1 = x[2] * np.uint64(_M1)
        n0 = (p1 >> np.uint64(32)) ^ x[1] ^ np.uint64(ka)
        n1 = p1 & np.uint64(_MASK)
        n2 = (p0 >> np.uint64(32)) ^ x[3] ^ np.uint64(kb)
        n3 = p0 & np.uint64(_MASK)
        x = [n0, n1, n2, n3]
        ka = (ka + _W0) & _MASK
        kb = (kb + _W1) & _MASK
    return [a.astype(np.uint32) for a in x]


def _u53(a, b):
    return ((a.astype(np.float64) // 32) * 67108864.0 + (b.astype(np.float64) // 64)) * 2.0 ** -53


def square_lattice(meta_pop: float) -> int:
    """SimulationParam::squareLattice (SimulationParam.cpp:50-88): nearest
    square, ties to the smaller one, perfect squares unchanged."""
    up = int(math.ceil(meta_pop ** 0.5)) ** 2
    down = int(math.floor(meta_pop ** 0.5)) ** 2
    up_diff = abs(meta_pop - up)
    down_diff = abs(meta_pop - down)
    if up_diff >= down_diff and up_diff != 0:
        return down
    if down_diff > up_diff and down_diff != 0:
        return up
    return int(meta_pop)


def weibull_scale_for_mean(mean: float, shape: float) -> float:
    return mean / math.gamma(1.0 + 1.0 / shape)


def _set(p, name: str, value):
    """Set one field; a name that is no field of the structure is an error
    (ctypes would otherwise keep it as a plain attribute)."""
    if name not in {f[0] for f in type(p)._fields_}:
        raise AttributeError(f"{type(p).__name__} has no field {name!r}")
    setattr(p, name, value)


def base_params(**over) -> abi.Params:
    p = abi.Params()
    p.n_serotypes = 4
    p.beta_hm = 0.3
    p.beta_mh = 0.3
    p.long_distance = 0.01
    p.external_infection = 0.0
    p.neighbourhood = 8
    p.include_self = 1
    p.dispersal_scale = 1.0
    p.iip_min, p.iip_max = 4, 7
    p.recovery = 5.0
    p.eip = 10.0
    p.h_infant_shape = 0.4
    p.h_infant_scale = 0.0          # infant term off
    p.h_loc = 0.0
    p.h_scale = weibull_scale_for_mean(70.0 * 365.0, 6.0)
    p.h_shape = 6.0
    p.m_scale = 10.0                # exponential, mean 10 d
    p.m_shape = 1.0
    p.m_ratio = 2.0
    p.season_amp = 0.5
    p.season_phase = 0.0
    p.seed_frac_h = 0.001
    p.seed_frac_m = 0.001
    p.immunity_mode = 0
    p.sero_prev = 0.0
    p.no_distance = 0.0
    p.count_law = abi.COUNT_POISSON
    for k, v in over.items():
        _set(p, k, v)
    return p


def reference_params(**over) -> abi.Params:
    """The reference's own defaults (SimulationParam.cpp:11-47): bi-Weibull
    humans, Weibull(4, 23 d) mosquitoes, von Neumann lattice with unit weights
    and no self entry, one serotype, 75% of events staying at the source
    (noDistance) and the stochastic-rounding event counts of
    nepidemic.cu:64-104."""
    p = base_params(
        n_serotypes=1, beta_hm=0.25 * 0.5, beta_mh=0.25 * 0.5, long_distance=0.01,
        external_infection=5.0, neighbourhood=4, include_self=0, dispersal_scale=0.0,
        iip_min=7, iip_max=7, recovery=4.0, eip=6.0,
        h_infant_shape=0.4, h_infant_scale=1.0 / 100000.0 / 365.0, h_loc=8.0 * 365.0,
        h_scale=75.0 * 365.0, h_shape=6.0, m_scale=23.0, m_shape=4.0,
        m_ratio=1.5, season_amp=0.0, seed_frac_h=0.0, seed_frac_m=0.0,
        no_distance=0.75, count_law=abi.COUNT_STOCH_ROUND)
    for k, v in over.items():
        _set(p, k, v)
    return p


def seasonal_climate(days: int, t_mean: float = 26.0, t_amp: float = 4.0, h_mean: float = 0.7,
                     h_amp: float = 0.1, p_mean: float = 0.4, p_amp: float = 0.4, phase: float = 0.0,
                     **coeffs) -> ClimateSpec:
    """A smooth synthetic climate, one row per day (the reference reads a
    smoothed daily file, climate_smooth.csv, that it does not ship): annual
    sinusoids of temperature (deg C), relative humidity and precipitation,
    humidity and rain in phase with the temperature.  The coefficients
    default to the means of the reference's MCMC priors (MCMC.h:37-46) and
    SimulationParam defaults (bitingRate 0.25, mosNetGrowth 0.5)."""
    d = np.arange(days, dtype=np.float64)
    s = np.sin(2.0 * np.pi * d / 365.0 + phase)
    return ClimateSpec((t_mean + t_amp * s).astype(np.float32), (h_mean + h_amp * s).astype(np.float32),
                       (p_mean + p_amp * s).astype(np.float32), **coeffs)


def lattice10(seed: int = 1) -> tuple[list[SimSpec], int]:
    """One simulation on a 10x10 lattice, 1,000 humans per community, 365
    days."""
    spec = SimSpec(10, 10, np.full(100, 1000, np.uint32), base_params(), seed)
    return [spec], 365


def lattice100(seed: int = 1, grid: int = 100) -> tuple[list[SimSpec], int]:
    """One simulation on a 100x100 lattice, 1,000 humans per community
    (1e7), Beta(2,5) initial immunity, 3,650 days."""
    spec = SimSpec(grid, grid, np.full(grid * grid, 1000, np.uint32),
                   base_params(immunity_mode=1), seed)
    return [spec], 3650


def batch_sim(i: int, seed: int = 2) -> SimSpec:
    """Simulation i of the 256-simulation batch."""
    r = philox4x32(i, 1, np.arange(4), 0, seed, 0)
    u = _u53(r[0], r[1])
    c_req = round(math.exp(math.log(400.0) + (math.log(1e4) - math.log(400.0)) * float(u[0])))
    C = square_lattice(float(c_req))
    g = int(round(C ** 0.5))
    r2 = philox4x32(i, 2, np.arange(C), 0, seed, 0)
    humans = (200 + ((r2[0].astype(np.uint64) * np.uint64(1801)) >> np.uint64(32))).astype(np.uint32)
    r3 = philox4x32(i, 3, np.arange(5), 0, seed, 0)
    v = _u53(r3[0], r3[1])
    p = base_params(
        beta_hm=float(0.05 + 0.55 * v[0]),
        beta_mh=float(0.05 + 0.55 * v[1]),
        long_distance=float(math.exp(math.log(1e-3) + (math.log(1e-1) - math.log(1e-3)) * v[2])),
        season_amp=float(0.9 * v[3]),
        season_phase=float(2.0 * math.pi * v[4]),
    )
    return SimSpec(g, g, humans, p, (i << 32) | (seed & 0xFFFFFFFF))


def batch256(n_sims: int = 256, seed: int = 2) -> tuple[list[SimSpec], int]:
    """The batch of 256 simulations (parameters and sizes drawn per
    simulation), 1,825 days."""
    return [batch_sim(i, seed) for i in range(n_sims)], 1825


def mcmc_shape(total_humans: int = 1_000_000, communities: int = 1000):
    """Model shape of the MCMC fit (1e6 humans over 1,000 requested
    communities): communities squared like squareLattice, humans spread as
    evenly as possible."""
    C = square_lattice(float(communities))
    g = int(round(C ** 0.5))
    base = total_humans // C
    extra = total_humans - base * C
    humans = np.full(C, base, np.uint32)
    humans[:extra] += 1
    return g, humans


def mcmc_spec(theta: dict | None = None, seed: int = 3, **kw) -> SimSpec:
    g, humans = mcmc_shape(**kw)
    p = base_params()
    for k, v in (theta or {}).items():
        _set(p, k, v)
    return SimSpec(g, g, humans, p, seed)


# ---------------------------------------------------------------------------
# The batched MCMC fit: 7 parameters (beta_hm, beta_mh, log10 long-distance
# fraction, seasonal amplitude, seasonal phase, reporting rate, NB
# dispersion) of a model of 1e6 humans in 1,000 requested communities (32 x
# 32 after squaring), 260 weekly counts (7 * 260 - 1 = 1,819 days,
# main.cpp:34), negative-binomial observations with dispersion d = 0.1, i.e.
# size 1 / d = 10 (variance mu + mu^2 / 10; MCMC.cpp:240-243), held fixed, so
# six parameters are fitted; synthetic data from the model at MCMC_TRUTH.

MCMC_WEEKS = 260
MCMC_DISPERSION = 0.1
MCMC_TRUTH = np.array([0.3, 0.3, -2.0, 0.5, 1.0, 0.1, MCMC_DISPERSION])
PRIOR_FIXED = 4


def mcmc_priors():
    """(kind, a, b) per parameter: the batch's parameter ranges
    (beta ~ U[0.05, 0.6], long-distance ~ LogU[1e-3, 1e-1], amplitude ~
    U[0, 0.9], phase ~ U[0, 2 pi) on the circle), the reference's
    observationRate prior Beta(10, 90) for the reporting rate (MCMC.h:47-48)
    and the dispersion fixed at MCMC_DISPERSION."""
    return [(0, 0.05, 0.6), (0, 0.05, 0.6), (0, -3.0, -1.0), (0, 0.0, 0.9), (1, 0.0, 2.0 * math.pi),
            (2, 10.0, 90.0), (PRIOR_FIXED, MCMC_DISPERSION, 0.0)]


def climate_priors():
    """The reference's fitted set (MCMC.h:37-50): TruncatedNormal priors of
    scaleEIP (2.5, 0.5), scaleMLE (3, 0.5), scaleCC (1, 0.4), scaleHumidity
    (1, 0.4), minMosToHuman (3, 1), Beta(10, 90) for the observation
    (reporting) rate and Beta(20, 80) for the NB dispersion; kind 3 is the
    normal truncated to (0, 10000]."""
    return [(3, 2.5, 0.5), (3, 3.0, 0.5), (3, 1.0, 0.4), (3, 1.0, 0.4), (3, 3.0, 1.0), (2, 10.0, 90.0),
            (2, 20.0, 80.0)]


def mcmc_data(engine, theta=MCMC_TRUTH, n_weeks: int = MCMC_WEEKS, seed: int = 3, obs_seed: int = 7,
              total_humans: int = 1_000_000, communities: int = 1000):
    """Synthetic weekly counts: one simulation of the model at `theta` (Philox
    seed `seed`) on `engine`, its weekly incidence (all serotypes) observed
    through NB(mean reporting * incidence, size 1 / dispersion) with numpy's
    PCG64 seeded by `obs_seed`.  Returns (observed, simulated weekly
    incidence)."""
    from . import abi as _abi
    from .mcmc import sim_params, weekly_incidence
    lib = _abi.load_hip_library()
    g, humans = mcmc_shape(total_humans, communities)
    p = sim_params(lib, theta, base_params())
    days = 7 * n_weeks - 1
    engine.init([SimSpec(g, g, humans, p, seed)], days)
    engine.step(days)
    wk = weekly_incidence(lib, engine.series(0)["n_incidence"], n_weeks)
    rng = np.random.Generator(np.random.PCG64(obs_seed))
    mu = float(theta[5]) * wk.astype(np.float64)
    nb_size = 1.0 / float(theta[6])
    obs = rng.negative_binomial(nb_size, nb_size / (nb_size + mu)).astype(np.uint32)
    return obs, wk
