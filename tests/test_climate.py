"""Climate-driven days (CPU): the day parameters of
SimulationParam::adjustClimate (SimulationParam.cpp:92-137), the logistic
emergence target (simulation.cpp:84-132) and mosquito lifetimes under a
per-day lifespan scale (mdemographic.cu:72), pinned on the oracle by what
the model fixes: zero transmission outside the Briere ranges on exactly the
days that read such rows, the target recursion recounted from the census,
and the first-passage law of each birth day's lifetimes."""
import math

import numpy as np
import pytest
from scipy import stats

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import ArboError, ClimateSpec, Engine, SimSpec, climate_day


def briere(T, c, lo, hi):
    return np.where((T < lo) | (T > hi), 0.0, c * T * (T - lo) * np.sqrt(np.maximum(hi - T, 0.0)))


def numpy_day(cl: ClimateSpec, row: int):
    """SimulationParam.cpp:92-137 in numpy double."""
    T = float(cl.temperature[row])
    H = float(cl.humidity[row])
    TK = T + 273.15
    rate = 24.0 * 0.003359 * (TK / 298.0) * math.exp((15000.0 / 1.986) * ((1.0 / 298.0) - (1.0 / TK)))
    mu = 0.8692 - 0.159 * T + 0.01116 * T ** 2 - 0.0003408 * T ** 3 + 0.000003809 * T ** 4
    f = {k: float(np.float32(getattr(cl, k))) for k in ("biting_rate", "scale_eip", "scale_mle", "scale_humidity")}
    return np.array([f["biting_rate"] * float(briere(T, 0.000491, 12.22, 37.46)),
                     f["biting_rate"] * float(briere(T, 0.000849, 17.05, 35.83)),
                     1.0 / (rate * f["scale_eip"]),
                     1.0 / (mu * f["scale_mle"] * (1.0 - H) ** f["scale_humidity"])])


def test_climate_day_parameters(hip_lib, oracle_lib):
    """Both libraries give the same bits, numpy's double formulas agree to
    1e-12, and the Briere curves vanish outside [12.22, 37.46] and
    [17.05, 35.83] deg C."""
    T = np.array([-5.0, 10.0, 12.22, 12.5, 17.0, 17.05, 17.1, 25.0, 30.0, 35.83, 36.0, 37.46, 38.0, 45.0], np.float32)
    cl = ClimateSpec(T, np.linspace(0.0, 0.95, T.size).astype(np.float32), np.zeros(T.size, np.float32),
                     scale_humidity=1.3)
    for r in range(T.size):
        a = climate_day(cl, r, hip_lib)
        b = climate_day(cl, r, oracle_lib)
        assert np.array_equal(a, b), r
        np.testing.assert_allclose(a, numpy_day(cl, r), rtol=1e-12, atol=0)
        t = float(T[r])
        assert (a[0] == 0.0) == (t < 12.22 or t > 37.46 or t == 12.22 or t == 37.46)
        assert (a[1] == 0.0) == (t < 17.05 or t > 35.83 or t == 17.05 or t == 35.83)
    # the thermal EIP of 25 deg C: 1/(0.080616 * 298.15/298 * e^(7552.87 (1/298 - 1/298.15))) = 12.241 d at scale 1
    one = ClimateSpec(np.array([25.0], np.float32), np.array([0.5], np.float32), np.array([0.0], np.float32),
                      scale_eip=1.0)
    assert abs(climate_day(one, 0, oracle_lib)[2] - 12.2412) < 1e-3
    with pytest.raises(ArboError):
        climate_day(one, 1, oracle_lib)


def test_cold_rows_stop_transmission_until_the_day_that_reads_a_warm_row(oracle_lib):
    """Rows 0..10 at 11 deg C (below both Briere ranges), then 28 deg C: day t
    reads row t + 1 (simulation.cpp:122), so no infection happens before day
    10, and infectious mosquitoes seeded at t = 0 infect humans on day 10."""
    days = 40
    temp = np.where(np.arange(days + 1) <= 10, 11.0, 28.0).astype(np.float32)
    cl = ClimateSpec(temp, np.full(days + 1, 0.6, np.float32), np.full(days + 1, 0.3, np.float32),
                     biting_rate=2.0)
    p = configs.base_params(seed_frac_h=0.0, seed_frac_m=0.2)
    spec = SimSpec(3, 3, np.full(9, 400, np.uint32), p, 17, climate=cl)
    eng = Engine(oracle_lib, flags=abi.F_HISTORY, threads=2)
    eng.init([spec], days)
    eng.step(days)
    hist = eng.history(0)
    hb = eng.state(0)["h_birth"].astype(np.int64)
    day = (hb[None, :] + hist)[hist >= 0]
    assert day.size > 0 and day.min() == 10
    eng.close()


def test_emergence_target_follows_the_logistic_recursion(oracle_lib):
    """With lifetimes of about a day every community refills to its target
    each day, so the alive mosquitoes after day t's births are
    sum_c stochRound(M(t+1) N_c / N): within one per community of the
    recursion M(t+1) = M(t) + g M(t) (1 - M(t)/K(t+1)), K = minMosToHuman N
    (1 + P)^scaleCC, restated here in float32 steps."""
    days = 120
    d = np.arange(days + 1)
    prec = (0.8 + 0.7 * np.sin(2 * np.pi * d / 60.0)).astype(np.float32)
    cl = ClimateSpec(np.full(days + 1, 27.0, np.float32), np.full(days + 1, 0.6, np.float32), prec,
                     scale_mle=40.0, scale_cc=1.4, min_mos_to_human=2.0, mos_net_growth=0.3)
    p = configs.base_params(seed_frac_h=0.0, seed_frac_m=0.0)
    humans = np.array([300, 500, 800, 1000, 900, 700, 400, 900, 600], np.uint32)
    spec = SimSpec(3, 3, humans, p, 23, climate=cl)
    eng = Engine(oracle_lib, threads=2)
    eng.init([spec], days)
    eng.step(days)
    m_count = eng.series(0)["m_count"].astype(np.float64)
    N = float(humans.sum())
    K = cl.min_mos_to_human * N * (1.0 + prec.astype(np.float64)) ** cl.scale_cc
    M = np.zeros(days + 1, np.float32)
    M[0] = np.float32(K[0])
    for t in range(days):
        m = float(M[t])
        M[t + 1] = np.float32(m + cl.mos_net_growth * m * (1.0 - m / K[t + 1]))
    assert np.abs(m_count - M).max() <= humans.size + 1, np.abs(m_count - M).max()
    assert M.max() / M.min() > 1.5   # the target moves
    eng.close()


def test_mosquito_lifetimes_follow_the_first_passage_law(oracle_lib):
    """A mosquito born on day b dies on the first day b + a with
    PLE <= F(a; scale(b + a)) (mdemographic.cu:72 under each day's scale),
    so P(lifetime <= a) = max_{a' <= a} F(a'; scale(b + a')).  Humidity
    alternating between 0.3 and 0.85 changes the scale 4.7-fold from day to
    day; the randomised probability-integral transform of the lifetimes of
    births on days 0..39 is uniform (KS)."""
    days = 120
    d = np.arange(days + 1)
    hum = np.where(d % 2 == 0, 0.3, 0.85).astype(np.float32)
    cl = ClimateSpec(np.full(days + 1, 26.0, np.float32), hum, np.full(days + 1, 0.2, np.float32), scale_mle=1.0)
    p = configs.base_params(seed_frac_h=0.0, seed_frac_m=0.0, m_shape=2.0)
    spec = SimSpec(4, 4, np.full(16, 300, np.uint32), p, 29, climate=cl)
    scale = np.array([climate_day(cl, t + 1, oracle_lib)[3] for t in range(days)])

    def F(a, s):
        return 1.0 - np.exp(-(a / s) ** p.m_shape)

    eng = Engine(oracle_lib, threads=2)
    eng.init([spec], days)
    prev = eng.state(0)["m_word"]
    rng = np.random.default_rng(5)
    pit = []
    for b in range(40):
        eng.step(1)
        cur = eng.state(0)["m_word"]
        prev_dead = ((prev & 3) == 3) | ((((b - (prev >> 4).astype(np.int64)) % 4096) < 2048) & ((prev & 3) != 3))
        born = prev_dead & ((cur & 3) != 3)
        life = ((cur[born] >> 4).astype(np.int64) - b) % 4096
        prev = cur
        if life.size == 0:   # nobody died yesterday
            continue
        amax = int(life.max()) + 1
        a = np.arange(amax + 1)
        G = np.maximum.accumulate(F(a, scale[np.minimum(b + a, days - 1)]) * (a > 0))
        lo, hi = G[life - 1], G[life]
        pit.append(lo + (hi - lo) * rng.random(life.size))
    pit = np.concatenate(pit)
    assert pit.size > 20000
    assert stats.kstest(pit, "uniform").pvalue > 1e-3
    eng.close()


def test_climate_validation(oracle_lib):
    """Short climates, humidity >= 1, precipitation <= -1 and bad
    coefficients are errors before anything runs."""
    p = configs.base_params()
    good = configs.seasonal_climate(31)
    bad = [configs.seasonal_climate(30),
           ClimateSpec(good.temperature, np.full(31, 1.0, np.float32), good.precipitation),
           ClimateSpec(good.temperature, good.humidity, np.full(31, -1.0, np.float32)),
           configs.seasonal_climate(31, scale_eip=0.0),
           configs.seasonal_climate(31, min_mos_to_human=0.0)]
    for lib in (oracle_lib,):
        eng = Engine(lib, threads=1)
        eng.init([SimSpec(2, 2, np.full(4, 100, np.uint32), p, 1, climate=good)], 30)
        for cl in bad:
            with pytest.raises(ArboError):
                eng.init([SimSpec(2, 2, np.full(4, 100, np.uint32), p, 1, climate=cl)], 30)
        eng.close()
