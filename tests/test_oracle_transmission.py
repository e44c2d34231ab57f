"""Pins the oracle's transmission and natural-history steps (CPU) against
closed forms.

* Force of infection.  In one isolated community every target is hit
  independently (Poisson splitting of the day's events), so the number of
  distinct targets hit on a day is Binomial(n, 1 - exp(-lambda/n)) given the
  infectious counts after the day's demography, with
  lambda_hm = (M_c / N_c) * beta_hm * I_h   (nepidemic.cu:66-70)
  lambda_mh = beta_mh * I_m                 (mepidemic.cu:65-66).
  Long-distance events of a one-community lattice land back on it, so the
  local / long-distance split must add up to lambda.  A dropped ratio, a
  swapped beta or a lost split share moves the mean by many standard errors.
* Periods.  An exposure during day t turns infectious during day
  t + iip + 1 (the reference's `exposed < age` test after the day's ++age,
  ndemographic.cu:57-103) with iip uniform on {iip_min..iip_max}, and the human
  recovers stochRound(recovery) days after that; a mosquito exposed during day
  t turns infectious stochRound(eip) + 1 days later (mdemographic.cu:88-96).
* t = 0: Bernoulli seeding of humans and alive mosquitoes, uniform seeded
  serotypes, per-serotype prior immunity, and exactly stochRound(E[0]) alive
  mosquitoes per community (demographicintial.cu:183-297,
  epidemicinitial.cu:27-99).
* External introductions: stochRound(iota / C) per community, serotype and
  day onto eligible humans (nepidemic.cu:261-309).
"""
import math

import numpy as np
import pytest
from scipy import stats

from arbo_gpu_fit_amd import configs
from arbo_gpu_fit_amd.engine import Engine, SimSpec


def _one_community(n, seed, **over):
    p = configs.base_params(n_serotypes=1, season_amp=0.0, **over)
    return SimSpec(1, 1, np.array([n], np.uint32), p, seed)


def _z(hits, n, lam):
    """z-score of summed hit counts against Binomial(n_r, 1 - exp(-lam_r/n_r))."""
    hits, n, lam = (np.asarray(a, float) for a in (hits, n, lam))
    p = 1.0 - np.exp(-lam / n)
    return (hits.sum() - (n * p).sum()) / math.sqrt((n * p * (1 - p)).sum())


def test_force_of_infection_mosquito_to_human(oracle_lib):
    N, R = 4000, 40
    hits, lam = [], []
    for r in range(R):
        spec = _one_community(N, 1000 + r, beta_hm=0.0, beta_mh=0.2, long_distance=0.3, seed_frac_h=0.0,
                              seed_frac_m=1.0, iip_min=20, iip_max=20)
        eng = Engine(oracle_lib)
        eng.init([spec], 1)
        eng.step(1)
        st = eng.state(0)
        i_m = int(((st["m_word"] & 3) == 2).sum())          # after day 0's demography
        lam.append(float(np.float32(0.2) * np.float32(i_m)))
        hits.append(int(((st["h_word"] & 3) == 1).sum()))
        eng.close()
    z = _z(hits, [N] * R, lam)
    assert abs(z) < 5, z
    # the alternative with the mosquito/human ratio in this direction is far off
    assert abs(_z(hits, [N] * R, 2 * np.array(lam))) > 20


def test_force_of_infection_human_to_mosquito(oracle_lib):
    N, R = 2000, 40
    hits, lam, alive = [], [], []
    for r in range(R):
        spec = _one_community(N, 2000 + r, beta_hm=0.5, beta_mh=0.0, long_distance=0.3, seed_frac_h=0.05,
                              seed_frac_m=0.0, recovery=30.0)
        eng = Engine(oracle_lib)
        eng.init([spec], 1)
        eng.step(1)
        st = eng.state(0)
        m_c = int(eng.series(0)["m_count"][1])               # alive after day 0's emergence
        i_h = int(((st["h_word"] & 3) == 2).sum())
        lam.append(float(np.float32(m_c) / np.float32(N) * np.float32(0.5) * np.float32(i_h)))
        hits.append(int(((st["m_word"] & 3) == 1).sum()))
        alive.append(m_c)
        eng.close()
    z = _z(hits, alive, lam)
    assert abs(z) < 5, z
    assert abs(_z(hits, alive, np.array(lam) * N / np.array(alive))) > 20   # ratio dropped


def _track(eng, T, word_key, birth_of):
    """Observed durations (in state snapshots) of E and of I per individual."""
    st = eng.state(0)
    prev = st[word_key] & 3
    born = birth_of(st)
    e_day = np.full(prev.size, -1)
    i_day = np.full(prev.size, -1)
    e_dur, i_dur = [], []
    for t in range(T):
        eng.step(1)
        st = eng.state(0)
        cur = st[word_key] & 3
        b = birth_of(st)
        reborn = b != born
        e_day[reborn] = -1
        i_day[reborn] = -1
        born = b
        day = t + 1
        new_e = (cur == 1) & ((prev != 1) | reborn)
        e_day[new_e] = day
        to_i = (prev == 1) & (cur == 2) & ~reborn & (e_day >= 0)
        e_dur += list(day - e_day[to_i])
        i_day[to_i] = day
        e_day[to_i] = -1
        to_s = (prev == 2) & (cur == 0) & ~reborn & (i_day >= 0)
        i_dur += list(day - i_day[to_s])
        i_day[to_s] = -1
        prev = cur
    return np.array(e_dur), np.array(i_dur)


def test_human_incubation_and_recovery_periods(oracle_lib):
    spec = _one_community(3000, 61, beta_hm=0.0, beta_mh=0.05, seed_frac_h=0.0, seed_frac_m=0.3,
                          iip_min=3, iip_max=6, recovery=4.3)
    eng = Engine(oracle_lib)
    eng.init([spec], 40)
    e_dur, i_dur = _track(eng, 40, "h_word", lambda st: st["h_birth"])
    eng.close()
    assert e_dur.size > 300 and i_dur.size > 300
    assert set(np.unique(e_dur)) <= {4, 5, 6, 7}
    counts = np.bincount(e_dur - 4, minlength=4)
    assert stats.chisquare(counts).pvalue > 1e-4
    assert set(np.unique(i_dur)) <= {4, 5}
    k, n = int((i_dur == 5).sum()), i_dur.size
    assert abs(k - 0.3 * n) < 5 * math.sqrt(0.21 * n)


def test_mosquito_extrinsic_incubation(oracle_lib):
    spec = _one_community(3000, 71, beta_hm=0.05, beta_mh=0.0, seed_frac_h=0.1, seed_frac_m=0.0,
                          recovery=60.0, eip=7.6)
    eng = Engine(oracle_lib)
    eng.init([spec], 40)
    e_dur, _ = _track(eng, 40, "m_word", lambda st: st["m_word"] >> 4)
    eng.close()
    assert e_dur.size > 200
    assert set(np.unique(e_dur)) <= {8, 9}
    k, n = int((e_dur == 9).sum()), e_dur.size
    assert abs(k - 0.6 * n) < 5 * math.sqrt(0.24 * n)


def test_initial_census(oracle_lib):
    p = configs.base_params(n_serotypes=3, seed_frac_h=0.02, seed_frac_m=0.05, sero_prev=0.25,
                            season_amp=0.4, season_phase=1.0)
    C, n = 36, 1500
    spec = SimSpec(6, 6, np.full(C, n, np.uint32), p, 81)
    eng = Engine(oracle_lib)
    eng.init([spec], 1)
    st = eng.state(0)
    w = st["h_word"]
    status, strain, imm = w & 3, (w >> 2) & 3, (w >> 4) & 15
    N = C * n
    seeded = status == 2
    assert np.all((status == 0) | seeded)
    assert abs(seeded.sum() - 0.02 * N) < 5 * math.sqrt(0.02 * 0.98 * N)
    assert stats.chisquare(np.bincount(strain[seeded], minlength=3)).pvalue > 1e-4
    assert np.all((imm[seeded] >> strain[seeded]) & 1)
    for v in range(3):
        frac = ((imm[~seeded] >> v) & 1).mean()
        assert abs(frac - 0.25) < 5 * math.sqrt(0.25 * 0.75 / (~seeded).sum())
    assert np.all(imm < 8)
    # mosquitoes: exactly stochRound(E[0]) alive per community, in the lowest slots
    _, _, slots, _ = eng.sizes(0)
    cap = slots // C
    m = st["m_word"].reshape(C, cap)
    alive = (m & 3) != 3
    e0 = 2.0 * n * (1 + 0.4 * math.sin(1.0))
    per = alive.sum(axis=1)
    assert np.all((per == math.floor(e0)) | (per == math.ceil(e0)))
    assert np.all(np.diff(alive.astype(int), axis=1) <= 0)        # alive slots form a prefix
    inf = (m & 3) == 2
    assert abs(inf.sum() - 0.05 * alive.sum()) < 5 * math.sqrt(0.05 * 0.95 * alive.sum())
    assert eng.series(0)["m_count"][0] == alive.sum()
    eng.close()


@pytest.mark.parametrize("iota", [0.6, 7.3])
def test_external_introductions(oracle_lib, iota):
    C, n, T = 16, 500, 30
    p = configs.base_params(n_serotypes=2, beta_hm=0.0, beta_mh=0.0, seed_frac_h=0.0, seed_frac_m=0.0,
                            external_infection=iota, iip_min=50, iip_max=50)
    spec = SimSpec(4, 4, np.full(C, n, np.uint32), p, 91)
    eng = Engine(oracle_lib)
    eng.init([spec], T)
    counts = []
    prev = eng.state(0)["h_word"] & 3
    for _ in range(T):
        eng.step(1)
        cur = eng.state(0)["h_word"] & 3
        counts.append(int(((prev == 0) & (cur == 1)).sum()))
        prev = cur
    eng.close()
    # per day: sum over 2 serotypes x 16 communities of stochRound(iota/16),
    # mean 2 iota, each term within one of iota/16
    counts = np.array(counts)
    lo, hi = 2 * C * math.floor(iota / C), 2 * C * math.ceil(iota / C)
    assert np.all((counts >= lo) & (counts <= hi))
    frac = iota / C - math.floor(iota / C)
    sd = math.sqrt(2 * C * frac * (1 - frac) * T)
    assert abs(counts.sum() - 2 * iota * T) < 5 * sd + 1
