"""Workload generators (CPU): lattice squaring as SimulationParam::squareLattice
(SimulationParam.cpp:50-88), the configured parameter values and the ranges
of the 256-simulation batch."""
import math

import numpy as np

from arbo_gpu_fit_amd import configs


def test_square_lattice_rule():
    # nearest square; ties go to the smaller square; squares are unchanged
    cases = {1000: 1024, 400: 400, 10000: 10000, 12: 9, 13: 16, 20: 16, 21: 25, 2: 1, 6: 4, 7: 9, 30: 25, 31: 36}
    for meta, want in cases.items():
        assert configs.square_lattice(float(meta)) == want, meta


def test_lattice10_values():
    specs, T = configs.lattice10()
    (s,) = specs
    p = s.params
    assert T == 365 and (s.grid_w, s.grid_h) == (10, 10) and int(s.humans.sum()) == 100_000
    assert p.n_serotypes == 4 and (p.iip_min, p.iip_max) == (4, 7)
    assert math.isclose(p.recovery, 5.0) and math.isclose(p.eip, 10.0)
    assert math.isclose(p.beta_hm, 0.3, rel_tol=1e-6) and math.isclose(p.beta_mh, 0.3, rel_tol=1e-6)
    assert math.isclose(p.long_distance, 0.01, rel_tol=1e-6)
    assert p.neighbourhood == 8 and p.include_self == 1 and p.dispersal_scale == 1.0
    # Weibull(6) with mean 70 years; exponential mosquitoes with mean 10 days
    mean_h = p.h_scale * math.gamma(1 + 1 / p.h_shape)
    assert abs(mean_h - 70 * 365) < 0.01
    assert p.m_shape == 1.0 and p.m_scale == 10.0
    assert p.m_ratio == 2.0 and p.season_amp == 0.5
    assert math.isclose(p.seed_frac_h, 1e-3, rel_tol=1e-6) and math.isclose(p.seed_frac_m, 1e-3, rel_tol=1e-6)


def test_lattice100_values():
    specs, T = configs.lattice100(grid=100)
    (s,) = specs
    assert T == 3650 and s.communities == 10_000 and int(s.humans.sum()) == 10_000_000
    assert s.params.immunity_mode == 1


def test_batch256_ranges():
    specs, T = configs.batch256(64)
    assert T == 1825 and len(specs) == 64
    for s in specs:
        C = s.communities
        assert s.grid_w == s.grid_h and s.grid_w ** 2 == C
        assert 361 <= C <= 10000          # squaring of [400, 10000]
        assert s.humans.min() >= 200 and s.humans.max() <= 2000
        p = s.params
        assert 0.05 <= p.beta_hm <= 0.6 and 0.05 <= p.beta_mh <= 0.6
        assert 1e-3 * (1 - 1e-6) <= p.long_distance <= 1e-1 * (1 + 1e-6)
        assert 0 <= p.season_amp <= 0.9 and 0 <= p.season_phase < 2 * math.pi + 1e-6
    # log-uniform community counts: median near sqrt(400 * 10000) = 2000
    cs = np.array([s.communities for s in configs.batch256(256)[0]])
    assert 1000 < np.median(cs) < 4000
    # distinct seeds, reproducible generation
    assert len({s.seed for s in specs}) == len(specs)
    again = configs.batch_sim(5)
    assert np.array_equal(again.humans, specs[5].humans) and again.seed == specs[5].seed


def test_mcmc_shape():
    g, humans = configs.mcmc_shape()
    assert g == 32 and humans.size == 1024 and int(humans.sum()) == 1_000_000
    assert humans.max() - humans.min() <= 1
