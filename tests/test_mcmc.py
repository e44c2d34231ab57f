"""The batched MCMC driver (CPU): its negative-binomial likelihood against
scipy, its weekly binning against the reference's rule, the adaptive
Metropolis core on targets with known answers (a correlated Gaussian; the
prior itself under a flat likelihood, which pins the circular wrap and the
Beta density), and the full driver over the CPU oracle's engine on a tiny
model (determinism, counters, support)."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

from arbo_gpu_fit_amd import abi, configs, mcmc
from arbo_gpu_fit_amd.diagnostics import ess, rhat
from arbo_gpu_fit_amd.engine import ArboError, Engine, SimSpec


@pytest.fixture(scope="module")
def lib():
    return mcmc.bind(abi.load_hip_library())


def test_nb_loglik_matches_scipy(lib):
    rng = np.random.default_rng(5)
    r = 10.0
    sim = rng.integers(0, 500, 40).astype(np.uint32)
    data = rng.integers(0, 80, 40).astype(np.uint32)
    rho = 0.13
    got = mcmc.nb_loglik(lib, sim, data, 1, rho, r, floor=-1e300)
    mu = rho * sim[1:].astype(float)
    ref = np.where(mu > 0, stats.nbinom.logpmf(data[1:], r, r / (r + np.where(mu > 0, mu, 1.0))),
                   np.where(data[1:] == 0, 0.0, -np.inf))
    ref = np.maximum(ref, -1e300)
    assert np.isclose(got, ref.sum(), rtol=1e-12, atol=1e-9)
    # the reference's parameterisation: dispersion 1/r, p = disp mu / (1 + disp mu) (MCMC.cpp:216-243)
    k, m = 7, 4.5
    disp = 1.0 / r
    p = disp * m / (1 + disp * m)
    ref_gsl = math.lgamma(k + r) - math.lgamma(k + 1) - math.lgamma(r) + k * math.log(p) + r * math.log(1 - p)
    one = mcmc.nb_loglik(lib, np.array([0, 45], np.uint32), np.array([0, k], np.uint32), 1, 0.1, r)
    assert np.isclose(one, ref_gsl, rtol=1e-13)
    # clamp at -250 per week, zero mean with a positive count is clamped, zero with zero is 0 (the
    # reference computes 0 log 0 = NaN there and clamps it: mcmc.F_ZERO_WEEK_FLOOR, tested below)
    z = mcmc.nb_loglik(lib, np.array([9, 0, 0, 100000], np.uint32), np.array([9, 0, 5, 0], np.uint32), 1, 0.5, r)
    assert np.isclose(z, 0.0 - 250.0 + max(-250.0, stats.nbinom.logpmf(0, r, r / (r + 50000.0))))


def test_zero_week_floor_flag_matches_the_reference(lib):
    """With F_ZERO_WEEK_FLOOR a week with zero simulated mean and a zero count
    scores the floor, as negbinomial_pdf's 0 log 0 = NaN under fmax(-250, .)
    does in the reference (MCMC.cpp:216-243); without it the week scores
    log 1 = 0.  Checked through chains over the oracle engine: a model with
    no transmission simulates zero incidence, so every week is zero/zero."""
    from tests.oracle_lib import load_oracle
    olib = load_oracle()
    W = 4
    p = configs.base_params(beta_hm=0.0, beta_mh=0.0, seed_frac_h=0.0, seed_frac_m=0.0)
    lls = {}
    for flags in (0, mcmc.F_ZERO_WEEK_FLOOR):
        cfg = mcmc.McmcConfig(n_chains=2, n_weeks=W, priors=configs.mcmc_priors(), adapt_after=50, seed=3,
                              flags=flags)
        m = mcmc.Mcmc(cfg, 2, 2, np.full(4, 20, np.uint32), p, np.zeros(W, np.uint32), lib=lib)
        eng = Engine(olib, threads=1)
        m.run(eng, 1, api=mcmc.engine_api(olib))
        lls[flags] = m.history()[1][0]
        m.close()
        eng.close()
    assert np.all(lls[0] == 0.0)
    assert np.all(lls[mcmc.F_ZERO_WEEK_FLOOR] == -250.0 * (W - 1))


def test_weekly_binning_reference_rule(lib):
    rng = np.random.default_rng(3)
    W = 9
    rows = 7 * W            # days 0..7W-1
    inc1 = rng.integers(0, 50, (rows, 1)).astype(np.uint32)
    got = mcmc.weekly_incidence(lib, inc1, W)
    max_time = rows - 1
    ref = []
    for w in range(W):                      # MCMC.cpp:199-209 with C_STRAINS = 1
        s = int(inc1[7 * w, 0])
        for t in range(1, 7):
            if w * 7 + t <= max_time:
                s += int(inc1[7 * w + t, 0])
        ref.append(s)
    assert list(got) == ref
    inc4 = rng.integers(0, 50, (rows - 3, 4)).astype(np.uint32)    # short last week
    got4 = mcmc.weekly_incidence(lib, inc4, W)
    ref4 = [int(inc4[7 * w:min(7 * w + 7, rows - 3)].sum()) for w in range(W)]
    assert list(got4) == ref4


def _gauss_cfg(n_chains=64, seed=11, **kw):
    priors = [(0, -20.0, 20.0)] * 6 + [(mcmc.PRIOR_FIXED, 0.1, 0.0)]
    return mcmc.McmcConfig(n_chains=n_chains, n_weeks=3, priors=priors, adapt_after=200, seed=seed, **kw)


def _dummy_model(cfg, lib, theta0=None):
    return mcmc.Mcmc(cfg, 1, 1, [10], configs.base_params(), np.zeros(cfg.n_weeks, np.uint32), theta0=theta0,
                     lib=lib)


def test_gaussian_target_recovery(lib):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(6, 6))
    cov = A @ A.T / 6 + 0.2 * np.eye(6)
    mean = rng.normal(size=6)
    prec = np.linalg.inv(cov)

    def loglik(th):
        d = th[:, :6] - mean
        return -0.5 * np.einsum("ni,ij,nj->n", d, prec, d)

    cfg = _gauss_cfg()
    m = _dummy_model(cfg, lib, theta0=np.tile(np.append(mean + 1.0, 0.1), (cfg.n_chains, 1)))
    m.run_target(loglik, 1500)
    th, _, acc = m.history()
    assert np.all(th[:, :, 6] == 0.1)                    # the fixed parameter never moves
    post = th[500:, :, :6].reshape(-1, 6)
    se = np.sqrt(np.diag(cov) / np.array([ess(th[500:, :, i]) for i in range(6)]))
    assert np.all(np.abs(post.mean(axis=0) - mean) < 5 * se)
    assert np.allclose(np.cov(post.T), cov, atol=0.12 * np.abs(cov).max())
    assert 0.1 < acc[500:].mean() < 0.5                  # adapted random walk in 6 d (~0.23 optimal)
    assert all(rhat(th[500:, :, i]) < 1.05 for i in range(6))     # ~50 effective samples per chain
    c = m.counters()
    assert c["proposals"] == 1500 * 64 and c["accepted"] == int(acc.sum())
    m.close()


def test_prior_is_sampled_under_flat_likelihood(lib):
    priors = configs.mcmc_priors()
    cfg = mcmc.McmcConfig(n_chains=64, n_weeks=3, priors=priors, adapt_after=300, seed=4)
    m = _dummy_model(cfg, lib)
    m.run_target(lambda th: np.zeros(len(th)), 1500)
    th, _, _ = m.history()
    s = th[300::5].reshape(-1, 7)
    assert np.all(s[:, 6] == configs.MCMC_DISPERSION)
    # every state inside the support; the phase wraps into [0, 2 pi)
    assert np.all((s[:, 4] >= 0) & (s[:, 4] < 2 * math.pi))
    assert stats.kstest(s[:, 4], stats.uniform(0, 2 * math.pi).cdf).pvalue > 1e-3
    assert stats.kstest(s[:, 5], stats.beta(10, 90).cdf).pvalue > 1e-3
    assert stats.kstest(s[:, 2], stats.uniform(-3, 2).cdf).pvalue > 1e-3
    for i, (k, a, b) in enumerate(priors):
        if k == 0:
            assert s[:, i].min() >= a and s[:, i].max() <= b
    # log prior of the reference's Beta(10, 90) reporting prior and the uniform box
    lp = m.log_prior(np.array([0.3, 0.3, -2.0, 0.5, 1.0, 0.1, 0.1]))
    ref = (-math.log(0.55) * 2 - math.log(2.0) - math.log(0.9) - math.log(2 * math.pi)
           + stats.beta(10, 90).logpdf(0.1))
    assert np.isclose(lp, ref, rtol=1e-12)
    assert m.log_prior(np.array([0.7, 0.3, -2.0, 0.5, 1.0, 0.1, 0.1])) == -np.inf
    assert m.log_prior(np.array([0.3, 0.3, -2.0, 0.5, 1.0, 0.1, 0.2])) == -np.inf   # off the fixed value
    m.close()


def test_invalid_descriptors(lib):
    cfg = _gauss_cfg(n_chains=0)
    with pytest.raises(ArboError):
        _dummy_model(cfg, lib)
    bad = _gauss_cfg()
    bad.priors = list(bad.priors)
    bad.priors[5] = (2, -1.0, 3.0)            # Beta with a negative shape
    with pytest.raises(ArboError):
        _dummy_model(bad, lib)
    with pytest.raises(ArboError):             # start outside the support
        _dummy_model(_gauss_cfg(), lib, theta0=np.full((64, 7), 30.0))


def _tiny_driver(lib, oracle_lib, seed=9, n_chains=6, n_particles=2):
    W = 6
    priors = configs.mcmc_priors()
    humans = np.full(4, 150, np.uint32)
    orc = Engine(oracle_lib, threads=2)
    p = mcmc.sim_params(lib, configs.MCMC_TRUTH, configs.base_params(seed_frac_h=0.02, seed_frac_m=0.02))
    orc.init([SimSpec(2, 2, humans, p, 5)], 7 * W - 1)
    orc.step(7 * W - 1)
    data = mcmc.weekly_incidence(lib, orc.series(0)["n_incidence"], W)
    cfg = mcmc.McmcConfig(n_chains=n_chains, n_weeks=W, priors=priors, n_particles=n_particles, adapt_after=50,
                          seed=seed)
    m = mcmc.Mcmc(cfg, 2, 2, humans, configs.base_params(seed_frac_h=0.02, seed_frac_m=0.02), data, lib=lib)
    return m, orc


def test_driver_over_oracle_engine(lib, oracle_lib):
    api = mcmc.engine_api(oracle_lib)
    runs = []
    for _ in range(2):
        m, orc = _tiny_driver(lib, oracle_lib)
        m.run(orc, 3, api=api)
        runs.append((m.history(), m.counters()))
        th, ll, acc = runs[-1][0]
        assert th.shape == (3, 6, 7) and np.all(np.isfinite(ll))
        for c in range(6):
            for it in range(3):
                assert np.isfinite(m.log_prior(th[it, c]))
        cnt = runs[-1][1]
        assert cnt["simulations"] == 2 * (6 + cnt["in_support"])      # starting values + proposals, 2 particles
        assert cnt["proposals"] == 18
        m.close()
        orc.close()
    (a, b) = runs
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)


def test_features_match_python(lib):
    from arbo_gpu_fit_amd.surrogate import features
    cfg = mcmc.McmcConfig(n_chains=2, n_weeks=3, priors=configs.mcmc_priors())
    m = _dummy_model(cfg, lib)
    th = np.array([[0.3, 0.1, -2.5, 0.8, 5.0, 0.1, 0.1], [0.05, 0.6, -1.0, 0.0, 0.0, 0.5, 0.1]])
    assert np.allclose(m.features(th), features(th, configs.mcmc_priors()), atol=1e-6)
    m.close()


def test_truncated_normal_priors_sampled_under_flat_likelihood(lib):
    """The reference's fitted set (MCMC.h:37-50): TruncatedNormal priors on
    (0, 10000], Beta(10, 90) for the observation rate and Beta(20, 80) for
    the dispersion; the chains sample them (KS) and the log prior equals
    scipy's truncnorm + beta."""
    priors = configs.climate_priors()
    cfg = mcmc.McmcConfig(n_chains=64, n_weeks=3, priors=priors, adapt_after=300, seed=6,
                          param_set=mcmc.FIT_SEASONAL)
    m = _dummy_model(cfg, lib, theta0=np.tile([2.5, 3.0, 1.0, 1.0, 3.0, 0.1, 0.2], (64, 1)))
    m.run_target(lambda th: np.zeros(len(th)), 4000)
    th, _, _ = m.history()
    s = th[400::25].reshape(-1, 7)   # thinned past the chains' autocorrelation
    for i in range(5):
        _, a, b = priors[i]
        dist = stats.truncnorm((0.0 - a) / b, (mcmc.TRUNC_HI - a) / b, loc=a, scale=b)
        assert s[:, i].min() > 0.0
        assert stats.kstest(s[:, i], dist.cdf).pvalue > 1e-3, i
    assert stats.kstest(s[:, 6], stats.beta(20, 80).cdf).pvalue > 1e-3
    x = np.array([2.2, 3.4, 0.3, 1.9, 0.5, 0.12, 0.25])
    ref = sum(stats.truncnorm((0.0 - a) / b, (mcmc.TRUNC_HI - a) / b, loc=a, scale=b).logpdf(x[i])
              for i, (_, a, b) in enumerate(priors[:5])) + stats.beta(10, 90).logpdf(0.12) \
        + stats.beta(20, 80).logpdf(0.25)
    assert np.isclose(m.log_prior(x), ref, rtol=1e-10)
    assert m.log_prior(np.array([2.2, 3.4, -0.1, 1.9, 0.5, 0.12, 0.25])) == -np.inf
    m.close()


def test_climate_fit_sends_each_chain_its_climate(lib, oracle_lib):
    """ARBO_FIT_CLIMATE (MCMC::copyParams, MCMC.cpp:79-86): every simulation
    of a proposal runs under the base climate with the vector's five
    coefficients; the other parameters stay the base's.  A recording
    batch_init in front of the oracle engine sees exactly the chains'
    vectors, two particles each."""
    W = 6
    days = 7 * W - 1
    cl = configs.seasonal_climate(days + 1)
    base = configs.base_params(seed_frac_h=0.02, seed_frac_m=0.02)
    humans = np.full(4, 150, np.uint32)
    seen = []
    init_t = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.POINTER(abi.Shape), C.POINTER(abi.Params),
                         C.POINTER(C.c_uint64), C.c_uint32)

    def recording_init(e, n, shapes, params, seeds, max_days):
        for k in range(n):
            c = shapes[k].climate.contents
            seen.append((c.scale_eip, c.scale_mle, c.scale_cc, c.scale_humidity, c.min_mos_to_human,
                         c.days, params[k].beta_hm))
        return oracle_lib.arbo_batch_init(e, n, shapes, params, seeds, max_days)

    cb = init_t(recording_init)
    api = mcmc.engine_api(oracle_lib)
    api.batch_init = C.cast(cb, C.c_void_p).value
    orc = Engine(oracle_lib, threads=2)
    cfg = mcmc.McmcConfig(n_chains=4, n_weeks=W, priors=configs.climate_priors(), n_particles=2, adapt_after=50,
                          seed=3, param_set=mcmc.FIT_CLIMATE)
    m = mcmc.Mcmc(cfg, 2, 2, humans, base, np.full(W, 20, np.uint32), lib=lib, climate=cl)
    m.run(orc, 2, api=api)
    th, ll, _ = m.history()
    assert np.all(np.isfinite(ll))
    f32 = np.float32
    starts = seen[:8]   # the starting vectors, two particles each
    cur = th[0]
    for k, rec in enumerate(starts):
        assert rec[5] == cl.temperature.size and rec[6] == f32(base.beta_hm)
    got = {tuple(r[:5]) for r in seen}
    for it in range(2):
        for c in range(4):
            v = tuple(float(f32(x)) for x in th[it, c, :5])
            assert v in got, (it, c, v)
    assert cur.shape == (4, 7)
    m.close()
    orc.close()
    with pytest.raises(ArboError):   # the climate set needs a climate
        mcmc.Mcmc(cfg, 2, 2, humans, base, np.full(W, 20, np.uint32), lib=lib)


def test_chain_and_prior_files(tmp_path, lib):
    """MCMC.csv and MCMC_priors.csv layouts (MCMC.cpp:383-440), one chain
    file per chain: numbers as C++'s default ostream prints them (%g, 6
    significant digits), read back to that precision."""
    rng = np.random.default_rng(2)
    th = rng.normal(size=(5, 3, 7))
    ll = rng.normal(size=(5, 3)) * 100
    path = tmp_path / "MCMC.csv"
    mcmc.write_chain_csv(str(path), th, ll, chain=1, names=mcmc.REFERENCE_NAMES)
    lines = path.read_text().split("\n")
    assert lines[0] == "step,logLikelihood,scaleEIP,scaleMLE,scaleCC,scaleHumidity,minMosToHuman,observationRate," \
                       "dispersion"
    back = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
    assert np.array_equal(back[:, 0], np.arange(5))
    assert np.allclose(back[:, 1], ll[:, 1], rtol=1e-5) and np.allclose(back[:, 2:], th[:, 1, :], rtol=1e-5)
    assert lines[1].split(",")[2] == f"{th[0, 1, 0]:g}"
    pp = tmp_path / "MCMC_priors.csv"
    mcmc.write_priors_csv(str(pp), configs.climate_priors(), mcmc.CLIMATE_PARAM_NAMES)
    rows = pp.read_text().split("\n")
    assert rows[0] == "PARAMETER_NAME,PRIOR_DISTRIBUTION,PARAMETERS"
    assert rows[1] == "scale_eip,TruncatedNormal,(2.5;0.5)" and rows[6] == "reporting,Beta,(10.0;90.0)"
    assert rows[7] == "dispersion,Beta,(20.0;80.0)"


def test_candidate_and_parameter_files(tmp_path, lib, oracle_lib):
    """MCMC_candidates.csv (MCMC::writeCandidates, MCMC.cpp:443-484) and
    parameterData.csv (Parameter::writeParameters, parameter.cpp:8-47) from
    a short run over the oracle engine: every iteration's candidate, its
    log posterior (-inf outside the support) and the accepted flag; an
    accepted candidate is the chain's next state with that log posterior."""
    m, orc = _tiny_driver(lib, oracle_lib, n_chains=3, n_particles=1)
    m.run(orc, 4, api=mcmc.engine_api(oracle_lib))
    th, ll, acc = m.history()
    cth, clp = m.candidates()
    assert cth.shape == th.shape and clp.shape == ll.shape
    for it in range(4):
        for c in range(3):
            inside = np.isfinite(m.log_prior(cth[it, c]))
            assert inside == np.isfinite(clp[it, c])
            if acc[it, c]:
                assert np.array_equal(cth[it, c], th[it, c])
                assert np.isclose(clp[it, c], ll[it, c] + m.log_prior(th[it, c]), rtol=1e-12)
    path = tmp_path / "MCMC_candidates.csv"
    mcmc.write_candidates_csv(str(path), cth, clp, acc, chain=2)
    lines = path.read_text().split("\n")
    assert lines[0] == "step,logLikelihood," + ",".join(mcmc.PARAM_NAMES) + ",accepted"
    assert len(lines) == 5
    for it, ln in enumerate(lines[1:]):
        f = ln.split(",")
        assert int(f[0]) == it and int(f[-1]) == int(acc[it, 2]) and len(f) == 2 + mcmc.D + 1
        assert f[1] == mcmc._num(clp[it, 2])
    rows = [mcmc.parameter_data_row(it, mcmc.sim_params(lib, cth[it, 0], m.base), np.full(4, 150))
            for it in range(4) if np.isfinite(clp[it, 0])]
    pd_path = tmp_path / "parameterData.csv"
    mcmc.write_parameter_data_csv(str(pd_path), rows)
    pl = pd_path.read_text().split("\n")
    assert pl[0].split(",") == list(mcmc.PARAMETER_DATA_COLUMNS) and len(pl[0].split(",")) == 28
    first = pl[1].split(",")
    assert first[1] == "600" and first[3] == "4" and first[4] == "150"
    assert np.isclose(float(first[7]) * float(first[5]), cth[0, 0, 0], rtol=1e-5)   # nmBitingSuccess * bitingRate
    assert np.isclose(float(first[14]), 10 ** cth[0, 0, 2], rtol=1e-5)              # longDistance
    m.close()
    orc.close()


def test_unsimulable_proposals_are_rejected_not_fatal(lib, oracle_lib):
    """A proposal inside the prior support that the engine's checks reject
    (here: mosquito capacity above ARBO_MAX_COMMUNITY_SLOTS for a large
    minMosToHuman) gets log-likelihood -inf and is counted; the run goes on
    and the other chains are simulated as usual."""
    W = 4
    cl = configs.seasonal_climate(7 * W, biting_rate=0.5)
    humans = np.full(4, 1000, np.uint32)
    orc = Engine(oracle_lib, threads=2)
    cfg = mcmc.McmcConfig(n_chains=3, n_weeks=W, priors=configs.climate_priors(), adapt_after=50, seed=5,
                          param_set=mcmc.FIT_CLIMATE)
    # chain 1 starts at a ratio whose capacity needs more than 65,535 slots per community
    th0 = np.array([[2.5, 3.0, 1.0, 1.0, 3.0, 0.1, 0.2], [2.5, 3.0, 1.0, 1.0, 80.0, 0.1, 0.2],
                    [2.5, 3.0, 1.0, 1.0, 2.0, 0.1, 0.2]])
    m = mcmc.Mcmc(cfg, 2, 2, humans, configs.base_params(seed_frac_h=0.02, seed_frac_m=0.02),
                  np.full(W, 20, np.uint32), theta0=th0, lib=lib, climate=cl)
    m.run(orc, 2, api=mcmc.engine_api(oracle_lib))
    th, ll, _ = m.history()
    cnt = m.counters()
    assert cnt["unsimulable"] >= 1
    assert np.isfinite(ll[0, 0]) and np.isfinite(ll[0, 2])
    assert ll[0, 1] == -np.inf or th[0, 1, 4] != 80.0   # stuck at -inf unless it moved to a simulable vector
    m.close()
    orc.close()


def test_fit_workflow_with_the_reference_parameter_set(lib, oracle_lib):
    """fit.model_spec puts a climate-set vector's five coefficients into the
    simulation's climate (MCMC::copyParams, MCMC.cpp:79-86): the training
    simulations equal direct runs under those coefficients, differ between
    vectors, and a plain fit over them summarises under the reference's
    parameter names."""
    from arbo_gpu_fit_amd import fit
    from arbo_gpu_fit_amd.engine import ClimateSpec
    W = 5
    cl = configs.seasonal_climate(7 * W, biting_rate=1.0)
    humans = np.full(4, 250, np.uint32)
    shape = (2, humans)
    th = fit.prior_sample(3, seed=4, priors=configs.climate_priors())
    assert np.all(th[:, :5] > 0)
    orc = Engine(oracle_lib, threads=2)
    weekly = fit.simulate_weekly(orc, th, seed=9, n_weeks=W, param_set=mcmc.FIT_CLIMATE, climate=cl, shape=shape,
                                 lib=lib)
    for i in range(3):
        c = ClimateSpec(cl.temperature, cl.humidity, cl.precipitation, cl.biting_rate, *(float(x) for x in th[i, :5]),
                        cl.mos_net_growth)
        spec = SimSpec(2, 2, humans, configs.base_params(), (i << 32) | 9, climate=c)
        e = Engine(oracle_lib, threads=2)
        e.init([spec], 7 * W - 1)
        e.step(7 * W - 1)
        assert np.array_equal(weekly[i], mcmc.weekly_incidence(lib, e.series(0)["n_incidence"], W)), i
        e.close()
    assert len({tuple(w) for w in weekly}) > 1
    m, out = fit.run_fit(orc, weekly[0], iters=2, chains=4, seed=3, param_set=mcmc.FIT_CLIMATE, climate=cl,
                         shape=shape, api=mcmc.engine_api(oracle_lib), lib=lib, burn=0)
    assert set(out["ess"]) == set(mcmc.CLIMATE_PARAM_NAMES)
    assert out["counters"]["proposals"] == 8
    m.close()
    orc.close()
