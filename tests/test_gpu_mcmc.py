"""The batched MCMC driver over the GPU engine: its chains equal, value for
value, the chains of the same driver over the CPU oracle's engine (the
simulations are bit-exact, so every likelihood and decision is too)."""
import numpy as np
import pytest

from arbo_gpu_fit_amd import abi, mcmc
from arbo_gpu_fit_amd.engine import Engine
from tests.test_mcmc import _tiny_driver

pytestmark = pytest.mark.gpu


def test_chains_equal_over_gpu_and_oracle(hip_lib, oracle_lib):
    lib = mcmc.bind(hip_lib)
    print("[mcmc] gpu vs oracle chains", flush=True)
    m_o, orc = _tiny_driver(lib, oracle_lib, seed=21, n_chains=8, n_particles=2)
    m_o.run(orc, 4, api=mcmc.engine_api(oracle_lib))
    m_g, _ = _tiny_driver(lib, oracle_lib, seed=21, n_chains=8, n_particles=2)
    gpu = Engine(hip_lib, flags=abi.F_GRAPH)
    m_g.run(gpu, 4)
    for x, y in zip(m_o.history(), m_g.history()):
        assert np.array_equal(x, y)
    assert m_o.counters()["simulations"] == m_g.counters()["simulations"]
    for m in (m_o, m_g):
        m.close()
    gpu.close()
    orc.close()


def test_climate_fit_chains_equal_over_gpu_and_oracle(hip_lib, oracle_lib):
    """The reference's fitted set (ARBO_FIT_CLIMATE): every chain's climate
    coefficients reach the GPU engine as they reach the oracle's."""
    from arbo_gpu_fit_amd import configs
    lib = mcmc.bind(hip_lib)
    print("[mcmc] climate fit, gpu vs oracle chains", flush=True)
    W = 6
    cl = configs.seasonal_climate(7 * W)
    base = configs.base_params(seed_frac_h=0.02, seed_frac_m=0.02)
    humans = np.full(9, 200, np.uint32)
    cfg = mcmc.McmcConfig(n_chains=8, n_weeks=W, priors=configs.climate_priors(), n_particles=2, adapt_after=20,
                          seed=5, param_set=mcmc.FIT_CLIMATE)
    data = np.array([0, 3, 8, 15, 20, 30], np.uint32)
    hist = []
    orc = Engine(oracle_lib, threads=4)
    gpu = Engine(hip_lib, flags=abi.F_GRAPH)
    for eng, api in ((orc, mcmc.engine_api(oracle_lib)), (gpu, None)):
        m = mcmc.Mcmc(cfg, 3, 3, humans, base, data, lib=lib, climate=cl)
        m.run(eng, 4, api=api)
        hist.append(m.history())
        m.close()
    for x, y in zip(*hist):
        assert np.array_equal(x, y)
    gpu.close()
    orc.close()
