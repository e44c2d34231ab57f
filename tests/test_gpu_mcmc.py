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
