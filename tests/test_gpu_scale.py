"""GPU engine at the configured sizes.

* the 100x100 lattice (1e7 humans, 3,650 days configured): bit-exact against the oracle for
  the first 30 days;
* the full 256-simulation batch for all 1,825 days, checked through properties
  that do not need the oracle at that size: constant human counts, mosquito
  counts within capacity, census recounts of the last series row,
  batch-vs-solo identity for sampled simulations and run-to-run determinism;
* error paths of the C ABI on the device;
* 1,000 replicates of a small model on the GPU and on the oracle with disjoint
  seeds: epidemic summaries agree by two-sample KS tests (family-wise
  alpha = 0.01, Bonferroni over the summaries).
"""
import hashlib

import numpy as np
import pytest
from scipy import stats

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import ArboError, Engine, SimSpec
from tests.test_gpu_parity import _compare_series, _compare_state

pytestmark = pytest.mark.gpu


def test_lattice100_first_month_bit_exact(hip_lib, oracle_lib):
    specs, _ = configs.lattice100()
    T = 30
    print("[scale] lattice100 30 days", flush=True)
    gpu = Engine(hip_lib)
    orc = Engine(oracle_lib, threads=16)
    for e in (gpu, orc):
        e.init(specs, T)
    for d in (1, T):
        gpu.step(d - gpu.day)
        orc.step(d - orc.day)
        _compare_state(gpu.state(0), orc.state(0), f"lattice100 day {d}")
    _compare_series(gpu.series(0), orc.series(0), "lattice100")
    gpu.close()
    orc.close()


def _series_digest(eng, n):
    h = hashlib.sha256()
    for i in range(n):
        for k, v in eng.series(i).items():
            h.update(v.tobytes())
    return h.hexdigest()


def test_full_batch_properties(hip_lib):
    specs, T = configs.batch256()
    print(f"[scale] full batch {len(specs)} sims x {T} days", flush=True)
    eng = Engine(hip_lib, flags=abi.F_GRAPH)
    eng.init(specs, T)
    eng.step(T)
    eng.sync()
    digest = _series_digest(eng, len(specs))
    for i, s in enumerate(specs):
        ser = eng.series(i)
        _, nh, nm, S = eng.sizes(i)
        assert np.all(ser["n_count"] == nh)
        assert np.all(ser["m_count"] <= nm)
        assert np.all(ser["n_incidence"][0] == 0)
    for i in (0, 101, 255):
        st = eng.state(i)
        ser = eng.series(i)
        w = st["h_word"]
        isI = (w & 3) == 2
        S = specs[i].params.n_serotypes
        assert np.array_equal(ser["n_infected"][T], np.bincount((w[isI] >> 2) & 3, minlength=4)[:S])
        assert ser["m_count"][T] == int(((st["m_word"] & 3) != 3).sum())
        inf = (w & 3) != 0
        assert np.all((((w >> 4) & 15) >> ((w >> 2) & 3))[inf] & 1)
    sampled = {i: (eng.state(i), eng.series(i)) for i in (3, 200)}
    eng.close()
    print("[scale] solo reruns", flush=True)
    for i, (st, ser) in sampled.items():
        solo = Engine(hip_lib)
        solo.init([specs[i]], T)
        solo.step(T)
        _compare_state(solo.state(0), st, f"solo sim {i}")
        _compare_series(solo.series(0), ser, f"solo sim {i}")
        solo.close()
    print("[scale] determinism rerun", flush=True)
    again = Engine(hip_lib)
    again.init(specs, T)
    again.step(T)
    assert _series_digest(again, len(specs)) == digest
    again.close()


def test_device_error_paths(hip_lib):
    eng = Engine(hip_lib)
    with pytest.raises(ArboError) as ei:
        eng.step(1)
    assert ei.value.code == abi.E_STATE
    bad = configs.base_params(n_serotypes=7)
    with pytest.raises(ArboError) as ei:
        eng.init([SimSpec(2, 2, np.full(4, 10, np.uint32), bad, 1)], 5)
    assert ei.value.code == abi.E_INVALID and "n_serotypes" in str(ei.value)
    with pytest.raises(ArboError):
        eng.init([SimSpec(2, 2, np.full(4, abi.MAX_COMMUNITY_HUMANS + 1, np.uint32), configs.base_params(), 1)], 5)
    eng.init([SimSpec(2, 2, np.full(4, 10, np.uint32), configs.base_params(), 1)], 5)
    eng.step(5)
    with pytest.raises(ArboError) as ei:
        eng.step(1)
    assert ei.value.code == abi.E_CAPACITY
    with pytest.raises(ArboError):
        eng.series(1)
    eng.close()


def _summaries(eng, n):
    out = []
    for i in range(n):
        s = eng.series(i)
        inc = s["n_incidence"].astype(np.int64)
        tot = inc.sum(axis=1)
        out.append([tot.max(), int(np.argmax(tot))] + list(inc.sum(axis=0)) + [inc.sum()])
    return np.array(out, float)


def test_replicates_ks_against_oracle(hip_lib, oracle_lib):
    p = configs.base_params(seed_frac_h=0.01, seed_frac_m=0.01, beta_hm=0.4, beta_mh=0.4)
    n, T = 1000, 150
    base = np.full(16, 250, np.uint32)
    gpu_specs = [SimSpec(4, 4, base, p, 1 + i) for i in range(n)]
    orc_specs = [SimSpec(4, 4, base, p, 1_000_001 + i) for i in range(n)]
    print("[scale] 1000 replicates", flush=True)
    g = Engine(hip_lib)
    g.init(gpu_specs, T)
    g.step(T)
    o = Engine(oracle_lib, threads=16)
    o.init(orc_specs, T)
    o.step(T)
    a, b = _summaries(g, n), _summaries(o, n)
    names = ["peak incidence", "peak day", "final size s0", "final size s1", "final size s2", "final size s3",
             "attack"]
    for j, name in enumerate(names):
        pv = stats.ks_2samp(a[:, j], b[:, j]).pvalue
        print(f"[scale] KS {name}: p={pv:.3f}", flush=True)
        assert pv > 0.01 / len(names), name
    g.close()
    o.close()
