"""GPU engine vs CPU oracle: bit-exact on every census word and series value.

Runs through the C ABI of libarbo_hip.so (GPU) and liboracle (CPU) on the same
seeded inputs."""
import os

import numpy as np
import pytest

from arbo_gpu_fit_amd.engine import Engine
from tests.cases import CASES, CHECK_DAYS
from tests.random_cases import N_RANDOM, random_case

pytestmark = pytest.mark.gpu


def _compare_state(a: dict, b: dict, where: str):
    for k in ("h_word", "h_next", "h_birth", "m_word", "m_next"):
        if not np.array_equal(a[k], b[k]):
            bad = np.nonzero(a[k] != b[k])[0]
            raise AssertionError(f"{where}: {k} differs at {bad.size} positions, first {bad[:5]} "
                                 f"gpu={a[k][bad[:5]]} oracle={b[k][bad[:5]]}")


def _compare_series(a: dict, b: dict, where: str):
    for k in a:
        if not np.array_equal(a[k], b[k]):
            rows = np.nonzero((a[k] != b[k]).reshape(a[k].shape[0], -1).any(axis=1))[0]
            raise AssertionError(f"{where}: series {k} differs from row {rows[0]}")


def _run_pair(specs, T, hip_lib, oracle_lib, flags=0, tuning=None):
    gpu = Engine(hip_lib, flags=flags)
    orc = Engine(oracle_lib, threads=8)
    gpu.init(specs, T)
    if tuning:
        gpu.set_tuning(**tuning)
    orc.init(specs, T)
    days = sorted(set([d for d in CHECK_DAYS if d <= T] + [T]))
    done = 0
    for d in days:
        if d > done:
            gpu.step(d - done)
            orc.step(d - done)
            done = d
        for i in range(len(specs)):
            _compare_state(gpu.state(i), orc.state(i), f"sim {i} day {d}")
    for i in range(len(specs)):
        _compare_series(gpu.series(i), orc.series(i), f"sim {i}")
    return gpu, orc


@pytest.mark.parametrize("name", list(CASES))
def test_parity(name, hip_lib, oracle_lib):
    print(f"[parity] case={name}", flush=True)
    specs, T = CASES[name]()
    gpu, orc = _run_pair(specs, T, hip_lib, oracle_lib)
    s = gpu.series(0)
    assert s["n_count"][0] == int(specs[0].humans.sum())
    gpu.close()
    orc.close()


def test_graph_replay_matches_direct(hip_lib, oracle_lib):
    """hipGraph replay of the day loop gives the oracle's census too."""
    from arbo_gpu_fit_amd.abi import F_GRAPH
    specs, T = CASES["lattice10"]()
    print("[parity] graph replay lattice10", flush=True)
    gpu, orc = _run_pair(specs, T, hip_lib, oracle_lib, flags=F_GRAPH)
    gpu.close()
    orc.close()


def test_batch_invariance(hip_lib):
    """A simulation's results do not depend on the batch it runs in."""
    a, Ta = CASES["reference_params"]()
    b, Tb = CASES["ragged"]()
    c, Tc = CASES["batch_pair"]()
    T = 30
    print("[parity] batch invariance", flush=True)
    batch = Engine(hip_lib)
    batch.init(a + b + c, T)
    batch.step(T)
    for k, spec in enumerate(a + b + c):
        solo = Engine(hip_lib)
        solo.init([spec], T)
        solo.step(T)
        _compare_state(batch.state(k), solo.state(0), f"batch member {k}")
        _compare_series(batch.series(k), solo.series(0), f"batch member {k}")
        solo.close()
    batch.close()


def test_checkpoint_resume(hip_lib):
    """T days == T/2 days + read_state + fresh engine + write_state + T/2 days."""
    specs, T = CASES["high_lambda"]()
    print("[parity] checkpoint resume", flush=True)
    full = Engine(hip_lib)
    full.init(specs, T)
    full.step(T // 2)
    mid = full.state(0)
    full.step(T - T // 2)
    resumed = Engine(hip_lib)
    resumed.init(specs, T)
    resumed.set_day(T // 2)
    resumed.write_state(0, mid)
    resumed.step(T - T // 2)
    _compare_state(resumed.state(0), full.state(0), "resumed")
    a, b = resumed.series(0), full.series(0)
    for k in a:
        assert np.array_equal(a[k][T // 2 + 1:], b[k][T // 2 + 1:]), k
    full.close()
    resumed.close()


def test_reinit_reuses_engine(hip_lib):
    """One engine re-initialised with a smaller, then a larger batch (its device
    buffers reused where large enough, the hipGraph recaptured) gives the
    results of fresh engines."""
    from arbo_gpu_fit_amd.abi import F_GRAPH
    big, _ = CASES["batch_pair"]()
    small, _ = CASES["high_lambda"]()
    T = 24
    print("[parity] re-init", flush=True)
    reused = Engine(hip_lib, flags=F_GRAPH)
    for specs in (big, small, big):
        reused.init(specs, T)
        reused.step(T)
        fresh = Engine(hip_lib)
        fresh.init(specs, T)
        fresh.step(T)
        for k in range(len(specs)):
            _compare_state(reused.state(k), fresh.state(k), f"re-init member {k}")
            _compare_series(reused.series(k), fresh.series(k), f"re-init member {k}")
        fresh.close()
    reused.close()


@pytest.mark.parametrize("name", ["lattice10", "batch_pair"])
def test_spatial_series_matches_oracle(name, hip_lib, oracle_lib):
    """The per-community series (ARBO_F_SPATIAL), bit for bit."""
    from arbo_gpu_fit_amd import abi
    print(f"[parity] spatial case={name}", flush=True)
    specs, T = CASES[name]()
    gpu = Engine(hip_lib, flags=abi.F_SPATIAL)
    orc = Engine(oracle_lib, flags=abi.F_SPATIAL, threads=8)
    for e in (gpu, orc):
        e.init(specs, T)
        e.step(T)
    for i in range(len(specs)):
        a, b = gpu.spatial(i), orc.spatial(i)
        assert a.shape == b.shape and np.array_equal(a, b), f"sim {i}"
        assert np.array_equal(a.sum(axis=2), gpu.series(i)["n_infected"])
    gpu.close()
    orc.close()


@pytest.mark.parametrize("name", ["lattice10", "reference_params", "batch_pair"])
def test_history_matches_oracle(name, hip_lib, oracle_lib):
    """The infection history (ARBO_F_HISTORY), bit for bit, and the census
    ages beside it."""
    from arbo_gpu_fit_amd import abi
    print(f"[parity] history case={name}", flush=True)
    specs, T = CASES[name]()
    gpu = Engine(hip_lib, flags=abi.F_HISTORY)
    orc = Engine(oracle_lib, flags=abi.F_HISTORY, threads=8)
    for e in (gpu, orc):
        e.init(specs, T)
        e.step(T)
    for i in range(len(specs)):
        (ag, hg), (ao, ho) = gpu.census(i), orc.census(i)
        assert np.array_equal(hg, ho), f"sim {i}"
        assert np.array_equal(ag, ao), f"sim {i}"
        assert (hg >= 0).any()
    gpu.close()
    orc.close()


@pytest.mark.parametrize("name", ["lattice10", "reference_params", "batch_pair"])
def test_aoi_rings_match_oracle(name, hip_lib, oracle_lib):
    """The age-of-infection rings (ARBO_F_AOI), bit for bit."""
    from arbo_gpu_fit_amd import abi
    print(f"[parity] aoi case={name}", flush=True)
    specs, T = CASES[name]()
    gpu = Engine(hip_lib, flags=abi.F_AOI)
    orc = Engine(oracle_lib, flags=abi.F_AOI, threads=8)
    for e in (gpu, orc):
        e.init(specs, T)
        e.step(T)
    recorded = 0
    for i in range(len(specs)):
        a, b = gpu.aoi(i), orc.aoi(i)
        assert np.array_equal(a, b), f"sim {i}: {np.flatnonzero(a != b)[:10]}"
        recorded += int((a != 0).sum())
        assert np.array_equal(gpu.history(i), orc.history(i))
    assert recorded > 0
    gpu.close()
    orc.close()


# ARBO_RANDOM_CASES=N widens the sweep to cases 0..N-1 (the default 12 run
# with every suite; a wider sweep is a one-off run).
@pytest.mark.parametrize("k", range(int(os.environ.get("ARBO_RANDOM_CASES", N_RANDOM))))
def test_random_parity(k, hip_lib, oracle_lib):
    """Seeded random combinations of every model switch (tests/random_cases.py),
    case k under launch tuning k (mod the candidates): one, two or four
    waves per community, every census-load shape, either XCD order."""
    specs, T = random_case(k)
    probe = Engine(hip_lib)
    cands = probe.tuning_candidates()
    probe.close()
    tuning = cands[k % len(cands)]
    print(f"[parity] random case {k}: {len(specs)} sims, {T} days, tuning {tuning}", flush=True)
    gpu, orc = _run_pair(specs, T, hip_lib, oracle_lib, tuning=tuning)
    gpu.close()
    orc.close()
