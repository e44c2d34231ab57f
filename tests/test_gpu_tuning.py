"""Launch tunings of the day kernels change no result (GPU).

Every candidate of arbo_tuning_candidates (census loads in flight, early
mosquito loads, XCD-contiguous community order, 8 waves per SIMD, several
waves per community) and the extremes of the
long-distance split give the oracle's census and series bit for bit, and an
autotuned run, graph replay included, equals an untuned one."""
import numpy as np
import pytest

from arbo_gpu_fit_amd import abi
from arbo_gpu_fit_amd.engine import Engine
from tests.cases import CASES
from tests.test_gpu_parity import _compare_series, _compare_state

pytestmark = pytest.mark.gpu


def _reference(oracle_lib, specs, T):
    orc = Engine(oracle_lib, threads=8)
    orc.init(specs, T)
    orc.step(T)
    out = [(orc.state(i), orc.series(i)) for i in range(len(specs))]
    orc.close()
    return out


@pytest.mark.parametrize("case", ["ragged", "batch_pair", "climate", "burst"])
def test_every_tuning_is_bit_identical(case, hip_lib, oracle_lib):
    specs, T = CASES[case]()
    ref = _reference(oracle_lib, specs, T)
    gpu = Engine(hip_lib)
    cands = gpu.tuning_candidates()
    cands += [dict(cands[0], ld_split=1), dict(cands[1], ld_split=64)]
    for cand in cands:
        print(f"[tuning] case={case} tuning={cand}", flush=True)
        gpu.init(specs, T)
        gpu.set_tuning(**cand)
        assert gpu.tuning() == cand
        gpu.step(T)
        for i, (st, se) in enumerate(ref):
            _compare_state(gpu.state(i), st, f"{cand} sim {i}")
            _compare_series(gpu.series(i), se, f"{cand} sim {i}")
    gpu.close()


@pytest.mark.parametrize("flags", [0, abi.F_GRAPH])
def test_autotuned_run_equals_untuned(flags, hip_lib, oracle_lib):
    specs, T = CASES["lattice10"]()
    ref = _reference(oracle_lib, specs, T)
    gpu = Engine(hip_lib, flags=flags)
    gpu.init(specs, T)
    gpu.step(10)
    n = len(gpu.tuning_candidates())
    chosen, ms = gpu.autotune(days_each=2)
    print(f"[tuning] autotune flags={flags}: {chosen} {ms:.3f} ms/day", flush=True)
    assert gpu.day == 10 + 2 * n
    assert chosen == gpu.tuning() and ms > 0
    gpu.step(T - gpu.day)
    _compare_state(gpu.state(0), ref[0][0], "autotuned")
    _compare_series(gpu.series(0), ref[0][1], "autotuned")
    gpu.close()


def test_unsupported_tuning_is_rejected(hip_lib):
    gpu = Engine(hip_lib)
    with pytest.raises(RuntimeError):
        gpu.set_tuning(census_unroll=3)
    with pytest.raises(RuntimeError):
        gpu.set_tuning(ld_split=0)
    gpu.close()


def test_every_tuning_keeps_the_outputs(hip_lib, oracle_lib):
    """The instantiations that keep the infection history, the per-community
    series and the age-of-infection rings, under every launch tuning."""
    specs, T = CASES["reference_params"]()
    flags = abi.F_SPATIAL | abi.F_AOI
    orc = Engine(oracle_lib, flags=flags, threads=4)
    orc.init(specs, T)
    orc.step(T)
    ref = (orc.spatial(0), orc.history(0), orc.aoi(0), orc.state(0))
    orc.close()
    gpu = Engine(hip_lib, flags=flags)
    for cand in gpu.tuning_candidates():
        print(f"[tuning] outputs kept, tuning={cand}", flush=True)
        gpu.init(specs, T)
        gpu.set_tuning(**cand)
        gpu.step(T)
        assert np.array_equal(gpu.spatial(0), ref[0]), cand
        assert np.array_equal(gpu.history(0), ref[1]), cand
        assert np.array_equal(gpu.aoi(0), ref[2]), cand
        _compare_state(gpu.state(0), ref[3], f"{cand}")
    gpu.close()


@pytest.mark.parametrize("flags", [0, abi.F_SPATIAL | abi.F_AOI, abi.F_GRAPH])
def test_large_communities_under_every_tuning(hip_lib, oracle_lib, flags):
    """Communities above 8,192 slots (up to 65,535 humans and slots) take the
    eight-wave demography kernel and the eight-pass event kernel whatever
    the tuning asks for; census, series and kept outputs equal the
    oracle's."""
    specs, T = CASES["large"]()
    orc = Engine(oracle_lib, flags=flags & ~abi.F_GRAPH, threads=8)
    orc.init(specs, T)
    orc.step(T)
    ref = (orc.state(0), orc.series(0))
    kept = (orc.spatial(0), orc.history(0), orc.aoi(0)) if flags & abi.F_AOI else None
    orc.close()
    gpu = Engine(hip_lib, flags=flags)
    cands = gpu.tuning_candidates()
    for cand in (cands[0], cands[-1]):   # one wave per community, and four
        print(f"[tuning] large communities, flags={flags}, tuning={cand}", flush=True)
        gpu.init(specs, T)
        gpu.set_tuning(**cand)
        gpu.step(T)
        _compare_state(gpu.state(0), ref[0], f"large {cand}")
        _compare_series(gpu.series(0), ref[1], f"large {cand}")
        if kept:
            assert np.array_equal(gpu.spatial(0), kept[0])
            assert np.array_equal(gpu.history(0), kept[1])
            assert np.array_equal(gpu.aoi(0), kept[2])
    gpu.close()
