"""GPU engine vs CPU oracle late in long runs, where the exact skips of the
day kernels matter most: the human scan skipped until a community's earliest
event day, the mosquito scan stopped at the high-water mark of alive slots,
and the event kernel leaving communities with no arrivals at once.

The GPU runs from day 0 with its own carried summaries; at day D its census
of a simulation is written into the oracle at that day (arbo_write_state,
arbo_set_day), and both run on.  Every census word and series row after D
must agree bit for bit:

* the 100x100 lattice (1e7 humans, 3,650 days configured) at days 1,000 and
  3,000, in the endemic regime;
* the three simulations of the 256-simulation batch with the most
  communities (10,000 each), inside the full batch, at day 120 (the
  epidemic peak of the batch, where most communities hold exposed and
  infectious humans and mosquitoes every day) and at days 400 and 1,500;
* runs across the wraps of the packed day fields: mosquito death days are
  kept mod 4,096 and the human and mosquito next-event days mod 65,536.
"""
import numpy as np
import pytest

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import Engine, SimSpec
from tests.test_gpu_parity import _compare_series, _compare_state

pytestmark = pytest.mark.gpu


def _series_after(ser: dict, day: int) -> dict:
    return {k: v[day + 1:] for k, v in ser.items()}


def _resume_and_compare(gpu: Engine, specs, sims, day: int, days: int, T: int, oracle_lib, where: str,
                        threads: int = 16):
    """Oracle resumed from the GPU's census of `sims` at `day`; both run
    `days` more days (the GPU with every simulation of its batch)."""
    orc = Engine(oracle_lib, threads=threads)
    orc.init([specs[i] for i in sims], T)
    orc.set_day(day)
    for j, i in enumerate(sims):
        orc.write_state(j, gpu.state(i))
    gpu.step(days)
    orc.step(days)
    for j, i in enumerate(sims):
        _compare_state(gpu.state(i), orc.state(j), f"{where} sim {i} day {day + days}")
        _compare_series(_series_after(gpu.series(i), day), _series_after(orc.series(j), day),
                        f"{where} sim {i} days {day + 1}..{day + days}")
    orc.close()


def test_lattice100_endemic_windows_bit_exact(hip_lib, oracle_lib):
    specs, T = configs.lattice100()
    gpu = Engine(hip_lib, flags=abi.F_GRAPH)
    gpu.init(specs, T)
    done = 0
    for day, days in ((1000, 30), (3000, 30)):
        print(f"[late] lattice100 to day {day}, then {days} days against the oracle", flush=True)
        gpu.step(day - done)
        _resume_and_compare(gpu, specs, [0], day, days, T, oracle_lib, "lattice100")
        done = day + days
    inf = gpu.series(0)["n_infected"]
    assert inf[1000:].sum() > 0   # still endemic: the windows saw transmission
    gpu.close()


def test_batch_largest_simulations_late_windows_bit_exact(hip_lib, oracle_lib):
    specs, T = configs.batch256()
    big = sorted(range(len(specs)), key=lambda i: (-specs[i].communities, i))[:3]
    gpu = Engine(hip_lib, flags=abi.F_GRAPH)
    gpu.init(specs, T)
    done = 0
    for day, days in ((120, 10), (400, 20), (1500, 20)):
        print(f"[late] batch of {len(specs)} to day {day}; simulations {big} "
              f"({[specs[i].communities for i in big]} communities) {days} days against the oracle", flush=True)
        gpu.step(day - done)
        _resume_and_compare(gpu, specs, big, day, days, T, oracle_lib, "batch")
        done = day + days
    for i in big:
        ser = gpu.series(i)
        assert ser["n_incidence"][121:131].sum() > 0   # the day-120 window is inside the epidemic
        assert ser["n_incidence"][:400].sum() > 0
    gpu.close()


def _wrap_spec(seed: int) -> SimSpec:
    """A small model kept endemic by introductions on every serotype."""
    p = configs.base_params(seed_frac_h=0.01, seed_frac_m=0.01, external_infection=4.0, beta_hm=0.35,
                            beta_mh=0.35)
    return SimSpec(6, 6, np.full(36, 300, np.uint32), p, seed)


@pytest.mark.parametrize("day", [4070, 65510])
def test_runs_across_day_field_wraps_bit_exact(hip_lib, oracle_lib, day):
    """The packed census keeps mosquito death days mod 4,096 and the
    next-event days mod 65,536; the GPU's earliest-event summaries compare
    days across those wraps.  40 days from just before each wrap."""
    T = day + 60
    specs = [_wrap_spec(41), _wrap_spec(42)]
    print(f"[late] wrap run to day {day}, then 40 days against the oracle", flush=True)
    gpu = Engine(hip_lib, flags=abi.F_GRAPH)
    gpu.init(specs, T)
    gpu.step(day)
    _resume_and_compare(gpu, specs, [0, 1], day, 40, T, oracle_lib, f"wrap {day}", threads=4)
    ser = gpu.series(0)
    assert ser["n_incidence"][day + 1:day + 41].sum() > 0   # events across the wrap
    gpu.close()


def test_scan_counts_account_for_the_skips(hip_lib, oracle_lib):
    """arbo_scan_counts: on day 0 every community scans every member (its
    summary is unknown), so the GPU's counts equal the oracle's full scan;
    later days read at most that, and fewer once summaries skip work."""
    specs, _ = configs.lattice10()
    T = 120
    gpu = Engine(hip_lib)
    orc = Engine(oracle_lib, threads=4)
    for e in (gpu, orc):
        e.init(specs, T)
        e.step(1)
    assert gpu.scan_counts() == orc.scan_counts()
    full = orc.scan_counts()
    assert full == (int(sum((specs[0].humans + 3) // 4)), int(orc.sizes(0)[2] // 8))
    for e in (gpu, orc):
        e.reset_timing()
        e.step(T - 1)
    h, m = gpu.scan_counts()
    fh, fm = orc.scan_counts()
    assert fh == full[0] * (T - 1) and fm == full[1] * (T - 1)
    assert 0 < h < fh and 0 < m <= fm
    gpu.close()
    orc.close()
