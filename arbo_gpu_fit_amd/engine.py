"""Host-side mirror of the reference's simulation driver over the C ABI.

`Engine` wraps one ArboEngine handle: create -> batch_init -> step -> read.
`simulation()` keeps the shape of the reference's simulation()
(simulation.cpp:18-191): initialise, run the day loop, return the incidence
series that MCMC::storeIncidence (MCMC.cpp:190-213) and writeIncidence
(writeoutput.cpp:13-98) consume.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


class ArboError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"arbo error {code}: {msg}")
        self.code = code


@dataclass
class ClimateSpec:
    """Daily climate and its coefficients (include/arbo_c.h ArboClimate,
    SimulationParam.cpp:92-137): temperature (deg C), relative humidity in
    [0, 1) and precipitation per day, at least max_days + 1 days."""
    temperature: np.ndarray
    humidity: np.ndarray
    precipitation: np.ndarray
    biting_rate: float = 0.25
    scale_eip: float = 2.5
    scale_mle: float = 3.0
    scale_cc: float = 1.0
    scale_humidity: float = 1.0
    min_mos_to_human: float = 3.0
    mos_net_growth: float = 0.5

    def struct(self, keep: list) -> abi.Climate:
        """The ABI struct; the arrays it points at are appended to keep."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (self.temperature, self.humidity, self.precipitation)]
        if not (arrs[0].size == arrs[1].size == arrs[2].size):
            raise ValueError("climate columns differ in length")
        keep.extend(arrs)
        return abi.Climate(*(_ptr(a, C.c_float) for a in arrs), arrs[0].size, self.biting_rate, self.scale_eip,
                           self.scale_mle, self.scale_cc, self.scale_humidity, self.min_mos_to_human,
                           self.mos_net_growth)


@dataclass
class SimSpec:
    """One simulation: lattice, humans per community, parameters and seed,
    optionally the mosquito slots and a climate."""
    grid_w: int
    grid_h: int
    humans: np.ndarray            # uint32 [grid_w*grid_h]
    params: abi.Params
    seed: int
    mosq_slots: np.ndarray | None = None
    climate: ClimateSpec | None = None
    mosq_share: np.ndarray | None = None   # float [communities], sums to 1

    @property
    def communities(self) -> int:
        return self.grid_w * self.grid_h


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def climate_day(climate: ClimateSpec, row: int, lib: C.CDLL | None = None) -> np.ndarray:
    """{beta_hm, beta_mh, eip, mosquito Weibull scale} that climate row
    `row` gives a day (arbo_climate_day); day t of a simulation uses row
    t + 1."""
    lib = lib if lib is not None else abi.load_hip_library()
    keep: list = []
    cl = climate.struct(keep)
    out = np.zeros(4, np.float64)
    rc = lib.arbo_climate_day(C.byref(cl), row, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise ArboError(rc, f"arbo_climate_day(row {row})")
    return out


class Engine:
    """An engine over a library exporting include/arbo_c.h (the HIP engine by
    default)."""

    def __init__(self, lib: C.CDLL | None = None, device: int = 0, flags: int = 0, threads: int = 0):
        self.lib = lib if lib is not None else abi.load_hip_library()
        self._h = abi.EnginePtr()
        desc = abi.EngineDesc(abi.ABI_VERSION, device, flags, threads)
        rc = self.lib.arbo_create(C.byref(desc), C.byref(self._h))
        if rc != 0:
            raise ArboError(rc, "arbo_create failed")
        self.specs: list[SimSpec] = []
        self.max_days = 0

    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.arbo_last_error(self._h)
            raise ArboError(rc, msg.decode() if msg else "")

    def close(self):
        if self._h:
            self.lib.arbo_destroy(self._h)
            self._h = abi.EnginePtr()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init(self, specs: list[SimSpec], max_days: int):
        n = len(specs)
        shapes = (abi.Shape * n)()
        params = (abi.Params * n)()
        seeds = (C.c_uint64 * n)()
        keep = []
        for i, s in enumerate(specs):
            hum = np.ascontiguousarray(s.humans, dtype=np.uint32)
            assert hum.size == s.communities
            keep.append(hum)
            shapes[i].grid_w = s.grid_w
            shapes[i].grid_h = s.grid_h
            shapes[i].humans = _ptr(hum, C.c_uint32)
            if s.mosq_slots is not None:
                ms = np.ascontiguousarray(s.mosq_slots, dtype=np.uint32)
                keep.append(ms)
                shapes[i].mosq_slots = _ptr(ms, C.c_uint32)
            if s.mosq_share is not None:
                sh = np.ascontiguousarray(s.mosq_share, dtype=np.float32)
                assert sh.size == s.communities
                keep.append(sh)
                shapes[i].mosq_share = _ptr(sh, C.c_float)
            if s.climate is not None:
                cl = s.climate.struct(keep)
                keep.append(cl)
                shapes[i].climate = C.pointer(cl)
            params[i] = s.params
            seeds[i] = s.seed & 0xFFFFFFFFFFFFFFFF
        self._check(self.lib.arbo_batch_init(self._h, n, shapes, params, seeds, max_days))
        self.specs = list(specs)
        self.max_days = max_days

    def step(self, n_days: int = 1):
        self._check(self.lib.arbo_batch_step(self._h, n_days))

    def sync(self):
        self._check(self.lib.arbo_sync(self._h))

    def set_day(self, day: int):
        self._check(self.lib.arbo_set_day(self._h, day))

    @property
    def day(self) -> int:
        d = C.c_uint32()
        self._check(self.lib.arbo_days_done(self._h, C.byref(d)))
        return d.value

    def sizes(self, sim: int):
        c, s = C.c_uint32(), C.c_uint32()
        h, m = C.c_uint64(), C.c_uint64()
        self._check(self.lib.arbo_sim_sizes(self._h, sim, C.byref(c), C.byref(h), C.byref(m), C.byref(s)))
        return c.value, h.value, m.value, s.value

    def series(self, sim: int) -> dict:
        _, _, _, S = self.sizes(sim)
        rows = self.day + 1
        out = {
            "n_count": np.zeros(rows, np.uint32),
            "m_count": np.zeros(rows, np.uint32),
            "n_infected": np.zeros((rows, S), np.uint32),
            "n_incidence": np.zeros((rows, S), np.uint32),
            "n_one_subpop": np.zeros((rows, S), np.uint32),
        }
        ser = abi.Series(*(_ptr(out[k], C.c_uint32) for k in
                           ("n_count", "m_count", "n_infected", "n_incidence", "n_one_subpop")))
        self._check(self.lib.arbo_read_series(self._h, sim, C.byref(ser)))
        return out

    def state(self, sim: int) -> dict:
        _, nh, nm, _ = self.sizes(sim)
        out = {
            "h_word": np.zeros(nh, np.uint32),
            "h_next": np.zeros(nh, np.uint16),
            "h_birth": np.zeros(nh, np.int32),
            "m_word": np.zeros(nm, np.uint16),
            "m_next": np.zeros(nm, np.uint16),
        }
        st = abi.State(_ptr(out["h_word"], C.c_uint32), _ptr(out["h_next"], C.c_uint16),
                       _ptr(out["h_birth"], C.c_int32), _ptr(out["m_word"], C.c_uint16),
                       _ptr(out["m_next"], C.c_uint16))
        self._check(self.lib.arbo_read_state(self._h, sim, C.byref(st)))
        return out

    def write_state(self, sim: int, state: dict):
        arrs = {k: np.ascontiguousarray(state[k]) for k in ("h_word", "h_next", "h_birth", "m_word", "m_next")}
        st = abi.State(_ptr(arrs["h_word"], C.c_uint32), _ptr(arrs["h_next"], C.c_uint16),
                       _ptr(arrs["h_birth"], C.c_int32), _ptr(arrs["m_word"], C.c_uint16),
                       _ptr(arrs["m_next"], C.c_uint16))
        self._check(self.lib.arbo_write_state(self._h, sim, C.byref(st)))

    def spatial(self, sim: int) -> np.ndarray:
        """Infectious humans per day, serotype and community,
        [days + 1, serotypes, communities] (engines created with
        abi.F_SPATIAL)."""
        C_, _, _, S = self.sizes(sim)
        out = np.zeros((self.day + 1, S, C_), np.uint32)
        self._check(self.lib.arbo_read_spatial(self._h, sim, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def history(self, sim: int) -> np.ndarray:
        """Infection history [serotypes, humans]: -1 never infected, -2
        immune from the t = 0 census, else the age in days at infection
        (engines created with abi.F_HISTORY)."""
        _, nh, _, S = self.sizes(sim)
        out = np.zeros((S, nh), np.int32)
        self._check(self.lib.arbo_read_history(self._h, sim, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def aoi(self, sim: int) -> np.ndarray:
        """Age-of-infection rings [4, 5000]: ring k holds ages at infection
        of humans with k earlier infections (engines created with
        abi.F_AOI)."""
        out = np.zeros((4, abi.AOI_RECORD), np.uint32)
        self._check(self.lib.arbo_read_aoi(self._h, sim, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def census(self, sim: int):
        """(age, history) of every human after the last simulated day, as
        writeCensus records them: age = (days done - 1) - birth day."""
        hb = self.state(sim)["h_birth"].astype(np.int64)
        return ((self.day - 1) - hb).astype(np.uint32), self.history(sim)

    def kernel_time(self, kind: int):
        ms, n = C.c_double(), C.c_uint64()
        self._check(self.lib.arbo_kernel_time(self._h, kind, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_timing(self):
        self._check(self.lib.arbo_reset_timing(self._h))

    def scan_counts(self) -> tuple[int, int]:
        """16-byte census vectors of humans and of mosquitoes the demography
        kernel read since init or the last reset_timing()."""
        h, m = C.c_uint64(), C.c_uint64()
        self._check(self.lib.arbo_scan_counts(self._h, C.byref(h), C.byref(m)))
        return h.value, m.value

    def tuning_candidates(self) -> list[dict]:
        n = C.c_uint32()
        self._check(self.lib.arbo_tuning_candidates(None, 0, C.byref(n)))
        arr = (abi.Tuning * n.value)()
        self._check(self.lib.arbo_tuning_candidates(arr, n.value, C.byref(n)))
        return [t.as_dict() for t in arr]

    def set_tuning(self, **knobs):
        """Set launch knobs (census_unroll, early_mosq, xcd_order, ld_split);
        unnamed knobs keep their values.  Results never depend on them."""
        t = abi.Tuning()
        self._check(self.lib.arbo_get_tuning(self._h, C.byref(t)))
        for k, v in knobs.items():
            setattr(t, k, v)
        self._check(self.lib.arbo_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = abi.Tuning()
        self._check(self.lib.arbo_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def autotune(self, days_each: int = 2) -> tuple[dict, float]:
        """Time every candidate tuning for days_each simulated days and keep
        the fastest; returns it and its ms per day.  The run advances by
        len(tuning_candidates()) * days_each days."""
        t, ms = abi.Tuning(), C.c_double()
        self._check(self.lib.arbo_autotune(self._h, days_each, C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value


def simulation(spec: SimSpec, max_time: int, lib: C.CDLL | None = None) -> dict:
    """One stochastic run (simulation.cpp:18-191): t=0 census, max_time days,
    and the series rows 0..max_time."""
    with Engine(lib) as eng:
        eng.init([spec], max_time)
        eng.step(max_time)
        return eng.series(0)


def weekly_incidence(n_incidence: np.ndarray, max_week: int) -> np.ndarray:
    """Weekly binning of MCMC::storeIncidence (MCMC.cpp:199-209), summed over
    serotypes (the reference indexes [7w+t] and is correct only for one
    serotype)."""
    daily = n_incidence.sum(axis=1) if n_incidence.ndim == 2 else n_incidence
    max_time = daily.shape[0] - 1
    out = np.zeros(max_week, np.int64)
    for w in range(max_week):
        for t in range(7):
            if 7 * w + t <= max_time:
                out[w] += int(daily[7 * w + t])
    return out
