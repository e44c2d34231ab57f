"""ctypes mirror of include/arbo_c.h.

The structures here are byte-for-byte the C structs of the engine's C ABI; the
function prototypes are attached to any shared library that exports that ABI.
"""
from __future__ import annotations

import ctypes as C
import os

ABI_VERSION = 7
MAX_SEROTYPES = 4
MAX_COMMUNITY_HUMANS = 65535
MAX_COMMUNITY_SLOTS = 65535

OK = 0
E_INVALID = -1
E_DEVICE = -2
E_ALLOC = -3
E_STATE = -4
E_CAPACITY = -5

COUNT_POISSON = 0
COUNT_STOCH_ROUND = 1

F_GRAPH = 1
F_TIMING = 2
F_SPATIAL = 4
F_HISTORY = 8
F_AOI = 16          # age-of-infection rings (keeps the history too)
AOI_RECORD = 5000

_u32 = C.c_uint32
_f32 = C.c_float
_u32p = C.POINTER(C.c_uint32)
_u16p = C.POINTER(C.c_uint16)
_i32p = C.POINTER(C.c_int32)


class EngineDesc(C.Structure):
    _fields_ = [("abi_version", _u32), ("device", C.c_int32), ("flags", _u32), ("threads", _u32)]


class Params(C.Structure):
    _fields_ = [
        ("n_serotypes", _u32),
        ("beta_hm", _f32), ("beta_mh", _f32), ("long_distance", _f32), ("external_infection", _f32),
        ("neighbourhood", _u32), ("include_self", _u32), ("dispersal_scale", _f32),
        ("iip_min", _u32), ("iip_max", _u32), ("recovery", _f32), ("eip", _f32),
        ("h_infant_shape", _f32), ("h_infant_scale", _f32), ("h_loc", _f32), ("h_scale", _f32), ("h_shape", _f32),
        ("m_scale", _f32), ("m_shape", _f32),
        ("m_ratio", _f32), ("season_amp", _f32), ("season_phase", _f32),
        ("seed_frac_h", _f32), ("seed_frac_m", _f32),
        ("immunity_mode", _u32), ("sero_prev", _f32),
        ("no_distance", _f32), ("count_law", _u32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Climate(C.Structure):
    _fields_ = [("temperature", C.POINTER(C.c_float)), ("humidity", C.POINTER(C.c_float)),
                ("precipitation", C.POINTER(C.c_float)), ("days", _u32),
                ("biting_rate", _f32), ("scale_eip", _f32), ("scale_mle", _f32), ("scale_cc", _f32),
                ("scale_humidity", _f32), ("min_mos_to_human", _f32), ("mos_net_growth", _f32)]


class Shape(C.Structure):
    _fields_ = [("grid_w", _u32), ("grid_h", _u32), ("humans", _u32p), ("mosq_slots", _u32p),
                ("climate", C.POINTER(Climate)), ("mosq_share", C.POINTER(C.c_float))]


class Series(C.Structure):
    _fields_ = [("n_count", _u32p), ("m_count", _u32p), ("n_infected", _u32p),
                ("n_incidence", _u32p), ("n_one_subpop", _u32p)]


class State(C.Structure):
    _fields_ = [("h_word", _u32p), ("h_next", _u16p), ("h_birth", _i32p),
                ("m_word", _u16p), ("m_next", _u16p)]


class Tuning(C.Structure):
    _fields_ = [("census_unroll", _u32), ("early_mosq", _u32), ("xcd_order", _u32), ("ld_split", _u32),
                ("waves_per_simd", _u32), ("waves_per_comm", _u32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


EnginePtr = C.c_void_p

# name -> (restype, argtypes); the exported symbols of include/arbo_c.h
PROTOTYPES = {
    "arbo_create": (C.c_int, [C.POINTER(EngineDesc), C.POINTER(EnginePtr)]),
    "arbo_batch_init": (C.c_int, [EnginePtr, _u32, C.POINTER(Shape), C.POINTER(Params),
                                  C.POINTER(C.c_uint64), _u32]),
    "arbo_batch_step": (C.c_int, [EnginePtr, _u32]),
    "arbo_sync": (C.c_int, [EnginePtr]),
    "arbo_days_done": (C.c_int, [EnginePtr, _u32p]),
    "arbo_sim_sizes": (C.c_int, [EnginePtr, _u32, _u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _u32p]),
    "arbo_read_series": (C.c_int, [EnginePtr, _u32, C.POINTER(Series)]),
    "arbo_read_state": (C.c_int, [EnginePtr, _u32, C.POINTER(State)]),
    "arbo_write_state": (C.c_int, [EnginePtr, _u32, C.POINTER(State)]),
    "arbo_set_day": (C.c_int, [EnginePtr, _u32]),
    "arbo_read_spatial": (C.c_int, [EnginePtr, _u32, _u32p]),
    "arbo_read_history": (C.c_int, [EnginePtr, _u32, C.POINTER(C.c_int32)]),
    "arbo_read_aoi": (C.c_int, [EnginePtr, _u32, _u32p]),
    "arbo_climate_day": (C.c_int, [C.POINTER(Climate), _u32, C.POINTER(C.c_double)]),
    "arbo_kernel_time": (C.c_int, [EnginePtr, _u32, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "arbo_reset_timing": (C.c_int, [EnginePtr]),
    "arbo_scan_counts": (C.c_int, [EnginePtr, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "arbo_tuning_candidates": (C.c_int, [C.POINTER(Tuning), _u32, _u32p]),
    "arbo_set_tuning": (C.c_int, [EnginePtr, C.POINTER(Tuning)]),
    "arbo_get_tuning": (C.c_int, [EnginePtr, C.POINTER(Tuning)]),
    "arbo_autotune": (C.c_int, [EnginePtr, _u32, C.POINTER(Tuning), C.POINTER(C.c_double)]),
    "arbo_destroy": (None, [EnginePtr]),
    "arbo_last_error": (C.c_char_p, [EnginePtr]),
}


def bind(lib: C.CDLL) -> C.CDLL:
    """Attach the ABI prototypes to a loaded library (raises if one is missing)."""
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(PKG_DIR, "_build", "libarbo_hip.so")

_hip_lib = None


def load_hip_library(path: str | None = None) -> C.CDLL:
    """Load the HIP engine (libarbo_hip.so).  There is no fallback: if the
    library has not been built this raises."""
    global _hip_lib
    if path is None and _hip_lib is not None:
        return _hip_lib
    p = path or HIP_LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP engine library not built: {p} (run __graft_entry__.build())")
    lib = bind(C.CDLL(p))
    if path is None:
        _hip_lib = lib
    return lib
