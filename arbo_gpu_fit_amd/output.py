"""ctypes mirror of include/arbo_output.h: the reference's nData, nAlive,
nAOI and nSpatial files."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .engine import ArboError

PROTOTYPES = {
    "arbo_write_ndata": (C.c_int, [C.c_char_p, C.POINTER(abi.Series), C.c_uint32, C.c_uint32, C.c_uint32]),
    "arbo_write_nspatial": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32]),
    "arbo_write_nalive": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.c_uint32,
                                    C.c_uint32, C.c_uint32]),
    "arbo_write_naoi": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.c_uint32]),
}


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def write_ndata(path: str, series: dict, sim_run: int, lib: C.CDLL | None = None):
    """Write (sim_run 0) or append one run's series (Engine.series) to an
    nData file (writeoutput.cpp:13-98)."""
    lib = bind(lib if lib is not None else abi.load_hip_library())
    arr = {k: np.ascontiguousarray(series[k], dtype=np.uint32) for k in
           ("n_count", "m_count", "n_infected", "n_incidence", "n_one_subpop")}
    rows, S = arr["n_incidence"].shape
    p = {k: v.ctypes.data_as(C.POINTER(C.c_uint32)) for k, v in arr.items()}
    s = abi.Series(p["n_count"], p["m_count"], p["n_infected"], p["n_incidence"], p["n_one_subpop"])
    rc = lib.arbo_write_ndata(path.encode(), C.byref(s), rows, S, sim_run)
    if rc:
        raise ArboError(rc, f"arbo_write_ndata({path})")


def write_nspatial(path: str, spatial: np.ndarray, sim_run: int, lib: C.CDLL | None = None):
    """Write (sim_run 0) or append one run's per-community series
    (Engine.spatial: [rows, serotypes, communities]) to an nSpatial file
    (writeoutput.cpp:194-241)."""
    lib = bind(lib if lib is not None else abi.load_hip_library())
    a = np.ascontiguousarray(spatial, dtype=np.uint32)
    rows, S, Cn = a.shape
    rc = lib.arbo_write_nspatial(path.encode(), a.ctypes.data_as(C.POINTER(C.c_uint32)), rows, S, Cn, sim_run)
    if rc:
        raise ArboError(rc, f"arbo_write_nspatial({path})")


def write_nalive(path: str, age: np.ndarray, history: np.ndarray, sim_run: int, lib: C.CDLL | None = None):
    """Write (sim_run 0) or append one run's partial census (Engine.census:
    ages [n] and history [serotypes, n]) to an nAlive file
    (writeoutput.cpp:100-148): the first min(5000, n) humans."""
    lib = bind(lib if lib is not None else abi.load_hip_library())
    a = np.ascontiguousarray(age, dtype=np.uint32)
    h = np.ascontiguousarray(history, dtype=np.int32)
    S, n = h.shape
    if a.shape != (n,):
        raise ValueError(f"age has shape {a.shape}, history {h.shape}")
    rc = lib.arbo_write_nalive(path.encode(), a.ctypes.data_as(C.POINTER(C.c_uint32)),
                               h.ctypes.data_as(C.POINTER(C.c_int32)), n, S, sim_run)
    if rc:
        raise ArboError(rc, f"arbo_write_nalive({path})")


def write_naoi(path: str, aoi: np.ndarray, sim_run: int, lib: C.CDLL | None = None):
    """Write (sim_run 0) or append one run's age-of-infection rings
    (Engine.aoi: [4, 5000]) to an nAOI file (writeoutput.cpp:152-191)."""
    lib = bind(lib if lib is not None else abi.load_hip_library())
    a = np.ascontiguousarray(aoi, dtype=np.uint32)
    if a.shape != (4, abi.AOI_RECORD):
        raise ValueError(f"aoi has shape {a.shape}, expected (4, {abi.AOI_RECORD})")
    rc = lib.arbo_write_naoi(path.encode(), a.ctypes.data_as(C.POINTER(C.c_uint32)), sim_run)
    if rc:
        raise ArboError(rc, f"arbo_write_naoi({path})")
