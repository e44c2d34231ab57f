"""ctypes mirror of include/arbo_data.h: the reference's climate and dengue
input files (ClimateData.cpp:53-126, DengueData.cpp:49-112) and
degree-heterogeneous community sizes (network.cpp:116-164)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .engine import ArboError, ClimateSpec

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)

PROTOTYPES = {
    "arbo_read_climate_csv": (C.c_int, [C.c_char_p, _f32p, _f32p, _f32p, C.c_uint32, _u32p]),
    "arbo_read_dengue_csv": (C.c_int, [C.c_char_p, _f32p, C.c_uint32, _u32p]),
    "arbo_community_sizes": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_float,
                                       C.c_uint64, _u32p]),
}


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _lib(lib):
    return bind(lib if lib is not None else abi.load_hip_library())


def read_climate_csv(path: str, lib: C.CDLL | None = None, **coeffs) -> ClimateSpec:
    """The T, H and P columns of a climate file as a ClimateSpec (the
    coefficients as keyword arguments)."""
    lib = _lib(lib)
    n = C.c_uint32()
    rc = lib.arbo_read_climate_csv(path.encode(), None, None, None, 0, C.byref(n))
    if rc:
        raise ArboError(rc, f"arbo_read_climate_csv({path})")
    cols = [np.zeros(n.value, np.float32) for _ in range(3)]
    rc = lib.arbo_read_climate_csv(path.encode(), *(c.ctypes.data_as(_f32p) for c in cols), n.value, C.byref(n))
    if rc:
        raise ArboError(rc, f"arbo_read_climate_csv({path})")
    return ClimateSpec(*cols, **coeffs)


def read_dengue_csv(path: str, lib: C.CDLL | None = None) -> np.ndarray:
    """The weekly incidence column of a dengue file."""
    lib = _lib(lib)
    n = C.c_uint32()
    rc = lib.arbo_read_dengue_csv(path.encode(), None, 0, C.byref(n))
    if rc:
        raise ArboError(rc, f"arbo_read_dengue_csv({path})")
    out = np.zeros(n.value, np.float32)
    rc = lib.arbo_read_dengue_csv(path.encode(), out.ctypes.data_as(_f32p), n.value, C.byref(n))
    if rc:
        raise ArboError(rc, f"arbo_read_dengue_csv({path})")
    return out


def community_sizes(grid_w: int, grid_h: int, total: int, hetero: float, seed: int, neighbourhood: int = 4,
                    include_self: int = 0, lib: C.CDLL | None = None) -> np.ndarray:
    """Individuals per community with probabilities proportional to
    degree^hetero (one per community first)."""
    lib = _lib(lib)
    out = np.zeros(grid_w * grid_h, np.uint32)
    rc = lib.arbo_community_sizes(grid_w, grid_h, neighbourhood, include_self, total, hetero,
                                  seed & 0xFFFFFFFFFFFFFFFF, out.ctypes.data_as(_u32p))
    if rc:
        raise ArboError(rc, "arbo_community_sizes")
    return out


def mosquito_shares(grid_w: int, grid_h: int, hetero: float, neighbourhood: int = 4,
                    include_self: int = 0) -> np.ndarray:
    """Expected mosquito shares degree^hetero / sum (network.cpp:138-159),
    float32, for SimSpec.mosq_share."""
    deg = np.zeros((grid_h, grid_w), np.float64)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if (dx == 0 and dy == 0) or (neighbourhood == 4 and dx != 0 and dy != 0):
                continue
            ys, xs = np.mgrid[0:grid_h, 0:grid_w]
            ok = (xs + dx >= 0) & (xs + dx < grid_w) & (ys + dy >= 0) & (ys + dy < grid_h)
            deg += ok
    deg += 1 if include_self else 0
    deg = np.maximum(deg, 1.0).ravel()
    w = deg ** hetero
    return (w / w.sum()).astype(np.float32)
