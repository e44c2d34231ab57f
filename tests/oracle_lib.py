"""Loader of the CPU oracle (oracle/_build/libarbo_oracle.so) for the tests.

The oracle exports the engine's C ABI plus a few inspection hooks of its own
building blocks (Philox, exp/log/sin, Poisson, the host tables)."""
import ctypes as C
import os
import subprocess

from arbo_gpu_fit_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libarbo_oracle.so")

_lib = None


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    lib = abi.bind(C.CDLL(ORACLE_SO))
    lib.oracle_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
    lib.oracle_philox.restype = None
    for name in ("oracle_exp", "oracle_log", "oracle_sin", "oracle_logfact"):
        fn = getattr(lib, name)
        fn.argtypes = [C.c_double]
        fn.restype = C.c_double
    lib.oracle_poisson.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.oracle_poisson.restype = C.c_uint32
    lib.oracle_sround.argtypes = [C.c_float, C.c_uint32]
    lib.oracle_sround.restype = C.c_uint32
    lib.oracle_source_counts.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_uint64, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint32)]
    lib.oracle_source_counts.restype = C.c_uint32
    lib.oracle_tables.argtypes = [abi.EnginePtr, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32,
                                  C.POINTER(C.c_uint32)]
    lib.oracle_tables.restype = C.c_int
    lib.oracle_dens.argtypes = [abi.EnginePtr, C.c_uint32, C.POINTER(C.c_float), C.c_uint32]
    lib.oracle_dens.restype = C.c_int
    _lib = lib
    return lib


def load_native_oracle():
    """The oracle built -O3 -march=native on this machine (make native, into
    oracle/_build_native), for timing the CPU baseline on the host that runs
    it; (library, build flags).  Falls back to the portable build when the
    native one cannot be built."""
    path = os.path.join(ROOT, "oracle", "_build_native", "libarbo_oracle.so")
    res = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "native"], capture_output=True, text=True)
    if res.returncode != 0 or not os.path.exists(path):
        return load_oracle(), "g++ -O3 -march=x86-64-v2"
    return abi.bind(C.CDLL(path)), "g++ -O3 -march=native"
