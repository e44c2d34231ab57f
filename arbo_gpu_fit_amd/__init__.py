"""arbo_gpu_fit_amd — MI355X-native engine for the per-day hot path of a
spatial, multi-serotype dengue individual-based model.

The engine itself is C++/HIP behind the C ABI of include/arbo_c.h
(libarbo_hip.so); this package holds its sources (csrc/), the build recipe
and the host-side Python mirror of the reference's simulation driver.
"""
from .abi import ABI_VERSION, HIP_LIB_PATH, Params, load_hip_library  # noqa: F401
from .engine import ArboError, Engine, SimSpec, simulation, weekly_incidence  # noqa: F401
