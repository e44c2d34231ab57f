"""Build recipe of libarbo_hip.so (hipcc for gfx950, in-tree).

    python -m arbo_gpu_fit_amd.build

Flags: -ffp-contract=off keeps every float/double operation a separate
IEEE-754 operation (the engine must match the CPU oracle bit for bit);
-fhip-fp32-correctly-rounded-divide-sqrt is hipcc's default and is spelled
out; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(OUT_DIR, "libarbo_hip.so")

SOURCES = ["arbo_kernels.hip", "arbo_engine.cpp", "arbo_tables.cpp", "arbo_mcmc.cpp", "arbo_surrogate.hip", "arbo_output.cpp",
           "arbo_data.cpp"]
HEADERS = ["arbo_rng.h", "arbo_device.h", "arbo_tables.h"]

ARCH = os.environ.get("ARBO_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: str = LIB) -> list[str]:
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fvisibility=hidden", "-Wall",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-o", out,
    ] + [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    """The CPU oracle (test infrastructure) under oracle/_build."""
    res = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if verbose:
        print(res.stdout, res.stderr, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return os.path.join(ROOT, "oracle", "_build", "libarbo_oracle.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
