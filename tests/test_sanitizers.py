"""Sanitizer runs of the CPU oracle: AddressSanitizer + UndefinedBehavior-
Sanitizer on one thread and ThreadSanitizer over the OpenMP paths, both over
the driver oracle/sanitize_main.cpp (every model switch of the parity
cases, the kept outputs, a checkpoint written back).  `make asan` and `make
tsan` in oracle/ build and run them; a report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _make(target: str):
    res = subprocess.run(["make", "-C", ORACLE, target], capture_output=True, text=True, timeout=900)
    assert res.returncode == 0, f"make {target} failed:\n{res.stdout[-3000:]}\n{res.stderr[-6000:]}"
    assert "all cases ran clean" in res.stderr


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    _make("asan")


@pytest.mark.skipif(not os.path.exists("/opt/rocm/lib/llvm/lib/libarcher.so") or not shutil.which("make"),
                    reason="LLVM's OpenMP runtime with its Archer tool is needed for ThreadSanitizer")
def test_oracle_openmp_under_thread_sanitizer():
    _make("tsan")
