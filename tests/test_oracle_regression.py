"""Drift guard of the oracle (CPU): SHA-256 digests of every parity case's
final census and series, and of the kept outputs (per-community series,
infection history, age-of-infection rings) of the reference's parameter
set, and of a short MCMC run over the oracle engine, against the digests
committed in tests/golden/oracle_digests.json
(made by tests/golden/make_oracle_digests.py).

The digests pin nothing about correctness: the closed forms, known-answer
vectors and distribution tests do that.  They catch a change that moves the
oracle's results at all, which the GPU parity tests cannot see when the same
change moves both sides.  A deliberate change of the model regenerates the
file with the script and says so in its commit."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from arbo_gpu_fit_amd import abi
from arbo_gpu_fit_amd.engine import Engine
from tests.cases import CASES

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_digests.json")


def _digest(arrays: dict) -> str:
    h = hashlib.sha256()
    for k in sorted(arrays):
        a = np.ascontiguousarray(arrays[k])
        h.update(f"{k}:{a.dtype.str}:{a.shape}".encode())
        h.update(a.tobytes())
    return h.hexdigest()


def case_digests(lib, name: str) -> dict:
    """Digests of one case run to its last day on the oracle engine."""
    specs, T = CASES[name]()
    out = {}
    with Engine(lib, threads=4) as e:
        e.init(specs, T)
        e.step(T)
        for i in range(len(specs)):
            out[f"state{i}"] = _digest(e.state(i))
            out[f"series{i}"] = _digest(e.series(i))
    return out


def kept_output_digests(lib) -> dict:
    specs, T = CASES["reference_params"]()
    with Engine(lib, flags=abi.F_SPATIAL | abi.F_AOI, threads=4) as e:
        e.init(specs, T)
        e.step(T)
        return {"spatial": _digest({"v": e.spatial(0)}), "history": _digest({"v": e.history(0)}),
                "aoi": _digest({"v": e.aoi(0)})}


def mcmc_digests(oracle_lib) -> dict:
    """Four iterations of six chains of the configured fit over the oracle
    engine: proposals, log-likelihoods and decisions."""
    from arbo_gpu_fit_amd import mcmc
    from tests.test_mcmc import _tiny_driver
    lib = mcmc.bind(abi.load_hip_library())
    m, orc = _tiny_driver(lib, oracle_lib)
    m.run(orc, 4, api=mcmc.engine_api(oracle_lib))
    th, ll, acc = m.history()
    m.close()
    orc.close()
    return {"chains": _digest({"theta": th, "loglik": ll, "accepted": acc})}


def _golden() -> dict:
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("name", sorted(CASES))
def test_case_outputs_unchanged(name, oracle_lib):
    assert case_digests(oracle_lib, name) == _golden()["cases"][name]


def test_kept_outputs_unchanged(oracle_lib):
    assert kept_output_digests(oracle_lib) == _golden()["kept_outputs"]


def test_mcmc_chains_unchanged(oracle_lib):
    assert mcmc_digests(oracle_lib) == _golden()["mcmc"]


def test_digest_sees_one_changed_word():
    a = {"x": np.arange(10, dtype=np.uint32)}
    b = {"x": a["x"].copy()}
    b["x"][7] ^= 1
    assert _digest(a) != _digest(b)
    assert _digest(a) != _digest({"x": a["x"].astype(np.int32)})
