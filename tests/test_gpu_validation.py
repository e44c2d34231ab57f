"""The GPU engine's host checks reject what the oracle rejects (GPU): every
invalid parameter with the same error code, bad community sizes, a day
count of zero, stepping before init and past capacity, and reads of outputs
that were not kept.  Nothing here launches a kernel on invalid input: each
check fails on the host before any launch."""
import numpy as np
import pytest

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import ArboError, Engine, SimSpec
from tests.test_oracle_model import INVALID_PARAMS, small_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hip_lib):
    e = Engine(hip_lib)
    yield e
    e.close()


@pytest.mark.parametrize("field,value,code", INVALID_PARAMS)
def test_invalid_parameters(field, value, code, gpu, oracle_lib):
    print(f"[validation] {field}={value}", flush=True)
    codes = []
    with Engine(oracle_lib) as orc:
        for eng in (gpu, orc):
            spec = small_spec()
            setattr(spec.params, field, value)
            with pytest.raises(ArboError) as ei:
                eng.init([spec], 10)
            codes.append(ei.value.code)
    assert codes == [code, code]


def test_invalid_shapes_and_order(hip_lib):
    print("[validation] shapes and order", flush=True)
    eng = Engine(hip_lib)
    with pytest.raises(ArboError) as ei:
        eng.step(1)
    assert ei.value.code == abi.E_STATE
    for humans in (np.array([0] * 4, np.uint32), np.array([abi.MAX_COMMUNITY_HUMANS + 1] * 4, np.uint32)):
        with pytest.raises(ArboError):
            eng.init([SimSpec(2, 2, humans, configs.base_params(), 1)], 10)
    with pytest.raises(ArboError):
        eng.init([SimSpec(2, 2, np.full(4, 5, np.uint32), configs.base_params(), 1,
                          mosq_slots=np.full(4, abi.MAX_COMMUNITY_SLOTS + 1, np.uint32))], 10)
    with pytest.raises(ArboError):
        eng.init([small_spec()], 0)
    eng.init([small_spec()], 5)
    eng.step(5)
    with pytest.raises(ArboError) as ei:
        eng.step(1)
    assert ei.value.code == abi.E_CAPACITY
    for read in (lambda: eng.series(3), lambda: eng.spatial(0), lambda: eng.history(0), lambda: eng.aoi(0)):
        with pytest.raises(ArboError):
            read()
    eng.close()
