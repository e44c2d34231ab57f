"""The C++ a maintainer would add on the reference's side
compiles against this library's headers and links against the built
library.  The reference's own classes are stood in for by declarations with
the members the snippets use, of the reference's types
(SimulationParam.h:12-58, parameter.h:11-19, ClimateData.h:11-19,
constant.h:8,17); the reference's MCMC class is stood in for by the one
method the binding calls."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")

PRELUDE = r"""
#include <cmath>
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

#define C_MAXSIMRUN 30
#define C_STRAINS 1

class SimulationParam {
public:
    float nSize, metaPop, minMosToHuman, mosNetGrowth;
    float nShapeInfantMortality, nScaleInfantMortality, nScaleLifeExpectancy, nShapeLifeExpectancy, nLocWeibull;
    float mScaleLifeExpectancy, mShapeLifeExpectancy;
    float netPower, netM, netSeed, nHeteroPopSize, mHeteroPopSize;
    float initialSeroPrev, mInitialInf, nInitialInf, bitingRate, mnBitingSuccess, nmBitingSuccess;
    float recovery, mExposed, nExposed, externalInfection, longDistance, exIncPeriodRange, noDistance;
    float scaleEIP, scaleMLE, scaleCC, scaleHumidity;
};
class Parameter {
public:
    SimulationParam simulation;
};
class ClimateData {
public:
    uint32_t maxTime;
    std::vector<float> temperature, humidity, precipitation;
};
class MCMC {
public:
    void storeIncidenceHost(const uint32_t* incidence, uint32_t run) { (void)incidence; (void)run; }
};
"""

MAIN = r"""
int main() { return 0; }
"""


def _blocks():
    text = open(DOC).read()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


def test_document_has_the_bindings():
    blocks = _blocks()
    assert len(blocks) == 3
    assert "arbo_batch_init" in blocks[0] and "arbo_mcmc_run" in blocks[1] and "ArboClimate" in blocks[2]


def _compile(tmp_path, src_name, body):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not found")
    src = tmp_path / src_name
    src.write_text(PRELUDE + body)
    obj = tmp_path / (src_name + ".o")
    res = subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c",
                          str(src), "-o", str(obj)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return obj


@pytest.mark.parametrize("k", [0, 1, 2])
def test_snippet_compiles(tmp_path, k):
    _compile(tmp_path, f"snippet{k}.cpp", _blocks()[k])


def test_snippets_link_against_the_library(tmp_path):
    from arbo_gpu_fit_amd import build
    if not os.path.exists(build.LIB):
        pytest.skip("libarbo_hip.so not built")
    objs = [str(_compile(tmp_path, f"snippet{k}.cpp", b)) for k, b in enumerate(_blocks())]
    objs.append(str(_compile(tmp_path, "main.cpp", MAIN)))
    libdir = os.path.dirname(build.LIB)
    res = subprocess.run(["g++", "-o", str(tmp_path / "driver")] + objs +
                         ["-L", libdir, "-larbo_hip", f"-Wl,-rpath,{libdir}", "-Wl,-rpath-link,/opt/rocm/lib"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
