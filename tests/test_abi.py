"""The C ABI boundary (CPU): libarbo_hip.so loads and exports every entry point
include/*.h declares (no compute is run), the oracle exports the engine ABI
(include/arbo_c.h), and the ctypes mirrors match the C struct layouts (sizes
and offsets printed by a program compiled from the headers)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from arbo_gpu_fit_amd import abi, data, mcmc, output, surrogate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"arbo_c.h": abi.PROTOTYPES, "arbo_mcmc.h": mcmc.PROTOTYPES, "arbo_surrogate.h": surrogate.PROTOTYPES,
           "arbo_output.h": output.PROTOTYPES, "arbo_data.h": data.PROTOTYPES}


def declared_functions(header=None):
    names = set()
    for h in ([header] if header else sorted(os.listdir(os.path.join(ROOT, "include")))):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"ARBO_API\s+[\w\s\*]+?\b(arbo_\w+)\s*\(", text))
    return sorted(names)


def test_headers_declare_the_abi():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == sorted(HEADERS)
    for h, protos in HEADERS.items():
        names = declared_functions(h)
        assert set(names) == set(protos), (h, set(names) ^ set(protos))


def test_hip_library_exports_every_symbol():
    if not os.path.exists(abi.HIP_LIB_PATH):
        pytest.skip("libarbo_hip.so not built")
    lib = C.CDLL(abi.HIP_LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    abi.bind(lib)
    mcmc.bind(lib)
    surrogate.bind(lib)
    output.bind(lib)
    data.bind(lib)
    out = subprocess.run(["nm", "-D", "--defined-only", abi.HIP_LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert {n for n in exported if n.startswith("arbo_")} == set(declared_functions())


def test_oracle_exports_the_engine_abi(oracle_lib):
    for name in declared_functions("arbo_c.h"):
        assert hasattr(oracle_lib, name), name


def test_create_rejects_bad_descriptor(oracle_lib):
    h = abi.EnginePtr()
    desc = abi.EngineDesc(abi.ABI_VERSION + 7, 0, 0, 0)
    assert oracle_lib.arbo_create(C.byref(desc), C.byref(h)) == abi.E_INVALID
    assert oracle_lib.arbo_create(None, C.byref(h)) == abi.E_INVALID


LAYOUT_PROGRAM = r"""
#include <stddef.h>
#include <stdio.h>
#include "arbo_c.h"
#include "arbo_mcmc.h"
#include "arbo_output.h"
#include "arbo_data.h"
#include "arbo_surrogate.h"
#define F(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f));
int main(void) {
    printf("ArboEngineDesc %zu\nArboParams %zu\nArboShape %zu\nArboSeries %zu\nArboState %zu\n",
           sizeof(ArboEngineDesc), sizeof(ArboParams), sizeof(ArboShape), sizeof(ArboSeries), sizeof(ArboState));
    printf("ArboPrior %zu\nArboMcmcDesc %zu\nArboEngineApi %zu\n", sizeof(ArboPrior), sizeof(ArboMcmcDesc),
           sizeof(ArboEngineApi));
    printf("ArboClimate %zu\n", sizeof(ArboClimate));
    FIELDS
    return 0;
}
"""


def test_ctypes_layout_matches_header(tmp_path):
    structs = {"ArboEngineDesc": abi.EngineDesc, "ArboParams": abi.Params, "ArboShape": abi.Shape,
               "ArboClimate": abi.Climate,
               "ArboSeries": abi.Series, "ArboState": abi.State, "ArboPrior": mcmc.Prior,
               "ArboMcmcDesc": mcmc.McmcDesc, "ArboEngineApi": mcmc.EngineApi}
    fields = "".join(f"F({cname}, {f})" for cname, cls in structs.items() for f, _ in cls._fields_)
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_PROGRAM.replace("FIELDS", fields))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(ln.rsplit(" ", 1) for ln in lines if ln)
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)
