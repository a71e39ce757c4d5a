"""Drop-in check of the C++ host API against the reference's own callers:
each compiles unchanged (g++ -std=c++20 -fsyntax-only) against
dxmclib_amd/include instead of the reference's include/.

  examples/pencilbeam/pencilbeam.cpp   Transport, World, PencilSource
  validation/validation.cpp            every source type, ProgressBar on a
                                       worker thread, raw I/O
  tests/test{beamfilters,exposure,material,random,source,tube,vectormath,
       world}.cpp                      the reference's unit tests

Left out: tests/testattenuationlut.cpp, testinterpolation.cpp and
testtransport.cpp include xraylib.h (absent here); tests/testdxmclib.cpp and
validatedxmclib.cpp include headers (dxmc/transport.h, dxmc/attenuationlut.h)
that the reference itself no longer has. Skipped without the reference
checkout.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = os.environ.get("DXMC_REFERENCE_ROOT", "/root/reference")
CALLERS = ["examples/pencilbeam/pencilbeam.cpp", "validation/validation.cpp"] + [
    f"tests/test{n}.cpp" for n in ("beamfilters", "exposure", "material", "random", "source", "tube", "vectormath", "world")]


def _compile(rel):
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "dxmclib_amd", "include"), "-I",
                        os.path.join(ROOT, "include"), os.path.join(REFERENCE, rel)], capture_output=True, text=True, timeout=300)
    return rel, r.returncode, r.stderr


def test_reference_callers_compile():
    if not os.path.isdir(os.path.join(REFERENCE, "examples")):
        pytest.skip("reference checkout not present")
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_compile, CALLERS))
    failed = [(rel, err[:2000]) for rel, rc, err in results if rc != 0]
    assert not failed, failed
