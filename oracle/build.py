"""Build recipe of the oracle (test infrastructure).

oracle/liboracle.so        the CPU restatement (oracle.cpp) + the shared
                           physics data provider (dxmclib_amd/src/material.cpp)
oracle/_ref/libdxmc_ref.so ref_driver.cpp compiled against the reference's
                           own headers from REFERENCE (default
                           /root/reference/include); skipped when that
                           checkout is absent (the prebuilt file is used)

Both are built with -ffp-contract=off and no -march=native, so the same
binary runs on any x86-64 host and rounds like the plain FP64 device path.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REFERENCE = os.environ.get("DXMC_REFERENCE_INCLUDE", "/root/reference/include")
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libdxmc_ref.so")

FLAGS = ["-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off", "-pthread"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + r.stdout)


def build_oracle():
    _run(["g++", *FLAGS, "-I", os.path.join(ROOT, "dxmclib_amd", "include"), "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "oracle.cpp"), os.path.join(ROOT, "dxmclib_amd", "src", "material.cpp"),
          "-o", ORACLE_LIB + ".tmp"])
    os.replace(ORACLE_LIB + ".tmp", ORACLE_LIB)
    return ORACLE_LIB


def build_reference():
    """Compile the reference-header driver; returns None without the reference."""
    if not os.path.isdir(os.path.join(REFERENCE, "dxmc")):
        return None
    os.makedirs(REF_DIR, exist_ok=True)
    _run(["g++", *FLAGS, "-I", REFERENCE, os.path.join(HERE, "ref_driver.cpp"), "-o", REF_LIB + ".tmp"])
    os.replace(REF_LIB + ".tmp", REF_LIB)
    return REF_LIB


def build_all():
    return build_oracle(), build_reference()


if __name__ == "__main__":
    print(build_all())
    sys.exit(0)
