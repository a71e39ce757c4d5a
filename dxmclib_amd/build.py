"""Build recipe for the native libraries of dxmclib_amd.

libdxmc_hip.so   HIP backend for gfx950 (include/dxmc_hip.h)
libdxmc_host.so  C++ host API + C facade (include/dxmc_host.h)

Run `python -m dxmclib_amd.build` (or call build_all()). Objects are built
in-tree next to this file so that the libraries travel with the repository.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DXMC_OFFLOAD_ARCH", "gfx950")

HIP_LIB = os.path.join(HERE, "libdxmc_hip.so")
HOST_LIB = os.path.join(HERE, "libdxmc_host.so")

# -packed-fp32-ops keeps the compiler from emitting packed FP32 arithmetic.
HIP_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
    "-Wno-unused-result", "-Wno-unused-value",
]
# per source: the FP64 plain path is compiled without FMA contraction so it
# rounds like the CPU oracle; the FP32 transport kernel may contract and uses
# the hardware transcendental instructions (v_exp_f32, v_log_f32, v_sin_f32,
# v_cos_f32) and reciprocal-based division and square root, and flushes FP32
# denormals (no FP32 value of the transport comes near 1e-38; with denormals
# kept every division is a scaled reciprocal of about six instructions).
HIP_SOURCES = {
    "dxmc_hip.hip": ["-ffp-contract=off"],
    "tables.hip": ["-ffp-contract=off"],
    "transport_fp32.hip": ["-fgpu-approx-transcendentals", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"],
    "autotune.hip": [],
    "world.hip": [],
    "noalloc.hip": ["-ffp-contract=off"],
}
HOST_SOURCES = ["src/material.cpp", "src/host_capi.cpp", "src/server.cpp"]
# -ftree-vectorize: integer passes over voxels (the material-index check)
# vectorise; FP reductions are not reordered without -ffast-math, so values
# are unchanged
HOST_FLAGS = ["-O2", "-ftree-vectorize", "-std=c++20", "-fPIC", "-ffp-contract=off",
              "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src, extra in HIP_SOURCES.items():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(HERE, "csrc", src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)
    return HIP_LIB


def build_host(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src).replace(".cpp", ".o"))
        objs.append(obj)
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(HERE, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    # the host library calls the HIP backend through its C ABI
    _run(["g++", "-shared", "-fPIC", *objs, "-o", HOST_LIB + ".tmp", "-L", HERE, "-ldxmc_hip",
          "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    return HOST_LIB


# programs written against the C++ host API the way a user of the reference
# API writes them (tests/cpp), linked to the two libraries
PROGRAMS = {"layered_pencil": os.path.join(ROOT, "tests", "cpp", "layered_pencil.cpp"),
            "progress_preview": os.path.join(ROOT, "tests", "cpp", "progress_preview.cpp"),
            "filter_weights": os.path.join(ROOT, "tests", "cpp", "filter_weights.cpp"),
            "copy_helpers": os.path.join(ROOT, "tests", "cpp", "copy_helpers.cpp"),
            "threads_cancel": os.path.join(ROOT, "tests", "cpp", "threads_cancel.cpp")}
PROGRAM_DIR = os.path.join(ROOT, "tests", "cpp", "bin")


def build_programs(verbose=False):
    os.makedirs(PROGRAM_DIR, exist_ok=True)
    out = []
    for name, src in PROGRAMS.items():
        exe = os.path.join(PROGRAM_DIR, name)
        _run(["g++", "-O2", "-std=c++20", "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"), src,
              "-o", exe + ".tmp", "-L", HERE, "-ldxmc_host", "-ldxmc_hip", "-Wl,-rpath,$ORIGIN/../../../dxmclib_amd", "-lpthread"])
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build_all(verbose=False):
    return build_hip(verbose), build_host(verbose), build_programs(verbose)


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
