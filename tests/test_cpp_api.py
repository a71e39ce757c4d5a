"""The C++ host API as a drop-in: tests/cpp/layered_pencil.cpp is written the
way a user of the reference's Transport writes a program (World, Material,
PencilSource, Transport<T>::operator(), reference examples/pencilbeam) and is
linked to libdxmc_host.so / libdxmc_hip.so.

CPU: without a GPU the program fails loudly (no CPU fallback).
GPU: its FP64 result equals the oracle's on identical streams (step and event
counts, energy per slab to 1e-9); its FP32 result agrees to 1e-3.

tests/cpp/progress_preview.cpp watches a ProgressBar from a second thread
while Transport runs (reference validation/validation.cpp:270-285): the
exposure counter only grows and ends at the total, and dose previews from
the device arrive during the run with the projected plane's size.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from dxmclib_amd import backend, build, capi

EXE = os.path.join(build.PROGRAM_DIR, "layered_pencil")
PREVIEW_EXE = os.path.join(build.PROGRAM_DIR, "progress_preview")
THREADS_EXE = os.path.join(build.PROGRAM_DIR, "threads_cancel")
NAMES = ["Air, Dry (near sea level)", "Water, Liquid", "Al"]


def _run(precision, per, exposures):
    assert os.path.exists(EXE), "build() first"
    r = subprocess.run([EXE, str(precision), str(per), str(exposures)], capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout, r.stderr


def test_program_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc, out, err = _run(0, 100, 2)
    assert rc == 1
    assert "HIP" in err and out == ""


def _oracle_slabs(per, exposures):
    import oracle
    dims = (48, 48, 48)
    z = np.arange(48)
    m = np.where(z < 16, 0, np.where(z < 32, 1, 2)).astype(np.uint8)
    mat = np.broadcast_to(m[:, None, None], dims).copy()
    dens_of = []
    out = np.zeros(5)
    for n in NAMES:
        d = C.c_double()
        capi.host().dxmc_host_material_data(n.encode(), 60.0, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(d))
        dens_of.append(d.value)
    dens = np.asarray(dens_of)[mat]
    ext = backend.world_extent(dims[::-1], (1.0, 1.0, 1.0))
    spec = backend.source_spec(capi.SOURCE_PENCIL, position=(0.5, 0.5, -48.0), direction_cosines=(1, 0, 0, 0, 1, 0), photon_energy=60.0,
                               histories_per_exposure=per, total_exposures=exposures)
    t = backend.HostTables(NAMES, dens_of, 60.0, 1.0)
    exps = t.exposures(spec)
    orc = oracle.Oracle(NAMES, dens_of, 60.0, 1.0)
    ow = oracle.world_view(mat, dens, (1.0, 1.0, 1.0), ext)
    r = orc.transport(ow, exps, L=1, seed=42)
    d, _ = oracle.normalize(r["dose"], r["variance"], capi.OUTPUT_EV_PER_HISTORY, per * exposures)
    flat = mat.ravel()
    return r, [d[flat == i].sum() for i in range(3)]


@pytest.mark.gpu
def test_program_matches_oracle():
    per, exposures = 20000, 10
    ref, slabs = _oracle_slabs(per, exposures)
    print(f"oracle steps {ref['steps']} events {ref['events']} slabs {slabs}", flush=True)
    rc, out, err = _run(capi.PRECISION_FP64, per, exposures)
    assert rc == 0, err
    g = json.loads(out)
    print(f"fp64 {g}", flush=True)
    assert g["histories"] == per * exposures
    assert g["steps"] == ref["steps"] and g["events"] == ref["events"]
    for k, name in enumerate(("air", "water", "aluminium")):
        assert abs(g[name] - slabs[k]) <= 1e-9 * abs(slabs[k]) + 1e-12
    rc, out, err = _run(capi.PRECISION_FP32, per, exposures)
    assert rc == 0, err
    f = json.loads(out)
    print(f"fp32 {f}", flush=True)
    for name in ("water", "aluminium"):
        assert abs(f[name] - g[name]) <= 1e-3 * g[name]


def test_preview_program_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert os.path.exists(PREVIEW_EXE), "build() first"
    r = subprocess.run([PREVIEW_EXE, "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    assert "HIP" in r.stderr and r.stdout == ""


@pytest.mark.gpu
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_progress_bar_preview(axis):
    assert os.path.exists(PREVIEW_EXE), "build() first"
    r = subprocess.run([PREVIEW_EXE, str(axis)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    g = json.loads(r.stdout)
    print(g, flush=True)
    dims = (40, 36, 32)  # x, y, z of the program's world
    plane = [d for k, d in enumerate(dims) if k != axis]
    assert g["histories"] == 50 * 20000
    assert g["exposures"] == g["total_exposures"] == 50
    assert g["monotone"]
    assert g["mid_run"] >= 1
    assert [g["width"], g["height"]] == plane
    assert g["max_pixel"] == 255


@pytest.mark.gpu
def test_cancel_returns_empty_result():
    """ProgressBar::setCancel(true) from another thread during a run gives the
    reference's cancelled result (progressbar.hpp:77-84, transport.hpp:
    176-185): numberOfHistories 0 and all-zero arrays of the world's size."""
    assert os.path.exists(THREADS_EXE), "build() first"
    print("threads_cancel cancel", flush=True)
    r = subprocess.run([THREADS_EXE, "cancel"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    g = json.loads(r.stdout)
    print(g, flush=True)
    assert 1 <= g["exposures_at_cancel"] < g["total_exposures"] == 400
    assert g["completed"] < g["total_exposures"]
    assert g["size"] == 40 ** 3
    res = g["result"]
    assert res["histories"] == 0 and res["zero"]
    assert res["dose"] == 0 and res["variance"] == 0 and res["n_events"] == 0


@pytest.mark.gpu
def test_two_threads_share_the_device_context():
    """Two Transport calls on two threads at once (one shared device context,
    each C ABI call holding it) give the results of the same calls run one
    after the other: equal step and event counts and per-voxel event counts,
    doses equal up to the order of FP64 atomic additions."""
    assert os.path.exists(THREADS_EXE), "build() first"
    print("threads_cancel concurrent", flush=True)
    r = subprocess.run([THREADS_EXE, "concurrent"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    g = json.loads(r.stdout)
    print(g, flush=True)
    for k in ("a", "b"):
        c, s = g[f"concurrent_{k}"], g[f"serial_{k}"]
        assert c["histories"] == s["histories"] > 0
        assert c["steps"] == s["steps"] > 0 and c["events"] == s["events"] > 0
        assert c["n_events"] == s["n_events"]
        assert abs(c["dose"] - s["dose"]) <= 1e-12 * abs(s["dose"])
        assert g[f"max_rel_{k}"] < 1e-12
