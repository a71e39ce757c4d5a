"""dxmc_hip_run_noalloc (the run entry point that allocates nothing) against
dxmc_hip_run on the same inputs, through the C ABI: two calls in a row, the
second with fewer exposures and voxels, both precisions. Equal step and event
counts and per-voxel event counts; doses equal up to the order of FP64
atomic additions (1e-12 relative). Limits are refused with clear errors."""
import numpy as np
import pytest

from dxmclib_amd import backend, capi

pytestmark = pytest.mark.gpu

NAMES = ["Water, Liquid", "Bone, Cortical (ICRU)"]
MAXD = [1.0, 1.92]


def _case(n, exposures, seed):
    rng = np.random.default_rng(seed)
    mat = (rng.random((n, n, n)) < 0.15).astype(np.uint8)
    dens = np.where(mat == 1, 1.92, 1.0).astype(np.float32)
    spacing = (1.5, 1.5, 1.5)
    ext = backend.world_extent((n, n, n), spacing)
    spec = backend.source_spec(capi.SOURCE_CT_AXIAL, tube_voltage=110.0, sdd=600.0, fov=200.0, collimation=20.0,
                               exposure_angle_step_deg=360.0 / exposures, histories_per_exposure=4000)
    tables = backend.HostTables(NAMES, MAXD, 110.0, 1.0)
    exps = tables.exposures(spec)
    return mat, dens, spacing, ext, tables, exps


def _both(device, case, precision, seed):
    mat, dens, spacing, ext, tables, exps = case
    w = device.upload_world(mat, dens, spacing, ext, len(NAMES))
    t = device.upload_tables(tables.desc())
    out = []
    try:
        for run in (device.run, device.run_noalloc):
            tally = device.tally(mat.size)
            try:
                stats = run(w, t, exps, tally, precision, 1, seed=seed)
                dose, ev, var = device.download(tally, mat.size)
            finally:
                device.destroy_tally(tally)
            out.append((stats, dose, ev, var))
    finally:
        device.destroy_tables(t)
        device.destroy_world(w)
    return out


@pytest.mark.parametrize("precision", [capi.PRECISION_FP32, capi.PRECISION_FP64])
def test_noalloc_equals_run(device, precision):
    cases = [(_case(40, 90, 1), 5), (_case(24, 36, 2), 6)]  # the second call smaller
    for case, seed in cases:
        print(f"run_noalloc precision={precision} voxels={case[0].size} exposures={len(case[5])}", flush=True)
        (s1, d1, e1, v1), (s2, d2, e2, v2) = _both(device, case, precision, seed)
        assert s1.histories == s2.histories == sum(x.n_histories for x in case[5])
        assert s1.steps == s2.steps > 0 and s1.events == s2.events > 0
        assert np.array_equal(e1, e2)
        assert np.allclose(d1, d2, rtol=1e-12, atol=0)
        assert np.allclose(v1, v2, rtol=1e-12, atol=0)
        assert d2.sum() > 0


def test_noalloc_limits(device):
    mat, dens, spacing, ext, tables, exps = _case(16, 12, 3)
    w = device.upload_world(mat, dens, spacing, ext, len(NAMES))
    t = device.upload_tables(tables.desc())
    tally = device.tally(mat.size)
    try:
        many = (capi.Exposure * (capi.NOALLOC_MAX_EXPOSURES + 1))()
        for i in range(len(many)):
            many[i] = exps[i % len(exps)]
        with pytest.raises(backend.DxmcError) as err:
            device.run_noalloc(w, t, many, tally, capi.PRECISION_FP32)
        assert "exceed the maximum" in str(err.value)
        cfg = capi.RunCfg()
        cfg.flags = capi.RUN_AUTOTUNE
        with pytest.raises(backend.DxmcError) as err:
            device.run_noalloc(w, t, exps, tally, capi.PRECISION_FP32, tuning=cfg)
        assert "tuned before" in str(err.value)
    finally:
        device.destroy_tally(tally)
        device.destroy_tables(t)
        device.destroy_world(w)
