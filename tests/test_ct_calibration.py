"""CT dose calibration (CTBaseSource::ctCalibration, reference
source.hpp:925-988, on CTDIPhantom, world.hpp:229-390): the nested transport
of a CT source through the CTDI phantom with forced interactions in the
measurement holes, and CTDIvol / CTDIw / mean beam weight.

CPU: the phantom geometry (disc, holes, measurement map).
GPU: the calibration factor the host API applies in DOSE mode equals the
factor restated here from the same phantom transported through the device C
ABI (identical streams, FP64), and the centre/periphery ratio is physical;
the factor's CTDIw is cached per the calibration run's input;
that transport of the CTDI phantom (forced scoring in the holes) equals the
oracle's on identical streams at reduced histories.
"""
import ctypes as C
import math

import numpy as np
import pytest

from dxmclib_amd import backend, capi, workloads


def _phantom(diameter):
    dims = (C.c_uint64 * 3)()
    sp = np.zeros(3)
    lib = capi.host()
    assert lib.dxmc_host_ctdi_phantom(diameter, dims, sp.ctypes.data_as(C.POINTER(C.c_double)), None, None, None, None, 0) == 0
    n = dims[0] * dims[1] * dims[2]
    mat = np.zeros(n, dtype=np.uint8)
    dens = np.zeros(n)
    meas = np.zeros(n, dtype=np.uint8)
    hole = np.zeros(n, dtype=np.uint8)
    u8 = C.POINTER(C.c_uint8)
    assert lib.dxmc_host_ctdi_phantom(diameter, dims, sp.ctypes.data_as(C.POINTER(C.c_double)), mat.ctypes.data_as(u8),
                                      dens.ctypes.data_as(C.POINTER(C.c_double)), meas.ctypes.data_as(u8), hole.ctypes.data_as(u8), n) == 0
    shape = (dims[2], dims[1], dims[0])
    return mat.reshape(shape), dens.reshape(shape), meas.reshape(shape), hole.reshape(shape), tuple(sp)


def test_ctdi_phantom_geometry():
    mat, dens, meas, hole, sp = _phantom(320)
    assert mat.shape == (60, 323, 323) and sp == (1.0, 1.0, 2.5)
    # PMMA disc of radius 160 mm: area within 1 % of pi r^2 (holes are air)
    slice_ = mat[30]
    pmma_or_hole = np.count_nonzero(slice_ >= 1)
    assert abs(pmma_or_hole - math.pi * 160 ** 2) / (math.pi * 160 ** 2) < 0.01
    # five holes of radius 6.55 mm, air, measured on |z| <= 50 mm (40 slices)
    for h in range(1, 6):
        vox = hole == h
        per_slice = np.count_nonzero(vox[30])
        assert abs(per_slice - math.pi * 6.55 ** 2) < 0.2 * math.pi * 6.55 ** 2
        assert np.count_nonzero(vox.any(axis=(1, 2))) == 40
        assert np.all(mat[vox] == 1) and np.allclose(dens[vox], 0.001205)
    assert np.array_equal(meas > 0, hole > 0)


def _ct_spec(histories_per_exposure, step_deg, ctdi_vol, tube_voltage=120.0, **kw):
    ang, w = workloads.bowtie_profile()
    return backend.source_spec(capi.SOURCE_CT_AXIAL, tube_voltage=tube_voltage, al_mm=6.0, sdd=1190.0, fov=500.0, collimation=38.4,
                               exposure_angle_step_deg=step_deg, histories_per_exposure=histories_per_exposure, bowtie_angles=ang,
                               bowtie_weights=w, heel_effect=1, ctdi_vol=ctdi_vol, **kw)


def _target_world():
    # a small target world: the factor does not depend on it
    return backend.host_world(np.ones((8, 16, 16), dtype=np.uint8), np.ones((8, 16, 16), dtype=np.float32), (2.0, 2.0, 2.0),
                              ["Air, Dry (near sea level)", "Water, Liquid"])


@pytest.mark.gpu
def test_ct_calibration_factor(device):
    per, step, ctdivol = 100_000, 10.0, 10.0
    spec = _ct_spec(per, step, ctdivol)
    tw = _target_world()
    cal = backend.transport(tw, spec, precision=capi.PRECISION_FP64, output_mode=capi.OUTPUT_DOSE, use_source_calibration=True)
    raw = backend.transport(tw, spec, precision=capi.PRECISION_FP64, output_mode=capi.OUTPUT_DOSE, use_source_calibration=False)
    m = raw["dose"] > 0
    factors = cal["dose"][m] / raw["dose"][m]
    factor = float(np.median(factors))
    assert np.allclose(factors, factor, rtol=1e-12)

    # the same calibration restated: mean beam weight of the source, then
    # the source moved to the phantom centre (cosines (-1,0,0,0,0,1)), one
    # rotation, histories raised to at least 1e8 in total, FP64, seed 42
    names = ["Air, Dry (near sea level)", "Air, Dry (near sea level)", "Polymethyl Methacralate (Lucite, Perspex)"]
    t0 = backend.HostTables(names, [0.001205, 0.001205, 1.19], 120.0, 1.0)
    mean_weight = float(np.mean([e.weight for e in t0.exposures(spec)]))
    per_rotation = int(2 * math.pi / (step * math.pi / 180.0))
    stat = max(1, 100_000_000 // (per_rotation * per))
    cspec = _ct_spec(per * stat, step, ctdivol, direction_cosines=(-1, 0, 0, 0, 0, 1), position=(0.0, 0.0, 0.0), scan_length=38.4)
    mat, dens, meas, hole, sp = _phantom(320)
    ext = backend.world_extent(mat.shape[::-1], sp)
    maxd = [float(dens[mat == i].max()) for i in range(3)]
    t = backend.HostTables(names, maxd, 120.0, 1.0)
    exps = t.exposures(cspec)
    assert len(exps) == per_rotation
    w = device.upload_world(mat, dens, sp, ext, 3, measurement=meas)
    tt = device.upload_tables(t.desc())
    tally = device.tally(mat.size)
    try:
        stats = device.run(w, tt, exps, tally, capi.PRECISION_FP64, 1, seed=42)
        device.normalize(tally, w, capi.OUTPUT_DOSE, stats.histories, 1.0)
        d, _, _ = device.download(tally, mat.size)
    finally:
        device.destroy_tally(tally)
        device.destroy_tables(tt)
        device.destroy_world(w)
    flat = hole.ravel()
    means = [d[flat == h].mean() for h in range(1, 6)]
    ctdiw = (means[0] + 2 * np.mean(means[1:])) / 3 * 100.0 / 38.4
    expected = ctdivol / ctdiw / mean_weight
    print(f"calibration factor {factor:.6e}, restated {expected:.6e}; centre/periphery {means[0] / np.mean(means[1:]):.3f}", flush=True)
    assert abs(factor - expected) <= 1e-9 * expected
    assert 0.25 < means[0] / np.mean(means[1:]) < 0.9


@pytest.mark.gpu
def test_ct_calibration_cached(device):
    """The CTDIw of a calibration run is kept per the run's complete input
    (exposures and tables of the nested transport, model, phantom): a
    repeated DOSE-mode call reuses it (the same factor in a fraction of the
    time), another CTDIvol scales the factor exactly, another tube voltage
    runs a calibration of its own."""
    import time
    tw = _target_world()

    def factor(spec):
        t0 = time.perf_counter()
        cal = backend.transport(tw, spec, precision=capi.PRECISION_FP64, output_mode=capi.OUTPUT_DOSE, use_source_calibration=True)
        dt = time.perf_counter() - t0
        raw = backend.transport(tw, spec, precision=capi.PRECISION_FP64, output_mode=capi.OUTPUT_DOSE, use_source_calibration=False)
        m = raw["dose"] > 0
        return float(np.median(cal["dose"][m] / raw["dose"][m])), dt

    # settings no other test uses, so the first call is not already cached
    f1, t1 = factor(_ct_spec(50_000, 12.0, 10.0))
    f2, t2 = factor(_ct_spec(50_000, 12.0, 10.0))
    f3, _ = factor(_ct_spec(50_000, 12.0, 25.0))
    f4, t4 = factor(_ct_spec(50_000, 12.0, 10.0, tube_voltage=100.0))
    print(f"factor {f1:.6e} in {t1:.3f} s, again {f2:.6e} in {t2:.3f} s; CTDIvol x2.5 {f3 / f1:.12f}; 100 kV {f4:.6e} in {t4:.3f} s", flush=True)
    assert f2 == pytest.approx(f1, rel=1e-12) and t2 < 0.5 * t1
    assert f3 == pytest.approx(2.5 * f1, rel=1e-12)
    assert t4 > 2 * t2 and abs(f4 / f1 - 1) > 1e-3


@pytest.mark.gpu
def test_ctdi_transport_against_oracle(device):
    """The nested transport of the calibration (CT rotation at the phantom
    centre, forced interactions in the measurement holes, source.hpp:925-988
    on world.hpp:229-390) at reduced histories, FP64 device path against the
    oracle on identical streams: totals, the five hole doses and the CTDIw
    built from them agree, so the device doses the calibration factor comes
    from are the oracle's."""
    import oracle
    per, step = 2000, 10.0
    cspec = _ct_spec(per, step, 10.0, direction_cosines=(-1, 0, 0, 0, 0, 1), position=(0.0, 0.0, 0.0), scan_length=38.4)
    names = ["Air, Dry (near sea level)", "Air, Dry (near sea level)", "Polymethyl Methacralate (Lucite, Perspex)"]
    mat, dens, meas, hole, sp = _phantom(320)
    ext = backend.world_extent(mat.shape[::-1], sp)
    maxd = [float(dens[mat == i].max()) for i in range(3)]
    t = backend.HostTables(names, maxd, 120.0, 1.0)
    exps = t.exposures(cspec)
    print(f"CTDI phantom: {len(exps)} exposures x {per} histories", flush=True)
    w = device.upload_world(mat, dens, sp, ext, 3, measurement=meas)
    tt = device.upload_tables(t.desc())
    tally = device.tally(mat.size)
    try:
        stats = device.run(w, tt, exps, tally, capi.PRECISION_FP64, 1, seed=42)
        d, _, v = device.download(tally, mat.size)
    finally:
        device.destroy_tally(tally)
        device.destroy_tables(tt)
        device.destroy_world(w)
    spectrum = _spectrum(cspec)
    orc = oracle.Oracle(names, maxd, 120.0, 1.0)
    orc.add_source_tables(t.source_arrays(), [spectrum])
    ow = oracle.world_view(mat, dens, sp, ext, measurement=meas)
    ref = orc.transport(ow, exps, L=1, seed=42)
    rel = abs(d.sum() - ref["dose"].sum()) / ref["dose"].sum()
    flat = hole.ravel()
    dev_holes = np.array([d[flat == h].mean() for h in range(1, 6)])
    ref_holes = np.array([ref["dose"][flat == h].mean() for h in range(1, 6)])
    hole_rel = np.abs(dev_holes / ref_holes - 1)
    ctdiw = lambda m: (m[0] + 2 * np.mean(m[1:])) / 3 * 100.0 / 38.4  # noqa: E731
    print(f"CTDI FP64 vs oracle: total rel {rel:.2e}, hole rel {hole_rel}, steps {stats.steps} vs {ref['steps']}, "
          f"events {stats.events} vs {ref['events']}", flush=True)
    assert rel < 5e-3
    assert np.all(hole_rel < 5e-3)
    assert abs(ctdiw(dev_holes) / ctdiw(ref_holes) - 1) < 2e-3
    assert abs(stats.steps - ref["steps"]) <= 1e-3 * ref["steps"]
    assert abs(stats.events - ref["events"]) <= 1e-3 * ref["events"]
    # every hole scores (forced interactions), on the same voxels on both sides
    for h in range(1, 6):
        assert np.count_nonzero(ref["dose"][flat == h]) > 0
    same = np.mean((d[flat > 0] > 0) == (ref["dose"][flat > 0] > 0))
    assert same > 0.999


def _spectrum(spec):
    n = C.c_uint64()
    lib = capi.host()
    lib.dxmc_host_source_spectrum(C.byref(spec), None, None, 0, C.byref(n))
    e, w = np.zeros(n.value), np.zeros(n.value)
    lib.dxmc_host_source_spectrum(C.byref(spec), e.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n))
    return w, e
