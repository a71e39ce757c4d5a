"""CT sources of the C++ host API: axial and helical single- and dual-source
CT, the topogram, and the XCare exposure weight (reference source.hpp:
1000-1700). The reference's CT sources need its xraylib-backed Material
(CTBaseSource filters its tube with aluminium), so they do not build here;
their exposure tables are pinned by the geometry the reference code fixes:

* each exposure of tube A of a dual source is the exposure of a single-source
  CT with tube A's settings; each of tube B is the single-source exposure at
  tube B's start angle (the tube position with the source's SDD, as in the
  reference) with tube B's fan angles (field of view and SDD of B);
* exposures alternate A, B; totals are 2 x angles per rotation x rotations
  (axial) and 2 x floor(scan 2 pi / (collimation pitch step)) (helical);
* the tube weights are 2 w_X / (w_A + w_B), w_X = mAs_X times the sum of tube
  X's unnormalised spectrum (the oracle's literal spectrum integral), and
  each tube's alias table is built from its spectrum divided by that sum;
* a topogram holds max(ceil(scan length), 1) exposures at the start angle,
  z evenly from 0 to the scan length;
* XCare multiplies each exposure's weight by the filter weight at its gantry
  angle, equal bit for bit to the reference's XCareFilter (oracle/_ref).
"""
import math
import os

import numpy as np
import pytest

import oracle
from dxmclib_amd import backend, capi

NAMES = ["Water, Liquid"]
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "anode_reference.npz")


def _exposures(spec):
    t = backend.HostTables(NAMES, [1.0], 150.0, 1.0)
    return t, t.exposures(spec)


def _field(ex, name):
    return np.array([list(getattr(e, name)) for e in ex])


def _scalar(ex, name):
    return np.array([getattr(e, name) for e in ex])


CT = dict(tube_voltage=120.0, al_mm=6.0, sdd=1000.0, fov=500.0, collimation=40.0, start_angle_deg=10.0, exposure_angle_step_deg=3.0,
          histories_per_exposure=100)
DUAL_B = dict(tube_b_voltage=80.0, tube_b_al_mm=4.0, sdd_b=900.0, fov_b=330.0, start_angle_b_deg=100.0, start_angle_b_set=1)


@pytest.mark.parametrize("helical", [False, True])
def test_dual_source_geometry(helical):
    single_type = capi.SOURCE_CT_SPIRAL if helical else capi.SOURCE_CT_AXIAL
    dual_type = capi.SOURCE_CT_SPIRAL_DUAL if helical else capi.SOURCE_CT_AXIAL_DUAL
    extra = dict(pitch=0.8, scan_length=95.0) if helical else dict(scan_length=80.0, axial_step=40.0)
    _, dual = _exposures(backend.source_spec(dual_type, **CT, **DUAL_B, **extra))
    _, a = _exposures(backend.source_spec(single_type, **CT, **extra))
    b_pose = dict(CT, start_angle_deg=DUAL_B["start_angle_b_deg"])
    _, b = _exposures(backend.source_spec(single_type, **b_pose, **extra))
    b_fan = dict(CT, sdd=DUAL_B["sdd_b"], fov=DUAL_B["fov_b"])
    _, bf = _exposures(backend.source_spec(single_type, **b_fan, **extra))
    n = len(a)
    if helical:
        per = math.floor(95.0 * 2 * math.pi / (40.0 * 0.8 * math.radians(3.0)))
    else:
        per = int(2 * math.pi / math.radians(3.0)) * 2
    assert n == per and len(dual) == 2 * n
    ea, eb = dual[0::2], dual[1::2]
    for f in ("position", "direction_cosines", "beam_direction"):
        assert np.array_equal(_field(ea, f), _field(a, f)), f
        assert np.array_equal(_field(eb, f), _field(b, f)), f
    assert np.array_equal(_field(ea, "collimation_angles"), _field(a, "collimation_angles"))
    assert np.array_equal(_field(eb, "collimation_angles"), _field(bf, "collimation_angles"))
    # the tubes' own spectrum and heel tables
    assert len(set(_scalar(ea, "spectrum"))) == 1 and len(set(_scalar(eb, "spectrum"))) == 1
    assert _scalar(ea, "spectrum")[0] != _scalar(eb, "spectrum")[0]
    assert _scalar(ea, "heel")[0] != _scalar(eb, "heel")[0]
    # tube B a quarter turn ahead of A, both at SDD/2 from the isocentre
    pa, pb = _field(ea, "position"), _field(eb, "position")
    assert np.allclose(np.hypot(pa[:, 0], pa[:, 1]), 500.0) and np.allclose(np.hypot(pb[:, 0], pb[:, 1]), 500.0)
    cosang = np.sum(pa[:, :2] * pb[:, :2], axis=1) / 500.0 ** 2
    assert np.allclose(cosang, math.cos(math.radians(90.0)), atol=1e-12)


def _raw_sum_and_weights(kvp, al_mm):
    anode = dict(np.load(FIXTURE))
    energies = np.arange(1.0, kvp + 0.5, 1.0)
    raw = oracle.tube_spectrum(anode, kvp, math.radians(12.0), energies, [(13, al_mm)], normalize=False)
    total = 0.0
    for v in raw:  # std::accumulate: left to right
        total += v
    return total, raw / total, energies


def test_dual_source_weights_and_spectra():
    _, e = _exposures(backend.source_spec(capi.SOURCE_CT_AXIAL_DUAL, **CT, **DUAL_B, mas_a=150.0, mas_b=60.0))
    sum_a, wa, ea = _raw_sum_and_weights(120.0, 6.0)
    sum_b, wb, eb = _raw_sum_and_weights(80.0, 4.0)
    wA, wB = 150.0 * sum_a, 60.0 * sum_b
    assert e[0].weight == pytest.approx(wA * 2 / (wA + wB), rel=1e-14)
    assert e[1].weight == pytest.approx(wB * 2 / (wA + wB), rel=1e-14)
    assert e[0].weight + e[1].weight == pytest.approx(2.0, rel=1e-15)
    # each tube's alias table is the one the oracle builds from spectrum / sum
    t, _ = _exposures(backend.source_spec(capi.SOURCE_CT_AXIAL_DUAL, **CT, **DUAL_B, mas_a=150.0, mas_b=60.0))
    src = t.source_arrays()
    off = src["spectrum_offset"]
    assert len(off) == 3
    for k, (w, en) in enumerate(((wa, ea), (wb, eb))):
        prob, alias = oracle.alias_table(w)
        assert np.array_equal(src["spectrum_energy"][off[k]:off[k + 1]], en)
        assert np.array_equal(src["spectrum_probability"][off[k]:off[k + 1]], prob)
        assert np.array_equal(src["spectrum_alias"][off[k]:off[k + 1]], alias)
    # equal tubes and currents: both weights 1
    same = dict(DUAL_B, tube_b_voltage=120.0, tube_b_al_mm=6.0)
    _, e2 = _exposures(backend.source_spec(capi.SOURCE_CT_AXIAL_DUAL, **CT, **same))
    assert e2[0].weight == e2[1].weight == 1.0


@pytest.mark.parametrize("scan", [37.0, 37.5, 0.4])
def test_topogram(scan):
    _, e = _exposures(backend.source_spec(capi.SOURCE_CT_TOPOGRAM, **CT, scan_length=scan))
    n = max(math.ceil(scan), 1)
    assert len(e) == n
    pos = _field(e, "position")
    # fixed gantry angle (the start angle), z from 0 to the scan length
    ang = np.degrees(np.arctan2(pos[:, 0], -pos[:, 1]))
    assert np.allclose(ang, ang[0], atol=1e-12)
    assert np.allclose(np.hypot(pos[:, 0], pos[:, 1]), 500.0)
    z = pos[:, 2] - pos[0, 2]
    expect = np.linspace(0.0, scan, n) if n > 1 else np.zeros(1)
    assert np.allclose(z, expect, atol=1e-12)
    assert np.all(_scalar(e, "weight") == 1.0)
    assert np.array_equal(_field(e, "direction_cosines"), np.repeat(_field(e[:1], "direction_cosines"), n, axis=0))


@pytest.mark.parametrize("kind", [capi.SOURCE_CT_AXIAL, capi.SOURCE_CT_SPIRAL, capi.SOURCE_CT_AXIAL_DUAL])
def test_xcare_exposure_weights(kind):
    ref = oracle.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout at build time)")
    xc = dict(xcare=1, xcare_filter_angle_deg=40.0, xcare_span_angle_deg=150.0, xcare_ramp_angle_deg=25.0, xcare_low_weight=0.45)
    extra = dict(scan_length=40.0) if kind != capi.SOURCE_CT_SPIRAL else dict(scan_length=60.0, pitch=1.0)
    _, on = _exposures(backend.source_spec(kind, **CT, **xc, **extra))
    _, off = _exposures(backend.source_spec(kind, **CT, **extra))
    dual = kind == capi.SOURCE_CT_AXIAL_DUAL
    step = math.radians(3.0)
    per = int(2 * math.pi / step)
    got, want = [], []
    for i, (a, b) in enumerate(zip(on, off)):
        k = i // 2 if dual else i
        start = math.radians(CT["start_angle_deg"] + (90.0 if dual and i % 2 else 0.0))
        angle = start + step * (k % per if kind != capi.SOURCE_CT_SPIRAL else k)
        w = ref.ref_xcare_weight(math.radians(40.0), math.radians(150.0), math.radians(25.0), 0.45, angle)
        got.append(a.weight)
        want.append(b.weight * w)
    assert np.array_equal(np.array(got), np.array(want))
    assert min(got) < 0.5 < max(got)
