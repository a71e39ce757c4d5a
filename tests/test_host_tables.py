"""Host table builders (the product's C++ AttenuationLut / RITA / spline /
shell / alias code) against the oracle's independent restatement: bit-exact.

Reference: AttenuationLut::generate (attenuationlut.hpp:64-96),
AttenuationLutInterpolator (attenuationinterpolator.hpp:48-205),
RITA (dxmcrandom.hpp:384-468), CubicSplineInterpolator
(interpolation.hpp:123-168), Material::getElectronConfiguration
(material.cpp:143-273), RandomDistribution (dxmcrandom.hpp:229-274).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from dxmclib_amd import backend, capi, workloads

SETS = [
    (["Water, Liquid"], [1.0], 60.0),
    (["Water, Liquid", "Bone, Cortical (ICRU)", "Lung (ICRU)", "Air, Dry (near sea level)", "Adipose Tissue (ICRP)"],
     [1.0, 1.92, 0.26, 0.001205, 0.95], 120.0),
    (["Air, Dry (near sea level)", "Air, Dry (near sea level)", "Polymethyl Methacralate (Lucite, Perspex)"], [0.001205, 0.001205, 1.19], 140.0),
    (["Al", "Cu", "Polyethylene", "Bone, Cortical (ICRP)"], [2.699, 8.96, 0.94, 1.85], 150.0),
    (["Tissue, Soft (ICRU-44)"], [1.06], 30.0),  # maxEnergy raised to 50 keV for world tables
]


@pytest.mark.parametrize("names,maxd,emax", SETS)
def test_material_tables_bit_exact(names, maxd, emax):
    a = backend.HostTables(names, maxd, emax, 1.0).arrays()
    b = oracle.Oracle(names, maxd, emax, 1.0).arrays()
    for key in ("grid", "coefficients", "majorant", "rita", "spline", "shells", "linear_index", "linear_step", "linear_energy"):
        assert np.array_equal(a[key], b[key]), key


def test_spectrum_alias_tables_bit_exact():
    for wl in (workloads.dx_thorax(histories_per_exposure=10, exposures=2),):
        t = backend.HostTables(wl.names, [0.001205, 1.0, 0.95, 1.92, 0.26], 120.0, 1.0)
        t.exposures(wl.spec)
        src = t.source_arrays()
        n = C.c_uint64()
        lib = capi.host()
        lib.dxmc_host_source_spectrum(C.byref(wl.spec), None, None, 0, C.byref(n))
        e = np.zeros(n.value)
        w = np.zeros(n.value)
        lib.dxmc_host_source_spectrum(C.byref(wl.spec), e.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)),
                                      n.value, C.byref(n))
        prob, alias = oracle.alias_table(w)
        assert np.array_equal(src["spectrum_probability"], prob)
        assert np.array_equal(src["spectrum_alias"], alias)
        assert np.array_equal(src["spectrum_energy"], e)
        # the table encodes the spectrum: P(i) = w_i / sum w
        assert np.allclose(oracle.alias_probabilities(prob, alias), w / w.sum(), rtol=1e-12, atol=1e-15)


def test_material_model_plausible():
    """Sanity bounds of the built-in physics model (xraylib is not available)."""
    out = np.zeros(5)
    dens = C.c_double()
    lib = capi.host()
    # water mu/rho at 60 keV within 15 % of the tabulated 0.206 cm2/g
    assert lib.dxmc_host_material_data(b"Water, Liquid", 60.0, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(dens)) == 0
    assert abs(out[3] - 0.206) / 0.206 < 0.15
    assert dens.value == 1.0
    # photo + Compton + Rayleigh = total; energy absorption below total
    assert np.isclose(out[0] + out[1] + out[2], out[3])
    assert 0 < out[4] < out[3]
    # bone attenuates more than water below 60 keV
    bone = np.zeros(5)
    lib.dxmc_host_material_data(b"Bone, Cortical (ICRU)", 30.0, bone.ctypes.data_as(C.POINTER(C.c_double)), C.byref(dens))
    water = np.zeros(5)
    lib.dxmc_host_material_data(b"Water, Liquid", 30.0, water.ctypes.data_as(C.POINTER(C.c_double)), C.byref(dens))
    assert bone[0] > 3 * water[0]
    assert lib.dxmc_host_material_data(b"Unobtainium", 30.0, bone.ctypes.data_as(C.POINTER(C.c_double)), C.byref(dens)) != 0


def test_exposure_tables():
    """Source geometry: counts, beam directions and collimation of the configurations."""
    t = backend.HostTables(["Water, Liquid"], [1.0], 120.0, 1.0)
    ex = t.exposures(workloads.water_pencil().spec)
    assert len(ex) == 10 and all(e.n_histories == 100000 for e in ex)
    assert list(ex[0].beam_direction) == [0.0, 0.0, 1.0] and ex[0].spectrum == -1 and ex[0].mono_energy == 60.0
    spec, nexp, per = workloads.ct_spiral_spec(1_000_000_000)
    ex = t.exposures(spec)
    assert len(ex) == nexp == 3600 and ex[0].n_histories == per == 277778
    assert ex[0].bowtie == 0 and ex[0].heel == 0 and ex[0].spectrum == 0
    # tube at SDD/2 from the isocentre, beam through the isocentre axis
    for e in (ex[0], ex[900], ex[1799]):
        p = np.array(e.position)
        d = np.array(e.beam_direction)
        assert abs(np.hypot(p[0], p[1]) - 595.0) < 1e-9
        closest = p - np.dot(p, d) * d
        assert np.hypot(closest[0], closest[1]) < 1e-6
    # z advances by collimation * pitch per rotation (720 exposures of 0.5 deg)
    assert abs((ex[720].position[2] - ex[0].position[2]) - 40.0) < 1e-9
    dx = t.exposures(workloads.dx_thorax(histories_per_exposure=10, exposures=3).spec)
    assert len(dx) == 3 and abs(dx[0].collimation_angles[1] - np.arctan(0.15)) < 1e-12
