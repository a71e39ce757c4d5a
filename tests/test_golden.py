"""The oracle against golden vectors of the reference's own code
(tests/golden/reference_vectors.npz, written by tests/golden/make_fixtures.py
from the reference headers compiled into oracle/_ref): PCG32 and uniforms
(dxmcrandom.hpp:37-167), integer draws (:125-147), the alias table
(:229-274), RITA sampling (:483-500), the cubic spline
(interpolation.hpp:123-176), rotate and peturb (vectormath.hpp:83-99,
197-221) and the L=0 interaction samplers (transport.hpp:216-340). Bit-exact.
These run without the reference checkout."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz"))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def test_pcg32_and_uniforms():
    for i, (s0, s1) in enumerate(G["pcg_states"]):
        a = np.zeros(64, dtype=np.uint32)
        oracle.lib().oracle_pcg32(int(s0), int(s1), 64, _p(a, C.c_uint32))
        assert np.array_equal(a, G["pcg32"][i])
        for f32, key in ((0, "uniform_f64"), (1, "uniform_f32")):
            u = np.zeros(64)
            oracle.lib().oracle_uniform(int(s0), int(s1), 64, f32, _p(u, C.c_double))
            assert np.array_equal(u, G[key][i])


def test_random_index():
    for i, n in enumerate(G["index_n"]):
        a = np.zeros(64, dtype=np.uint64)
        oracle.lib().oracle_random_index(17, 99, int(n), 64, _p(a, C.c_uint64))
        assert np.array_equal(a, G["index"][i])


def test_alias_table():
    p, a = oracle.alias_table(G["alias_weights"])
    assert np.array_equal(p, G["alias_prob"])
    assert np.array_equal(a.astype(np.uint64), G["alias_index"])


def test_rita_sampling():
    tab = oracle.rita_build(0.0, 4.0, lambda x: x * x)
    assert np.array_equal(oracle.rita_sample(tab, (123, 457), 3.0, 256), G["rita_x2_0_4_max3"])


def test_spline():
    c = oracle.spline_build(0.0, 7.3, lambda q: 1 - 1 / (1 + (q / 0.9) ** 2) ** 2)
    assert np.array_equal(oracle.spline_eval(c, G["spline_q"]), G["spline_value"])


def test_rotate_and_peturb():
    for row_in, row_out in zip(G["rotate_peturb_in"], G["rotate_peturb_out"]):
        v, ax, ang, th, ph = row_in[:3].copy(), row_in[3:6].copy(), row_in[6], row_in[7], row_in[8]
        a, b = v.copy(), v.copy()
        oracle.lib().oracle_rotate(_p(a, C.c_double), _p(ax, C.c_double), ang)
        oracle.lib().oracle_peturb(_p(b, C.c_double), th, ph)
        assert np.array_equal(a, row_out[:3])
        assert np.array_equal(b, row_out[3:])


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_free_electron_interactions(kind):
    orc = oracle.Oracle(["Water, Liquid"], [1.0], 150.0)
    for j in range(3):
        ours, e, _ = orc.interact(0, kind, np.zeros(64), G["free_in"][kind, j], 42, 1000)
        assert np.array_equal(ours, G["free_out"][kind, j])
        assert np.array_equal(e, G["free_energy"][kind, j])
