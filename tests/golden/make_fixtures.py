"""Write tests/golden/reference_vectors.npz: input/output vectors of the
reference's own code (its headers compiled behind oracle/ref_driver.cpp into
oracle/_ref/libdxmc_ref.so by oracle/build.py, which needs the reference
checkout). The vectors pin the oracle where the reference build is absent
(tests/test_golden.py).

    python tests/golden/make_fixtures.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (run oracle/build.py with the reference checkout)")
    out = {}
    states = np.array([(0, 1), (1, 1), (42, 7), (0xDEADBEEF, 2), (2**64 - 1, 2**63 + 5), (oracle.splitmix64(42), 85)], dtype=np.uint64)
    out["pcg_states"] = states
    out["pcg32"] = np.zeros((len(states), 64), dtype=np.uint32)
    out["uniform_f64"] = np.zeros((len(states), 64))
    out["uniform_f32"] = np.zeros((len(states), 64))
    for i, (s0, s1) in enumerate(states):
        a = np.zeros(64, dtype=np.uint32)
        ref.ref_pcg32(int(s0), int(s1), 64, _p(a, C.c_uint32))
        out["pcg32"][i] = a
        for f32, key in ((0, "uniform_f64"), (1, "uniform_f32")):
            u = np.zeros(64)
            ref.ref_uniform(int(s0), int(s1), 64, f32, _p(u, C.c_double))
            out[key][i] = u
    ns = np.array([1, 2, 3, 7, 120, 1000, 2**31 + 11], dtype=np.uint64)
    out["index_n"] = ns
    out["index"] = np.zeros((len(ns), 64), dtype=np.uint64)
    for i, n in enumerate(ns):
        a = np.zeros(64, dtype=np.uint64)
        ref.ref_random_index(17, 99, int(n), 64, _p(a, C.c_uint64))
        out["index"][i] = a
    rng = np.random.default_rng(0)
    w = rng.random(141) ** 3
    w[[3, 77]] = 0.0
    p = np.zeros(len(w))
    a = np.zeros(len(w), dtype=np.uint64)
    ref.ref_alias_table(_p(w, C.c_double), len(w), _p(p, C.c_double), _p(a, C.c_uint64))
    out["alias_weights"], out["alias_prob"], out["alias_index"] = w, p, a
    pdf = oracle.PDF(lambda x: x * x)  # testrandom.cpp:151-188
    r = np.zeros(256)
    ref.ref_rita_sample(0.0, 4.0, pdf, 123, 457, 3.0, 256, _p(r, C.c_double))
    out["rita_x2_0_4_max3"] = r
    f = lambda q: 1 - 1 / (1 + (q / 0.9) ** 2) ** 2
    q = np.linspace(-1, 9, 128)
    sv = np.zeros(len(q))
    ref.ref_spline_eval(0.0, 7.3, oracle.PDF(f), len(q), _p(q, C.c_double), _p(sv, C.c_double))
    out["spline_q"], out["spline_value"] = q, sv
    rng = np.random.default_rng(5)
    rot_in = np.zeros((64, 9))
    rot_out = np.zeros((64, 6))
    for i in range(64):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        ang = rng.uniform(-7, 7)
        th, ph = rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi)
        a1, a2 = v.copy(), v.copy()
        ref.ref_rotate(_p(a1, C.c_double), _p(ax, C.c_double), ang)
        ref.ref_peturb(_p(a2, C.c_double), th, ph)
        rot_in[i] = [*v, *ax, ang, th, ph]
        rot_out[i] = [*a1, *a2]
    out["rotate_peturb_in"], out["rotate_peturb_out"] = rot_in, rot_out
    fe_in = np.zeros((3, 3, 64, 8))
    fe_out = np.zeros((3, 3, 64, 8))
    fe_e = np.zeros((3, 3, 64))
    for k in range(3):
        for j, energy in enumerate((15.0, 60.0, 140.0)):
            rng = np.random.default_rng(k)
            parts = np.zeros((64, 8))
            d = rng.normal(size=(64, 3))
            d /= np.linalg.norm(d, axis=1)[:, None]
            parts[:, 3:6] = d
            parts[:, 6] = energy
            parts[:, 7] = 1.0
            theirs = parts.copy()
            e = np.zeros(64)
            ref.ref_interact_l0(k, 64, 42, 1000, _p(theirs, C.c_double), _p(e, C.c_double))
            fe_in[k, j], fe_out[k, j], fe_e[k, j] = parts, theirs, e
    out["free_in"], out["free_out"], out["free_energy"] = fe_in, fe_out, fe_e
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_vectors.npz"), sorted(out))


if __name__ == "__main__":
    main()
