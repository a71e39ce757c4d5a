"""ORACLE - test infrastructure (only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline use it, as the checker; product code never does).

ctypes bindings of oracle/liboracle.so (the CPU restatement, oracle.cpp) and
of oracle/_ref/libdxmc_ref.so (the reference's own headers compiled behind a
small driver, ref_driver.cpp) used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdxmc_ref.so")

f64p = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int)
PDF = C.CFUNCTYPE(C.c_double, C.c_double)


class WorldView(C.Structure):
    _fields_ = [
        ("dims", C.c_uint64 * 3),
        ("spacing", C.c_double * 3),
        ("extent", C.c_double * 6),
        ("material", C.POINTER(C.c_uint8)),
        ("density", f64p),
        ("measurement", C.POINTER(C.c_uint8)),
    ]


_lib = None
_ref = None


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (python oracle/build.py)")
        _lib = C.CDLL(ORACLE_LIB)
        L = _lib
        L.oracle_rng_streams.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, u32p]
        L.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, u32p]
        L.oracle_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, f64p]
        L.oracle_random_index.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, u64p]
        L.oracle_splitmix64.argtypes = [C.c_uint64]
        L.oracle_splitmix64.restype = C.c_uint64
        L.oracle_alias_table.argtypes = [f64p, C.c_uint64, f64p, u32p]
        L.oracle_rita_build.argtypes = [C.c_double, C.c_double, PDF, C.c_int, f64p, f64p, f64p, f64p]
        L.oracle_rita_sample.argtypes = [f64p, f64p, f64p, f64p, C.c_int, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, f64p]
        L.oracle_spline_build.argtypes = [C.c_double, C.c_double, PDF, f64p]
        L.oracle_spline_eval.argtypes = [f64p, C.c_double]
        L.oracle_spline_eval.restype = C.c_double
        L.oracle_rotate.argtypes = [f64p, f64p, C.c_double]
        L.oracle_peturb.argtypes = [f64p, C.c_double, C.c_double]
        L.oracle_create.argtypes = [C.c_uint32, C.POINTER(C.c_char_p), f64p, C.c_double, C.c_double]
        L.oracle_create.restype = C.c_void_p
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_table_sizes.argtypes = [C.c_void_p, u64p, f64p]
        L.oracle_table_arrays.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p]
        L.oracle_add_spectrum.argtypes = [C.c_void_p, f64p, f64p, C.c_uint32]
        L.oracle_add_heel.argtypes = [C.c_void_p, C.c_void_p, f64p]
        L.oracle_add_bowtie.argtypes = [C.c_void_p, f64p, f64p, C.c_uint32]
        L.oracle_lut_lookup.argtypes = [C.c_void_p, C.c_int, C.c_uint64, u32p, f64p, f64p, f64p]
        L.oracle_interact.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint64, u32p, C.c_double, C.c_uint64, C.c_uint64, f64p, f64p, i32p]
        L.oracle_enter_world.argtypes = [C.POINTER(WorldView), C.c_uint64, f64p, i32p, u64p]
        L.oracle_bowtie_weight.argtypes = [C.c_void_p, C.c_int, C.c_uint64, f64p, f64p]
        L.oracle_heel_weight.argtypes = [C.c_void_p, C.c_int, C.c_uint64, f64p, f64p, f64p]
        L.oracle_compton_ia.argtypes = [f64p, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, f64p]
        L.oracle_sample_particles.argtypes = [C.c_void_p, C.POINTER(WorldView), C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, f64p, i32p]
        L.oracle_transport.argtypes = [C.c_void_p, C.POINTER(WorldView), C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                       C.c_uint32, f64p, u32p, f64p, u64p]
        L.oracle_normalize.argtypes = [C.c_uint64, f64p, f64p, C.c_int, C.c_uint64, f64p, f64p, C.c_double]
        L.oracle_tube_spectrum.argtypes = [C.POINTER(AnodeTables), C.c_double, C.c_double, f64p, C.c_uint64, i32p, f64p, C.c_int32, C.c_int32, f64p]
    return _lib


class AnodeTables(C.Structure):
    _fields_ = [("u", f64p), ("nu", C.c_int32), ("depth_f", f64p), ("depth_m", f64p), ("nx", C.c_int32), ("density_f", f64p),
                ("density_m", f64p), ("tw_voltage", f64p), ("tw_constant", f64p), ("ntw", C.c_int32), ("k_energy", f64p),
                ("k_fraction", f64p), ("nk", C.c_int32)]


def tube_spectrum(data, kvp, takeoff, energies, filters=(), normalize=True):
    """The reference's tube spectrum by its own nested-loop integral, on the
    tables in 'data' (a mapping of the fixture's arrays). filters: (Z, mm)
    applied in order."""
    keep = {k: np.ascontiguousarray(data[k], dtype=np.float64) for k in
            ("u", "depth_f", "depth_m", "density_f", "density_m", "tw_voltage", "tw_constant", "k_energy", "k_fraction")}
    t = AnodeTables(_p(keep["u"], C.c_double), len(keep["u"]), _p(keep["depth_f"], C.c_double), _p(keep["depth_m"], C.c_double),
                    len(keep["depth_f"]), _p(keep["density_f"], C.c_double), _p(keep["density_m"], C.c_double),
                    _p(keep["tw_voltage"], C.c_double), _p(keep["tw_constant"], C.c_double), len(keep["tw_voltage"]),
                    _p(keep["k_energy"], C.c_double), _p(keep["k_fraction"], C.c_double), len(keep["k_energy"]))
    e = np.ascontiguousarray(energies, dtype=np.float64)
    z = np.array([f[0] for f in filters], dtype=np.int32)
    mm = np.array([f[1] for f in filters], dtype=np.float64)
    out = np.zeros(len(e))
    lib().oracle_tube_spectrum(C.byref(t), kvp, takeoff, _p(e, C.c_double), len(e), _p(z, C.c_int32), _p(mm, C.c_double), len(filters),
                               int(normalize), _p(out, C.c_double))
    return out


def ref():
    """The reference-header driver, or None when it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        _ref = C.CDLL(REF_LIB)
        R = _ref
        R.ref_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, u32p]
        R.ref_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, f64p]
        R.ref_random_index.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, u64p]
        R.ref_alias_table.argtypes = [f64p, C.c_uint64, f64p, u64p]
        R.ref_spectrum_sample.argtypes = [f64p, f64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, f64p]
        R.ref_rita_sample.argtypes = [C.c_double, C.c_double, PDF, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, f64p]
        R.ref_spline_eval.argtypes = [C.c_double, C.c_double, PDF, C.c_uint64, f64p, f64p]
        R.ref_rotate.argtypes = [f64p, f64p, C.c_double]
        R.ref_peturb.argtypes = [f64p, C.c_double, C.c_double]
        R.ref_enter_world.argtypes = [u64p, f64p, f64p, C.c_uint64, f64p, i32p, u64p, f64p]
        R.ref_interact_l0.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, f64p, f64p]
        R.ref_sample_exposure.argtypes = [f64p, f64p, f64p, C.c_double, C.c_double, f64p, f64p, C.c_uint64, f64p, C.c_uint64, C.c_uint64,
                                          C.c_uint64, f64p, f64p]
        R.ref_xcare_weight.argtypes = [C.c_double] * 5
        R.ref_xcare_weight.restype = C.c_double
    return _ref


def rng_streams(seed, first, n, draws):
    out = np.zeros(n * draws, dtype=np.uint32)
    lib().oracle_rng_streams(seed, first, n, draws, _p(out, C.c_uint32))
    return out.reshape(n, draws)


def compton_ia(shells, energy, n, seed, first=0):
    """Impulse-approximation Compton sampler (L=2) on an explicit shell table
    of up to 12 (binding keV, electrons, J0) rows; photons start along +z.
    Returns [n][5]: returned energy, E', direction."""
    sh = np.zeros((12, 3))
    sh[:len(shells)] = shells
    sh = np.ascontiguousarray(sh, dtype=np.float64)
    out = np.zeros(n * 5)
    lib().oracle_compton_ia(_p(sh, C.c_double), energy, seed, first, n, _p(out, C.c_double))
    return out.reshape(n, 5)


def enter_world(world, particles):
    p = np.ascontiguousarray(particles, dtype=np.float64).copy()
    n = len(p)
    inside = np.zeros(n, dtype=np.int32)
    index = np.zeros(n, dtype=np.uint64)
    lib().oracle_enter_world(C.byref(world), n, _p(p, C.c_double), _p(inside, C.c_int), _p(index, C.c_uint64))
    return p, inside, index


def splitmix64(x):
    return lib().oracle_splitmix64(x)


def history_state(seed, h):
    """(state0, state1) of history h's PCG stream."""
    return splitmix64((seed ^ splitmix64(h)) & 0xFFFFFFFFFFFFFFFF), 2 * h + 1


def alias_table(weights):
    w = np.ascontiguousarray(weights, dtype=np.float64)
    prob = np.zeros(len(w))
    alias = np.zeros(len(w), dtype=np.uint32)
    lib().oracle_alias_table(_p(w, C.c_double), len(w), _p(prob, C.c_double), _p(alias, C.c_uint32))
    return prob, alias


def rita_build(lo, hi, pdf, n=56):
    cb = PDF(pdf)
    x, e, a, b = (np.zeros(n) for _ in range(4))
    lib().oracle_rita_build(lo, hi, cb, n, _p(x, C.c_double), _p(e, C.c_double), _p(a, C.c_double), _p(b, C.c_double))
    return x, e, a, b


def rita_sample(tab, state, max_value, n):
    x, e, a, b = (np.ascontiguousarray(t, dtype=np.float64) for t in tab)
    out = np.zeros(n)
    lib().oracle_rita_sample(_p(x, C.c_double), _p(e, C.c_double), _p(a, C.c_double), _p(b, C.c_double), len(x), state[0], state[1],
                             max_value, n, _p(out, C.c_double))
    return out


def spline_build(start, stop, f):
    out = np.zeros(63)
    cb = PDF(f)
    lib().oracle_spline_build(start, stop, cb, _p(out, C.c_double))
    return out


def spline_eval(c, q):
    c = np.ascontiguousarray(c, dtype=np.float64)
    return np.array([lib().oracle_spline_eval(_p(c, C.c_double), float(x)) for x in np.atleast_1d(q)])


def world_view(material, density, spacing, extent, measurement=None):
    """WorldView over (z, y, x) arrays; keeps the arrays alive."""
    mat = np.ascontiguousarray(material, dtype=np.uint8)
    dens = np.ascontiguousarray(density, dtype=np.float64)
    w = WorldView()
    nz, ny, nx = mat.shape
    for i, v in enumerate((nx, ny, nz)):
        w.dims[i] = v
    for i in range(3):
        w.spacing[i] = spacing[i]
    for i in range(6):
        w.extent[i] = extent[i]
    w.material = _p(mat, C.c_uint8)
    w.density = _p(dens, C.c_double)
    keep = [mat, dens]
    if measurement is not None:
        meas = np.ascontiguousarray(measurement, dtype=np.uint8)
        w.measurement = _p(meas, C.c_uint8)
        keep.append(meas)
    w._keep = keep
    return w


class Oracle:
    """Oracle tables for a material list and the transport over them."""

    def __init__(self, names, max_density, max_energy, min_energy=1.0):
        self._names = (C.c_char_p * len(names))(*[n.encode() for n in names])
        md = np.ascontiguousarray(max_density, dtype=np.float64)
        self.h = lib().oracle_create(len(names), self._names, _p(md, C.c_double), max_energy, min_energy)
        if not self.h:
            raise ValueError("oracle: unknown material")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def arrays(self):
        sizes = np.zeros(3, dtype=np.uint64)
        lin = np.zeros(2)
        lib().oracle_table_sizes(self.h, _p(sizes, C.c_uint64), _p(lin, C.c_double))
        nm, nn = int(sizes[0]), int(sizes[1])
        grid, coef, maj = np.zeros(nn), np.zeros(nm * nn * 6), np.zeros(nn * 2)
        rita, spline, shells = np.zeros(nm * 224), np.zeros(nm * 63), np.zeros(nm * 12 * 11)
        lib().oracle_table_arrays(self.h, *(_p(a, C.c_double) for a in (grid, coef, maj, rita, spline, shells)))
        return {"grid": grid, "coefficients": coef, "majorant": maj, "rita": rita, "spline": spline,
                "shells": shells.reshape(nm * 12, 11), "linear_index": int(sizes[2]), "linear_step": lin[0], "linear_energy": lin[1]}

    def add_spectrum(self, weights, energies):
        w = np.ascontiguousarray(weights, dtype=np.float64)
        e = np.ascontiguousarray(energies, dtype=np.float64)
        return lib().oracle_add_spectrum(self.h, _p(w, C.c_double), _p(e, C.c_double), len(w))

    def add_heel(self, heel_struct, weights):
        w = np.ascontiguousarray(weights, dtype=np.float64)
        return lib().oracle_add_heel(self.h, C.byref(heel_struct), _p(w, C.c_double))

    def add_bowtie(self, angles, weights):
        a = np.ascontiguousarray(angles, dtype=np.float64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        return lib().oracle_add_bowtie(self.h, _p(a, C.c_double), _p(w, C.c_double), len(a))

    def bowtie_weight(self, table, angles):
        q = np.ascontiguousarray(angles, dtype=np.float64)
        out = np.zeros(len(q))
        lib().oracle_bowtie_weight(self.h, table, len(q), _p(q, C.c_double), _p(out, C.c_double))
        return out

    def heel_weight(self, table, angles, energies):
        a = np.ascontiguousarray(angles, dtype=np.float64)
        e = np.ascontiguousarray(energies, dtype=np.float64)
        out = np.zeros(len(a))
        lib().oracle_heel_weight(self.h, table, len(a), _p(a, C.c_double), _p(e, C.c_double), _p(out, C.c_double))
        return out

    def add_source_tables(self, src, spectra):
        """Register source tables in the order of a product HostTables.source_arrays():
        spectra as the raw (weights, energies) the alias tables are built from,
        heel grids and bowtie pairs as the product holds them."""
        for weights, energies in spectra:
            self.add_spectrum(weights, energies)
        for h in src.get("heel", []):
            n = h.energy_size * h.angle_size
            self.add_heel(h, src["heel_weight"][h.weight_offset:h.weight_offset + n])
        off = src.get("bowtie_offset")
        if off is not None:
            for i in range(len(off) - 1):
                a, b = off[i], off[i + 1]
                self.add_bowtie(src["bowtie_angle"][a:b], src["bowtie_weight"][a:b])

    def lut_lookup(self, f32, material, energy):
        m = np.ascontiguousarray(material, dtype=np.uint32)
        e = np.ascontiguousarray(energy, dtype=np.float64)
        att, maj = np.zeros(len(e) * 3), np.zeros(len(e))
        lib().oracle_lut_lookup(self.h, f32, len(e), _p(m, C.c_uint32), _p(e, C.c_double), _p(att, C.c_double), _p(maj, C.c_double))
        return att.reshape(-1, 3), maj

    def interact(self, L, kind, material, particles, seed, first=0, forced_probability=0.0):
        m = np.ascontiguousarray(material, dtype=np.uint32)
        p = np.ascontiguousarray(particles, dtype=np.float64).copy()
        e = np.zeros(len(m))
        alive = np.zeros(len(m), dtype=np.int32)
        lib().oracle_interact(self.h, L, kind, len(m), _p(m, C.c_uint32), forced_probability, seed, first, _p(p, C.c_double),
                              _p(e, C.c_double), _p(alive, C.c_int))
        return p, e, alive

    def sample_particles(self, world, exposure, seed, first, n):
        out = np.zeros(n * 8)
        inside = np.zeros(n, dtype=np.int32)
        lib().oracle_sample_particles(self.h, C.byref(world), C.byref(exposure), seed, first, n, _p(out, C.c_double), _p(inside, C.c_int))
        return out.reshape(n, 8), inside

    def transport(self, world, exposures, L=1, f32=False, seed=42, first=0, threads=None, dose=None, events=None, variance=None):
        n = world.dims[0] * world.dims[1] * world.dims[2]
        dose = np.zeros(n) if dose is None else dose
        events = np.zeros(n, dtype=np.uint32) if events is None else events
        variance = np.zeros(n) if variance is None else variance
        counters = np.zeros(3, dtype=np.uint64)
        threads = threads or os.cpu_count()
        lib().oracle_transport(self.h, C.byref(world), exposures, len(exposures), L, 1 if f32 else 0, seed, first, threads, _p(dose, C.c_double),
                               _p(events, C.c_uint32), _p(variance, C.c_double), _p(counters, C.c_uint64))
        return {"dose": dose, "n_events": events, "variance": variance, "histories": int(counters[0]), "steps": int(counters[1]),
                "events": int(counters[2])}


def alias_probabilities(prob, alias):
    """Bin probabilities encoded by an alias table: (prob_i + sum_{alias_k = i} (1 - prob_k)) / n."""
    n = len(prob)
    p = np.array(prob, dtype=np.float64)
    for k in range(n):
        p[alias[k]] += 1.0 - prob[k]
    return p / n


def normalize(dose, variance, mode, n_histories, density=None, spacing=None, calibration=1.0):
    d = np.ascontiguousarray(dose, dtype=np.float64).copy()
    v = np.ascontiguousarray(variance, dtype=np.float64).copy()
    dens = np.ascontiguousarray(density if density is not None else np.zeros(len(d)), dtype=np.float64)
    sp = np.ascontiguousarray(spacing if spacing is not None else np.ones(3), dtype=np.float64)
    lib().oracle_normalize(len(d), _p(d, C.c_double), _p(v, C.c_double), mode, n_histories, _p(dens, C.c_double), _p(sp, C.c_double), calibration)
    return d, v
