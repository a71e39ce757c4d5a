"""numpy-level wrappers of the C ABIs.

Device       a HIP context (dxmc_hip_ctx_*) with resident worlds, tables
             and tallies, the transport run and the parity entry points
HostTables   host-built material + source tables (dxmc_host_tables_*)
transport()  the full Transport<T>::operator() through dxmc_host_transport

All compute goes through the native libraries; nothing here computes
transport physics.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi


class DxmcError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"dxmc status {status}: {message}")
        self.status = status


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _check_host(status):
    if status != capi.DXMC_OK:
        raise DxmcError(status, capi.host().dxmc_host_last_error().decode())


def source_spec(type, **kw):
    """Build a SourceSpec; array fields (bowtie, spectrum) are kept alive on the object."""
    s = capi.SourceSpec()
    s.type = type
    s.heel_effect = 1
    keep = []
    for k, v in kw.items():
        if k in ("bowtie_angles", "bowtie_weights", "spectrum_weights", "spectrum_energies", "bowtie_b_angles", "bowtie_b_weights"):
            a = np.ascontiguousarray(v, dtype=np.float64)
            keep.append(a)
            setattr(s, k, _ptr(a, C.c_double))
        elif k in ("position", "direction_cosines", "field_size", "source_angles_deg", "collimation_angles"):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    if "bowtie_angles" in kw:
        s.bowtie_n = len(kw["bowtie_angles"])
    if "spectrum_weights" in kw:
        s.spectrum_n = len(kw["spectrum_weights"])
    if "bowtie_b_angles" in kw:
        s.bowtie_b_n = len(kw["bowtie_b_angles"])
    s._keep = keep
    return s


class HostTables:
    """Material tables (+ the source tables appended by exposures())."""

    def __init__(self, names, max_density, max_energy, min_energy=1.0):
        lib = capi.host()
        self._names = (C.c_char_p * len(names))(*[n.encode() for n in names])
        md = np.ascontiguousarray(max_density, dtype=np.float64)
        h = C.c_void_p()
        _check_host(lib.dxmc_host_tables_create(len(names), self._names, _ptr(md, C.c_double), max_energy, min_energy, C.byref(h)))
        self.handle = h

    def exposures(self, spec, world_cosines=None):
        lib = capi.host()
        wc = np.zeros(6) if world_cosines is None else np.ascontiguousarray(world_cosines, dtype=np.float64)
        n = C.c_uint64()
        _check_host(lib.dxmc_host_source_exposures(C.byref(spec), _ptr(wc, C.c_double), self.handle, None, 0, C.byref(n)))
        out = (capi.Exposure * n.value)()
        _check_host(lib.dxmc_host_source_exposures(C.byref(spec), _ptr(wc, C.c_double), self.handle, out, n.value, C.byref(n)))
        return out

    def desc(self):
        d = capi.TablesDesc()
        _check_host(capi.host().dxmc_host_tables_desc(self.handle, C.byref(d)))
        return d

    def arrays(self):
        """Copies of the material tables as numpy arrays."""
        d = self.desc()
        nm, nn = d.n_materials, d.lut_nodes

        def get(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy()

        shells = np.ctypeslib.as_array(C.cast(d.shells, C.POINTER(C.c_double)), shape=(nm * 12 * 11,)).copy().reshape(nm * 12, 11)
        return {
            "grid": get(d.lut_grid, nn), "coefficients": get(d.lut_coefficients, nm * nn * 6), "majorant": get(d.lut_majorant, nn * 2),
            "rita": get(d.rita, nm * 4 * 56), "spline": get(d.spline, nm * 63), "shells": shells,
            "linear_index": d.lut_linear_index, "linear_step": d.lut_linear_step, "linear_energy": d.lut_linear_energy,
            "n_spectra": d.n_spectra, "n_heel": d.n_heel, "n_bowtie": d.n_bowtie,
        }

    def source_arrays(self):
        d = self.desc()
        out = {}
        if d.n_spectra:
            off = np.ctypeslib.as_array(d.spectrum_offset, shape=(d.n_spectra + 1,)).copy()
            n = off[-1]
            out["spectrum_offset"] = off
            out["spectrum_probability"] = np.ctypeslib.as_array(d.spectrum_probability, shape=(n,)).copy()
            out["spectrum_alias"] = np.ctypeslib.as_array(d.spectrum_alias, shape=(n,)).copy()
            out["spectrum_energy"] = np.ctypeslib.as_array(d.spectrum_energy, shape=(n,)).copy()
        if d.n_heel:
            heels = [d.heel[i] for i in range(d.n_heel)]
            nw = max(h.weight_offset + h.energy_size * h.angle_size for h in heels)
            out["heel"] = heels
            out["heel_weight"] = np.ctypeslib.as_array(d.heel_weight, shape=(nw,)).copy()
        if d.n_bowtie:
            off = np.ctypeslib.as_array(d.bowtie_offset, shape=(d.n_bowtie + 1,)).copy()
            out["bowtie_offset"] = off
            out["bowtie_angle"] = np.ctypeslib.as_array(d.bowtie_angle, shape=(off[-1],)).copy()
            out["bowtie_weight"] = np.ctypeslib.as_array(d.bowtie_weight, shape=(off[-1],)).copy()
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            capi.host().dxmc_host_tables_destroy(h)
            self.handle = None


def world_extent(dims, spacing, origin=(0, 0, 0)):
    """World::matrixExtentSafe of a grid."""
    w = capi.HostWorld()
    for i in range(3):
        w.dims[i] = int(dims[i])
        w.spacing[i] = float(spacing[i])
        w.origin[i] = float(origin[i])
    ext = np.zeros(6)
    _check_host(capi.host().dxmc_host_world_extent(C.byref(w), _ptr(ext, C.c_double)))
    return ext


@dataclass
class RunResult:
    dose: np.ndarray
    n_events: np.ndarray
    variance: np.ndarray
    stats: capi.RunStats


class Device:
    """A HIP context on one device."""

    def __init__(self, device=0):
        self.lib = capi.hip()
        h = C.c_void_p()
        st = self.lib.dxmc_hip_ctx_create(device, C.byref(h))
        if st != capi.DXMC_OK:
            raise DxmcError(st, self.lib.dxmc_hip_last_error(None).decode())
        self.ctx = h

    def _check(self, st):
        if st != capi.DXMC_OK:
            raise DxmcError(st, self.lib.dxmc_hip_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.dxmc_hip_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def mem_info(self):
        f, t = C.c_uint64(), C.c_uint64()
        self._check(self.lib.dxmc_hip_mem_info(self.ctx, C.byref(f), C.byref(t)))
        return f.value, t.value

    # ------------------------------------------------------------ objects
    def upload_world(self, material, density, spacing, extent_safe, n_materials, measurement=None):
        mat = np.ascontiguousarray(material, dtype=np.uint8)
        d = capi.WorldDesc()
        nz, ny, nx = mat.shape
        for i, v in enumerate((nx, ny, nz)):
            d.dims[i] = v
        for i in range(3):
            d.spacing[i] = spacing[i]
        for i in range(6):
            d.extent_safe[i] = extent_safe[i]
        d.material = _ptr(mat, C.c_uint8)
        keep = [mat]
        if density.dtype == np.float64:
            dens = np.ascontiguousarray(density)
            d.density_f64 = _ptr(dens, C.c_double)
        else:
            dens = np.ascontiguousarray(density, dtype=np.float32)
            d.density_f32 = _ptr(dens, C.c_float)
        keep.append(dens)
        if measurement is not None:
            meas = np.ascontiguousarray(measurement, dtype=np.uint8)
            d.measurement = _ptr(meas, C.c_uint8)
            keep.append(meas)
        d.n_materials = n_materials
        h = C.c_void_p()
        self._check(self.lib.dxmc_hip_world_upload(self.ctx, C.byref(d), C.byref(h)))
        return h

    def world_bytes(self, h):
        b = C.c_uint64()
        self._check(self.lib.dxmc_hip_world_bytes(h, C.byref(b)))
        return b.value

    def world_max_density(self, h, n_materials):
        """Per-material maximum density, from the device scan at upload."""
        out = np.zeros(n_materials, dtype=np.float64)
        self._check(self.lib.dxmc_hip_world_max_density(h, _ptr(out, C.c_double), n_materials))
        return out

    def destroy_world(self, h):
        self._check(self.lib.dxmc_hip_world_destroy(self.ctx, h))

    def upload_tables(self, desc):
        h = C.c_void_p()
        self._check(self.lib.dxmc_hip_tables_upload(self.ctx, C.byref(desc), C.byref(h)))
        return h

    def destroy_tables(self, h):
        self._check(self.lib.dxmc_hip_tables_destroy(self.ctx, h))

    def tally(self, n):
        h = C.c_void_p()
        self._check(self.lib.dxmc_hip_tally_create(self.ctx, n, C.byref(h)))
        return h

    def destroy_tally(self, h):
        self._check(self.lib.dxmc_hip_tally_destroy(self.ctx, h))

    def reset_tally(self, h):
        self._check(self.lib.dxmc_hip_tally_reset(self.ctx, h))

    def download(self, tally, n):
        dose = np.zeros(n)
        ev = np.zeros(n, dtype=np.uint32)
        var = np.zeros(n)
        self._check(self.lib.dxmc_hip_tally_download(self.ctx, tally, _ptr(dose, C.c_double), _ptr(ev, C.c_uint32), _ptr(var, C.c_double)))
        return dose, ev, var

    def normalize(self, tally, world, mode, n_histories, calibration=1.0):
        self._check(self.lib.dxmc_hip_tally_normalize(self.ctx, tally, world, mode, n_histories, calibration))

    def run(self, world, tables, exposures, tally, precision, low_energy_correction=1, seed=42, first_history=0, tuning=None,
            cancel=None, progress=None):
        cfg = tuning if tuning is not None else capi.RunCfg()
        cfg.precision = precision
        cfg.low_energy_correction = low_energy_correction
        cfg.seed = seed
        cfg.first_history = first_history
        stats = capi.RunStats()
        self._check(self.lib.dxmc_hip_run(self.ctx, world, tables, exposures, len(exposures), C.byref(cfg), tally,
                                          cancel, progress, C.byref(stats)))
        return stats

    def run_noalloc(self, world, tables, exposures, tally, precision, low_energy_correction=1, seed=42, first_history=0, tuning=None):
        """run() through dxmc_hip_run_noalloc (no allocation during the call;
        at most capi.NOALLOC_MAX_EXPOSURES exposures)."""
        cfg = tuning if tuning is not None else capi.RunCfg()
        cfg.precision = precision
        cfg.low_energy_correction = low_energy_correction
        cfg.seed = seed
        cfg.first_history = first_history
        stats = capi.RunStats()
        self._check(self.lib.dxmc_hip_run_noalloc(self.ctx, world, tables, exposures, len(exposures), C.byref(cfg), tally, C.byref(stats)))
        return stats

    def run_monitored(self, world, tables, exposures, tally, precision, low_energy_correction=1, seed=42, tuning=None, preview=None,
                      preview_axis=2, preview_interval=0.0, cancel=None, progress=None):
        """run() with a live preview: preview(image) receives the projection
        of the tally along preview_axis (a 2-D float32 array, rows along the
        second remaining axis) between launches."""
        cfg = tuning if tuning is not None else capi.RunCfg()
        cfg.precision = precision
        cfg.low_energy_correction = low_energy_correction
        cfg.seed = seed

        def on_preview(_user, image, width, height):
            preview(np.ctypeslib.as_array(image, shape=(height, width)).copy())

        fn = capi.PreviewFn(on_preview) if preview is not None else capi.PreviewFn()
        mon = capi.RunMonitor(cancel, progress, fn, None, preview_axis, preview_interval)
        stats = capi.RunStats()
        self._check(self.lib.dxmc_hip_run_monitored(self.ctx, world, tables, exposures, len(exposures), C.byref(cfg), tally, C.byref(mon),
                                                    C.byref(stats)))
        return stats

    def tally_mip(self, tally, world, dims, axis):
        """Maximum of the tally's dose along axis (0 x, 1 y, 2 z); dims (nx, ny, nz)."""
        other = [d for i, d in enumerate(dims) if i != axis]
        out = np.zeros(other[0] * other[1], dtype=np.float32)
        self._check(self.lib.dxmc_hip_tally_mip(self.ctx, tally, world, axis, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out.reshape(other[1], other[0])

    def autotune(self, world, tables, exposures, precision, low_energy_correction=1, trial_histories=1 << 23, base=None):
        """dxmc_hip_autotune: tuned run configuration and its report."""
        cfg = base if base is not None else capi.RunCfg()
        cfg.precision = precision
        cfg.low_energy_correction = low_energy_correction
        best = capi.RunCfg()
        rep = capi.TuneReport()
        self._check(self.lib.dxmc_hip_autotune(self.ctx, world, tables, exposures, len(exposures), C.byref(cfg), trial_histories, C.byref(best),
                                               C.byref(rep)))
        return best, rep

    def synchronize(self):
        self._check(self.lib.dxmc_hip_synchronize(self.ctx))

    # ------------------------------------------------------------ parity
    def rng_streams(self, seed, first, n, draws):
        out = np.zeros(n * draws, dtype=np.uint32)
        self._check(self.lib.dxmc_hip_rng_streams(self.ctx, seed, first, n, draws, _ptr(out, C.c_uint32)))
        return out.reshape(n, draws)

    def lut_lookup(self, tables, precision, material, energy):
        m = np.ascontiguousarray(material, dtype=np.uint32)
        e = np.ascontiguousarray(energy, dtype=np.float64)
        att = np.zeros(len(e) * 3)
        maj = np.zeros(len(e))
        self._check(self.lib.dxmc_hip_lut_lookup(self.ctx, tables, precision, len(e), _ptr(m, C.c_uint32), _ptr(e, C.c_double),
                                                 _ptr(att, C.c_double), _ptr(maj, C.c_double)))
        return att.reshape(-1, 3), maj

    def interact(self, tables, precision, L, kind, material, particles, seed, first=0, forced_probability=0.0):
        m = np.ascontiguousarray(material, dtype=np.uint32)
        p = np.ascontiguousarray(particles, dtype=np.float64).copy()
        e = np.zeros(len(m))
        alive = np.zeros(len(m), dtype=np.int32)
        self._check(self.lib.dxmc_hip_interact(self.ctx, tables, precision, L, kind, len(m), _ptr(m, C.c_uint32), forced_probability, seed,
                                               first, _ptr(p, C.c_double), _ptr(e, C.c_double), _ptr(alive, C.c_int32)))
        return p, e, alive

    def sample_particles(self, world, tables, precision, exposure, seed, first, n):
        out = np.zeros(n * 8)
        inside = np.zeros(n, dtype=np.int32)
        self._check(self.lib.dxmc_hip_sample_particles(self.ctx, world, tables, precision, C.byref(exposure), seed, first, n,
                                                       _ptr(out, C.c_double), _ptr(inside, C.c_int32)))
        return out.reshape(n, 8), inside


def host_world(material, density, spacing, names, origin=(0, 0, 0), measurement=None, cosines=None):
    """HostWorld over numpy arrays (z, y, x order); keeps references alive."""
    w = capi.HostWorld()
    mat = np.ascontiguousarray(material, dtype=np.uint8)
    dens = np.ascontiguousarray(density, dtype=np.float32)
    nz, ny, nx = mat.shape
    for i, v in enumerate((nx, ny, nz)):
        w.dims[i] = v
    for i in range(3):
        w.spacing[i] = spacing[i]
        w.origin[i] = origin[i]
    if cosines is not None:
        for i in range(6):
            w.direction_cosines[i] = cosines[i]
    w.material = _ptr(mat, C.c_uint8)
    w.density = _ptr(dens, C.c_float)
    keep = [mat, dens]
    if measurement is not None:
        meas = np.ascontiguousarray(measurement, dtype=np.uint8)
        w.measurement = _ptr(meas, C.c_uint8)
        keep.append(meas)
    names_arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    w.material_names = names_arr
    w.n_materials = len(names)
    w._keep = (keep, names_arr)
    return w


def transport(world, spec, precision=capi.PRECISION_FP64, low_energy_correction=1, output_mode=capi.OUTPUT_DOSE,
              use_source_calibration=False, seed=42, device=0, tuning=None):
    """Transport<T>::operator() on the GPU (dxmc_host_transport)."""
    cfg = capi.HostTransportCfg()
    cfg.device = device
    cfg.precision = precision
    cfg.low_energy_correction = low_energy_correction
    cfg.output_mode = output_mode
    cfg.use_source_calibration = 1 if use_source_calibration else 0
    cfg.seed = seed
    if tuning is not None:
        cfg.tuning = tuning
    n = world.dims[0] * world.dims[1] * world.dims[2]
    dose = np.zeros(n)
    ev = np.zeros(n, dtype=np.uint32)
    var = np.zeros(n)
    stats = capi.RunStats()
    secs = C.c_double()
    nh = C.c_uint64()
    _check_host(capi.host().dxmc_host_transport(C.byref(world), C.byref(spec), C.byref(cfg), _ptr(dose, C.c_double), _ptr(ev, C.c_uint32),
                                                _ptr(var, C.c_double), C.byref(stats), C.byref(secs), C.byref(nh)))
    return {"dose": dose, "n_events": ev, "variance": var, "stats": stats, "seconds": secs.value, "histories": nh.value}


class Server:
    """Request server over one GPU (dxmc_host_server_*): requests are
    prepared on host threads, admitted to device slots by memory and run
    concurrently; results and per-request timing are collected by id."""

    def __init__(self, device=0, precision=capi.PRECISION_FP32, low_energy_correction=1, slots=4, prep_threads=4, memory_budget=0,
                 tuning=None):
        cfg = capi.ServerCfg()
        cfg.device = device
        cfg.precision = precision
        cfg.low_energy_correction = low_energy_correction
        cfg.slots = slots
        cfg.prep_threads = prep_threads
        cfg.memory_budget = memory_budget
        if tuning is not None:
            cfg.tuning = tuning
        self.lib = capi.host()
        h = C.c_void_p()
        _check_host(self.lib.dxmc_host_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            self.lib.dxmc_host_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def clock(self):
        return self.lib.dxmc_host_server_clock(self.h)

    def pause(self, paused=True):
        _check_host(self.lib.dxmc_host_server_pause(self.h, 1 if paused else 0))

    def submit(self, world, spec, seed=42, output_mode=capi.OUTPUT_EV_PER_HISTORY):
        """world: a HostWorld (host_world()); spec: a SourceSpec."""
        rid = C.c_uint64()
        _check_host(self.lib.dxmc_host_server_submit(self.h, C.byref(world), C.byref(spec), seed, output_mode, C.byref(rid)))
        self.sizes[rid.value] = world.dims[0] * world.dims[1] * world.dims[2]
        return rid.value

    def wait(self, rid, timeout=-1.0):
        st = C.c_int32()
        _check_host(self.lib.dxmc_host_server_wait(self.h, rid, timeout, C.byref(st)))
        return st.value

    def collect(self, rid):
        n = self.sizes.pop(rid)
        dose = np.zeros(n)
        ev = np.zeros(n, dtype=np.uint32)
        var = np.zeros(n)
        info = capi.RequestInfo()
        _check_host(self.lib.dxmc_host_server_collect(self.h, rid, _ptr(dose, C.c_double), _ptr(ev, C.c_uint32), _ptr(var, C.c_double),
                                                      C.byref(info)))
        return {"dose": dose, "n_events": ev, "variance": var, "info": info}
