"""ctypes bindings of the two C ABIs (include/dxmc_hip.h, include/dxmc_host.h).

The structures mirror the C declarations field by field. Loading fails
loudly when a library has not been built: there is no Python or CPU
fallback for the transport.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libdxmc_hip.so")
HOST_LIB = os.path.join(HERE, "libdxmc_host.so")

DXMC_OK = 0
DXMC_ERR_ARGUMENT = 1
DXMC_ERR_DEVICE = 2
DXMC_ERR_MEMORY = 3
DXMC_ERR_CANCELLED = 4
DXMC_ERR_ADMISSION = 5
DXMC_ERR_STATE = 6

PRECISION_FP64 = 0
PRECISION_FP32 = 1
OUTPUT_EV_PER_HISTORY = 0
OUTPUT_DOSE = 1
RUN_COUNTERS = 1
RUN_AUTOTUNE = 2
RUN_EXACT_STEPS = 4

NOALLOC_MAX_EXPOSURES = 8192  # DXMC_NOALLOC_MAX_EXPOSURES

SOURCE_PENCIL = 0
SOURCE_ISOTROPIC = 1
SOURCE_ISOTROPIC_CT = 2
SOURCE_DX = 3
SOURCE_CBCT = 4
SOURCE_CT_AXIAL = 5
SOURCE_CT_SPIRAL = 6
SOURCE_CT_AXIAL_DUAL = 7
SOURCE_CT_SPIRAL_DUAL = 8
SOURCE_CT_TOPOGRAM = 9

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class WorldDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_uint64 * 3),
        ("spacing", C.c_double * 3),
        ("extent_safe", C.c_double * 6),
        ("material", u8p),
        ("density_f32", f32p),
        ("density_f64", f64p),
        ("measurement", u8p),
        ("n_materials", C.c_uint32),
    ]


class Shell(C.Structure):
    _fields_ = [
        ("binding_energy", C.c_double),
        ("number_electrons", C.c_double),
        ("hartree_fock_j0", C.c_double),
        ("photo_ionization_probability", C.c_double),
        ("fluorescence_yield", C.c_double),
        ("line_probability", C.c_double * 3),
        ("line_energy", C.c_double * 3),
    ]


class Heel(C.Structure):
    _fields_ = [
        ("energy_start", C.c_double),
        ("energy_step", C.c_double),
        ("angle_start", C.c_double),
        ("angle_step", C.c_double),
        ("energy_size", C.c_uint32),
        ("angle_size", C.c_uint32),
        ("weight_offset", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class TablesDesc(C.Structure):
    _fields_ = [
        ("n_materials", C.c_uint32),
        ("lut_nodes", C.c_uint32),
        ("lut_grid", f64p),
        ("lut_coefficients", f64p),
        ("lut_majorant", f64p),
        ("lut_linear_index", C.c_uint32),
        ("lut_linear_step", C.c_double),
        ("lut_linear_energy", C.c_double),
        ("rita", f64p),
        ("spline", f64p),
        ("shells", C.POINTER(Shell)),
        ("n_spectra", C.c_uint32),
        ("spectrum_offset", u32p),
        ("spectrum_probability", f64p),
        ("spectrum_alias", u32p),
        ("spectrum_energy", f64p),
        ("n_heel", C.c_uint32),
        ("heel", C.POINTER(Heel)),
        ("heel_weight", f64p),
        ("n_bowtie", C.c_uint32),
        ("bowtie_offset", u32p),
        ("bowtie_angle", f64p),
        ("bowtie_weight", f64p),
    ]


class Exposure(C.Structure):
    _fields_ = [
        ("position", C.c_double * 3),
        ("direction_cosines", C.c_double * 6),
        ("beam_direction", C.c_double * 3),
        ("collimation_angles", C.c_double * 4),
        ("weight", C.c_double),
        ("mono_energy", C.c_double),
        ("n_histories", C.c_uint64),
        ("spectrum", C.c_int32),
        ("heel", C.c_int32),
        ("bowtie", C.c_int32),
        ("reserved", C.c_int32),
    ]


class RunCfg(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("low_energy_correction", C.c_int32),
        ("seed", C.c_uint64),
        ("first_history", C.c_uint64),
        ("histories_per_launch", C.c_uint64),
        ("block_size", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("regeneration_threshold", C.c_uint32),
        ("flags", C.c_uint32),
        ("interaction_threshold", C.c_uint32),
        ("queue_depth", C.c_uint32),
        ("waves_per_simd", C.c_uint32),
        ("lds_tables", C.c_uint32),
        ("block_bounds", C.c_uint32),
        ("reserved0", C.c_uint32),
    ]


class TuneReport(C.Structure):
    _fields_ = [
        ("trials", C.c_uint32),
        ("reserved", C.c_uint32),
        ("default_rate", C.c_double),
        ("best_rate", C.c_double),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("histories", C.c_uint64),
        ("steps", C.c_uint64),
        ("events", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("launches", C.c_uint64),
    ]


PreviewFn = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_float), C.c_uint64, C.c_uint64)


class RunMonitor(C.Structure):
    _fields_ = [
        ("cancel", C.POINTER(C.c_int32)),
        ("exposures_done", C.POINTER(C.c_uint64)),
        ("preview", PreviewFn),
        ("preview_user", C.c_void_p),
        ("preview_axis", C.c_int32),
        ("preview_interval_seconds", C.c_float),
    ]


class SourceSpec(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("heel_effect", C.c_int32),
        ("position", C.c_double * 3),
        ("direction_cosines", C.c_double * 6),
        ("histories_per_exposure", C.c_uint64),
        ("total_exposures", C.c_uint64),
        ("photon_energy", C.c_double),
        ("tube_voltage", C.c_double),
        ("anode_angle_deg", C.c_double),
        ("energy_resolution", C.c_double),
        ("al_mm", C.c_double),
        ("cu_mm", C.c_double),
        ("sn_mm", C.c_double),
        ("sdd", C.c_double),
        ("field_size", C.c_double * 2),
        ("source_angles_deg", C.c_double * 2),
        ("dap", C.c_double),
        ("span_angle_deg", C.c_double),
        ("step_angle_deg", C.c_double),
        ("collimation", C.c_double),
        ("fov", C.c_double),
        ("start_angle_deg", C.c_double),
        ("scan_length", C.c_double),
        ("exposure_angle_step_deg", C.c_double),
        ("pitch", C.c_double),
        ("gantry_tilt_deg", C.c_double),
        ("ctdi_vol", C.c_double),
        ("axial_step", C.c_double),
        ("bowtie_n", C.c_uint32),
        ("bowtie_angles", f64p),
        ("bowtie_weights", f64p),
        ("collimation_angles", C.c_double * 4),
        ("spectrum_n", C.c_uint32),
        ("spectrum_weights", f64p),
        ("spectrum_energies", f64p),
        ("air_dose", C.c_double),
        ("xcare", C.c_int32),
        ("start_angle_b_set", C.c_int32),
        ("xcare_filter_angle_deg", C.c_double),
        ("xcare_span_angle_deg", C.c_double),
        ("xcare_ramp_angle_deg", C.c_double),
        ("xcare_low_weight", C.c_double),
        ("tube_b_voltage", C.c_double),
        ("tube_b_anode_angle_deg", C.c_double),
        ("tube_b_al_mm", C.c_double),
        ("tube_b_cu_mm", C.c_double),
        ("tube_b_sn_mm", C.c_double),
        ("sdd_b", C.c_double),
        ("fov_b", C.c_double),
        ("start_angle_b_deg", C.c_double),
        ("mas_a", C.c_double),
        ("mas_b", C.c_double),
        ("bowtie_b_n", C.c_uint32),
        ("bowtie_b_angles", f64p),
        ("bowtie_b_weights", f64p),
    ]


class HostWorld(C.Structure):
    _fields_ = [
        ("dims", C.c_uint64 * 3),
        ("spacing", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("direction_cosines", C.c_double * 6),
        ("material", u8p),
        ("density", f32p),
        ("measurement", u8p),
        ("n_materials", C.c_uint32),
        ("material_names", C.POINTER(C.c_char_p)),
    ]


class HostTransportCfg(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("precision", C.c_int32),
        ("low_energy_correction", C.c_int32),
        ("output_mode", C.c_int32),
        ("use_source_calibration", C.c_int32),
        ("reserved", C.c_int32),
        ("seed", C.c_uint64),
        ("tuning", RunCfg),
    ]


class ServerCfg(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("precision", C.c_int32),
        ("low_energy_correction", C.c_int32),
        ("slots", C.c_uint32),
        ("prep_threads", C.c_uint32),
        ("reserved", C.c_uint32),
        ("memory_budget", C.c_uint64),
        ("tuning", RunCfg),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("submitted", C.c_double),
        ("prepared", C.c_double),
        ("admitted", C.c_double),
        ("finished", C.c_double),
        ("histories", C.c_uint64),
        ("steps", C.c_uint64),
        ("events", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("status", C.c_int32),
        ("slot", C.c_int32),
        ("uploaded", C.c_double),
        ("transported", C.c_double),
        ("downloaded", C.c_double),
        ("released", C.c_double),
        ("prep_started", C.c_double),
        ("world_built", C.c_double),
        ("source_ready", C.c_double),
    ]


REQUEST_QUEUED = 0
REQUEST_RUNNING = 1
REQUEST_DONE = 2
REQUEST_UNKNOWN = 3
REQUEST_READY = 4

# every symbol include/dxmc_hip.h declares, with its ctypes signature
HIP_SYMBOLS = {
    "dxmc_hip_abi_version": (C.c_int, []),
    "dxmc_hip_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "dxmc_hip_ctx_destroy": (C.c_int, [C.c_void_p]),
    "dxmc_hip_last_error": (C.c_char_p, [C.c_void_p]),
    "dxmc_hip_mem_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "dxmc_hip_synchronize": (C.c_int, [C.c_void_p]),
    "dxmc_hip_world_upload": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.POINTER(C.c_void_p)]),
    "dxmc_hip_world_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dxmc_hip_world_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "dxmc_hip_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "dxmc_hip_host_free": (C.c_int, [C.c_void_p]),
    "dxmc_hip_world_max_density": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_uint32]),
    "dxmc_hip_tables_upload": (C.c_int, [C.c_void_p, C.POINTER(TablesDesc), C.POINTER(C.c_void_p)]),
    "dxmc_hip_tables_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dxmc_hip_tally_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "dxmc_hip_tally_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dxmc_hip_tally_reset": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dxmc_hip_tally_download": (C.c_int, [C.c_void_p, C.c_void_p, f64p, u32p, f64p]),
    "dxmc_hip_tally_normalize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_double]),
    "dxmc_hip_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Exposure), C.c_uint64, C.POINTER(RunCfg), C.c_void_p,
                               C.POINTER(C.c_int32), C.POINTER(C.c_uint64), C.POINTER(RunStats)]),
    "dxmc_hip_run_monitored": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Exposure), C.c_uint64, C.POINTER(RunCfg), C.c_void_p,
                                         C.POINTER(RunMonitor), C.POINTER(RunStats)]),
    "dxmc_hip_run_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Exposure), C.c_uint64, C.POINTER(RunCfg), C.c_void_p,
                                       C.POINTER(RunStats)]),
    "dxmc_hip_tally_mip": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_float)]),
    "dxmc_hip_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Exposure), C.c_uint64, C.POINTER(RunCfg), C.c_uint64,
                                    C.POINTER(RunCfg), C.POINTER(TuneReport)]),
    "dxmc_hip_rng_streams": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, u32p]),
    "dxmc_hip_lut_lookup": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, u32p, f64p, f64p, f64p]),
    "dxmc_hip_interact": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, u32p, C.c_double, C.c_uint64,
                                    C.c_uint64, f64p, f64p, C.POINTER(C.c_int32)]),
    "dxmc_hip_sample_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(Exposure), C.c_uint64, C.c_uint64,
                                            C.c_uint64, f64p, C.POINTER(C.c_int32)]),
}

HOST_SYMBOLS = {
    "dxmc_host_tables_create": (C.c_int, [C.c_uint32, C.POINTER(C.c_char_p), f64p, C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "dxmc_host_tables_desc": (C.c_int, [C.c_void_p, C.POINTER(TablesDesc)]),
    "dxmc_host_tables_destroy": (None, [C.c_void_p]),
    "dxmc_host_last_error": (C.c_char_p, []),
    "dxmc_host_source_exposures": (C.c_int, [C.POINTER(SourceSpec), f64p, C.c_void_p, C.POINTER(Exposure), C.c_uint64, C.POINTER(C.c_uint64)]),
    "dxmc_host_source_spectrum": (C.c_int, [C.POINTER(SourceSpec), f64p, f64p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "dxmc_host_read_raw": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64]),
    "dxmc_host_write_raw": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64]),
    "dxmc_host_world_extent": (C.c_int, [C.POINTER(HostWorld), f64p]),
    "dxmc_host_transport": (C.c_int, [C.POINTER(HostWorld), C.POINTER(SourceSpec), C.POINTER(HostTransportCfg), f64p, u32p, f64p,
                                      C.POINTER(RunStats), f64p, C.POINTER(C.c_uint64)]),
    "dxmc_host_material_data": (C.c_int, [C.c_char_p, C.c_double, f64p, f64p]),
    "dxmc_host_material_scatter_factors": (C.c_int, [C.c_char_p, C.c_uint64, f64p, f64p, f64p]),
    "dxmc_host_ctdi_phantom": (C.c_int, [C.c_uint64, C.POINTER(C.c_uint64), f64p, C.POINTER(C.c_uint8), f64p, C.POINTER(C.c_uint8),
                                         C.POINTER(C.c_uint8), C.c_uint64]),
    "dxmc_host_server_create": (C.c_int, [C.POINTER(ServerCfg), C.POINTER(C.c_void_p)]),
    "dxmc_host_server_destroy": (C.c_int, [C.c_void_p]),
    "dxmc_host_server_submit": (C.c_int, [C.c_void_p, C.POINTER(HostWorld), C.POINTER(SourceSpec), C.c_uint64, C.c_int32,
                                          C.POINTER(C.c_uint64)]),
    "dxmc_host_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "dxmc_host_server_collect": (C.c_int, [C.c_void_p, C.c_uint64, f64p, u32p, f64p, C.POINTER(RequestInfo)]),
    "dxmc_host_server_pause": (C.c_int, [C.c_void_p, C.c_int32]),
    "dxmc_host_server_clock": (C.c_double, [C.c_void_p]),
}

_libs = {}


def _load(path, symbols):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{os.path.basename(path)} is not built: run __graft_entry__.build() "
                           "(python -m dxmclib_amd.build); there is no fallback transport")
    lib = C.CDLL(path)
    for name, (res, args) in symbols.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def hip():
    """The HIP backend library (loads libamdhip64; needs no GPU to load)."""
    return _load(HIP_LIB, HIP_SYMBOLS)


def host():
    """The host facade library (loads the HIP backend as a dependency)."""
    hip()
    return _load(HOST_LIB, HOST_SYMBOLS)
