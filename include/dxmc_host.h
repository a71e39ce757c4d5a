/*
 * dxmc_host.h - C facade of the C++ host API (dxmclib_amd/include/dxmc).
 *
 * The host side builds what the device samples: the material tables of
 * AttenuationLut (reference attenuationlut.hpp:64-96), the exposure tables of
 * the sources (reference source.hpp getExposure/totalExposures) and runs the
 * complete Transport<T>::operator() (reference transport.hpp:135-202) on the
 * GPU backend of dxmc_hip.h. It exists for bindings that cannot use the C++
 * templates directly (ctypes, cgo, JNI); C++ users include dxmc.hpp.
 */
#ifndef DXMC_HOST_H
#define DXMC_HOST_H

#include "dxmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dxmc_host_tables dxmc_host_tables;

/* Material tables for n materials given by name (NIST-style mixture name,
 * chemical formula or element symbol), with the per-material maximum voxel
 * density used by the Woodcock majorant (attenuationinterpolator.hpp:48-84).
 * max_energy / min_energy in keV as AttenuationLut::generate(world, max, min). */
int dxmc_host_tables_create(uint32_t n_materials, const char* const* names, const double* max_density, double max_energy, double min_energy,
    dxmc_host_tables** out);
/* The C ABI view of the tables; valid while the object lives. */
int dxmc_host_tables_desc(const dxmc_host_tables* t, dxmc_tables_desc* out);
void dxmc_host_tables_destroy(dxmc_host_tables* t);
const char* dxmc_host_last_error(void);

enum dxmc_source_type {
    DXMC_SOURCE_PENCIL = 0,
    DXMC_SOURCE_ISOTROPIC = 1,
    DXMC_SOURCE_ISOTROPIC_CT = 2,
    DXMC_SOURCE_DX = 3,
    DXMC_SOURCE_CBCT = 4,
    DXMC_SOURCE_CT_AXIAL = 5,
    DXMC_SOURCE_CT_SPIRAL = 6,
    DXMC_SOURCE_CT_AXIAL_DUAL = 7, /* CTAxialDualSource, source.hpp:1369-1472 */
    DXMC_SOURCE_CT_SPIRAL_DUAL = 8, /* CTSpiralDualSource, source.hpp:1474-1592 */
    DXMC_SOURCE_CT_TOPOGRAM = 9 /* CTTopogramSource, source.hpp:1609-1700 */
};

/* Parameters of one source; fields not used by a type are ignored. A zero
 * field keeps the reference default of the source class. */
typedef struct {
    int32_t type;
    int32_t heel_effect; /* DX/CBCT/CT: 1 model the heel effect (default in the reference), 0 off */
    double position[3]; /* mm */
    double direction_cosines[6]; /* all zero: the source's default */
    uint64_t histories_per_exposure;
    uint64_t total_exposures; /* pencil, isotropic, DX */
    double photon_energy; /* pencil, keV */
    /* tube (DX, CBCT, CT) */
    double tube_voltage; /* kV */
    double anode_angle_deg;
    double energy_resolution; /* keV */
    double al_mm, cu_mm, sn_mm; /* filtration; negative = none */
    /* DX / CBCT */
    double sdd; /* mm */
    double field_size[2]; /* mm at the detector */
    double source_angles_deg[2]; /* DX primary / secondary angle */
    double dap; /* Gy cm2 */
    double span_angle_deg, step_angle_deg; /* CBCT */
    /* CT */
    double collimation, fov, start_angle_deg, scan_length, exposure_angle_step_deg, pitch, gantry_tilt_deg, ctdi_vol;
    double axial_step;
    uint32_t bowtie_n;
    const double* bowtie_angles; /* rad */
    const double* bowtie_weights;
    /* isotropic */
    double collimation_angles[4]; /* x0 x1 y0 y1 rad */
    uint32_t spectrum_n;
    const double* spectrum_weights;
    const double* spectrum_energies;
    /* air dose (pencil, Gy cm2) */
    double air_dose;
    /* XCare organ-based modulation of CT axial, spiral and dual sources
     * (CTSource::setUseXCareFilter, XCareFilter setters; degrees, weight);
     * a zero field keeps the filter's default */
    int32_t xcare;
    /* dual-source CT: 1 = tube B starts at start_angle_b_deg, 0 = at tube
     * A's start angle + 90 degrees (the reference default) */
    int32_t start_angle_b_set;
    double xcare_filter_angle_deg, xcare_span_angle_deg, xcare_ramp_angle_deg, xcare_low_weight;
    /* dual-source CT: tube B as the tube fields above (0 keeps the default,
     * a negative filtration clears the filters; tube B starts with tube A's
     * default 7 mm Al), its SDD, field of view and start angle, the tube
     * currents (mAs, default 100 each) and tube B's bowtie */
    double tube_b_voltage, tube_b_anode_angle_deg, tube_b_al_mm, tube_b_cu_mm, tube_b_sn_mm;
    double sdd_b, fov_b, start_angle_b_deg, mas_a, mas_b;
    uint32_t bowtie_b_n;
    const double* bowtie_b_angles; /* rad */
    const double* bowtie_b_weights;
} dxmc_source_spec;

/* Exposure table of a source aligned to a world basis (Exposure::
 * alignToDirectionCosines). Spectrum, heel and bowtie tables are appended to
 * 'tables' and the records refer to them. out may be NULL to query *n_out. */
int dxmc_host_source_exposures(const dxmc_source_spec* spec, const double world_cosines[6], dxmc_host_tables* tables, dxmc_exposure* out,
    uint64_t capacity, uint64_t* n_out);

/* Raw spectrum of a source's tube (weights per energy bin, normalised), as
 * used to build its alias table. */
int dxmc_host_source_spectrum(const dxmc_source_spec* spec, double* energies, double* weights, uint64_t capacity, uint64_t* n_out);

/* A world as plain arrays. */
typedef struct {
    uint64_t dims[3];
    double spacing[3]; /* mm */
    double origin[3]; /* mm */
    double direction_cosines[6]; /* all zero: identity */
    const uint8_t* material;
    const float* density;
    const uint8_t* measurement; /* may be NULL */
    uint32_t n_materials;
    const char* const* material_names;
} dxmc_host_world;

/* Raw binary arrays (no header, host byte order), as the reference's
 * validation program keeps phantoms and dose maps (validation.cpp:128-135,
 * 1280-1306). dxmc_host_read_raw fails unless the file holds exactly
 * 'bytes' bytes. */
int dxmc_host_read_raw(const char* path, void* data, uint64_t bytes);
int dxmc_host_write_raw(const char* path, const void* data, uint64_t bytes);

/* The nextafter-shrunk extents of a world (World::matrixExtentSafe). */
int dxmc_host_world_extent(const dxmc_host_world* w, double extent_safe[6]);

typedef struct {
    int32_t device;
    int32_t precision; /* dxmc_precision */
    int32_t low_energy_correction; /* 0, 1 (default), 2 */
    int32_t output_mode; /* dxmc_output_mode */
    int32_t use_source_calibration; /* DOSE mode: 1 = mGy via the source calibration, 0 = keV/kg */
    int32_t reserved;
    uint64_t seed;
    dxmc_run_cfg tuning; /* precision / L / seed fields are ignored */
} dxmc_host_transport_cfg;

/* Transport<T>::operator()(world, source) on the GPU. Arrays receive the
 * normalised result (any may be NULL). Returns DXMC_OK, or an error code with
 * dxmc_host_last_error(); invalid worlds/sources return DXMC_ERR_ARGUMENT. */
int dxmc_host_transport(const dxmc_host_world* world, const dxmc_source_spec* source, const dxmc_host_transport_cfg* cfg, double* dose,
    uint32_t* n_events, double* variance, dxmc_run_stats* stats, double* seconds, uint64_t* n_histories);

/* Material data probes: photo, Compton, Rayleigh, total, mass energy
 * absorption [cm2/g] at energy E (keV) and the standard density. */
int dxmc_host_material_data(const char* name, double energy, double out[5], double* density);

/* Scatter factor probes at n momentum transfers q (1/Angstrom, sin(theta/2)
 * over wavelength): the squared Rayleigh form factor and the normalised
 * Compton scatter function the Livermore samplers are built from
 * (Material::getRayleightFormFactorSquared, getComptonNormalizedScatterFactor,
 * reference material.hpp:72-73). Either output may be NULL. */
int dxmc_host_material_scatter_factors(const char* name, uint64_t n, const double* q, double* form_factor_squared, double* scatter_function);

/* The CTDI phantom the CT sources calibrate on (CTDIPhantom, reference
 * world.hpp:229-390): PMMA cylinder of 'diameter_mm' with five air holes,
 * 1 x 1 x 2.5 mm voxels, materials 0 air, 1 air (holes), 2 PMMA. hole[i] is
 * the measured hole of voxel i (1 centre, 2 west, 3 east, 4 south, 5 north,
 * 0 none); the measurement map flags the same voxels. With the arrays NULL
 * only dims and spacing are returned. */
int dxmc_host_ctdi_phantom(uint64_t diameter_mm, uint64_t dims[3], double spacing[3], uint8_t* material, double* density, uint8_t* measurement,
    uint8_t* hole, uint64_t capacity);

/* ------------------------------------------------------------ serving
 * A request server over one GPU (no counterpart in the reference, whose
 * Transport runs one call at a time on the host cores). Each request is one
 * Transport<T>::operator() call: a world and a source. Requests are
 * prepared (attenuation tables, exposure table) by a pool of host threads,
 * then admitted to the GPU in arrival order, as soon as one of 'slots'
 * device slots is idle and the request's device memory fits in what the
 * admitted requests leave of 'memory_budget'; a request that does not fit
 * waits while later ones that fit go ahead. Each slot is a context with its
 * own stream, so the kernels of admitted requests run concurrently. World,
 * tables and tally stay on the device for the request's lifetime. Results
 * depend only on the request (world, source, seed): each is the same alone
 * and in any mix. */
typedef struct dxmc_server dxmc_server;

typedef struct {
    int32_t device;
    int32_t precision; /* dxmc_precision */
    int32_t low_energy_correction; /* 0, 1, 2 */
    uint32_t slots; /* requests on the device at once, 1..8 (0: 4) */
    uint32_t prep_threads; /* host threads preparing requests (0: 4) */
    uint32_t reserved;
    uint64_t memory_budget; /* device bytes for admitted requests (0: 80 % of the free memory) */
    dxmc_run_cfg tuning; /* precision, L and seed fields are ignored */
} dxmc_server_cfg;

enum dxmc_request_state {
    DXMC_REQUEST_QUEUED = 0, /* submitted, being prepared */
    DXMC_REQUEST_RUNNING = 1, /* admitted to a slot */
    DXMC_REQUEST_DONE = 2, /* result ready (or failed: see status) */
    DXMC_REQUEST_UNKNOWN = 3, /* no such request (never submitted or already collected) */
    DXMC_REQUEST_READY = 4 /* prepared, waiting for a slot and memory */
};

typedef struct {
    double submitted, prepared, admitted, finished; /* seconds on the server clock */
    uint64_t histories, steps, events;
    uint64_t device_bytes; /* memory charged against the budget */
    int32_t status; /* DXMC_OK or the error of the request */
    int32_t slot;
    /* device phases on the server clock: host result arrays taken, world,
     * tables and tally on the device (uploaded); transport and
     * normalisation done (transported); tally on the host (downloaded);
     * device objects released (released); host objects released and the
     * result published at 'finished' */
    double uploaded, transported, downloaded, released;
    /* preparation phases on the server clock: a preparation thread takes the
     * request (prep_started), the World is built from the submitted arrays
     * (world_built), the source's spectrum and filters are computed
     * (source_ready); attenuation tables and exposure table end at
     * 'prepared' */
    double prep_started, world_built, source_ready;
} dxmc_request_info;

int dxmc_host_server_create(const dxmc_server_cfg* cfg, dxmc_server** out);
/* Stops accepting requests, finishes the admitted ones, drops the queued
 * ones and frees everything. */
int dxmc_host_server_destroy(dxmc_server* server);
/* Submit a request; the world arrays and the source spec are copied before
 * the call returns. output_mode: DXMC_OUTPUT_EV_PER_HISTORY, or
 * DXMC_OUTPUT_DOSE in keV/kg (no source calibration). Fails with
 * DXMC_ERR_ADMISSION if the request can never fit the memory budget. */
int dxmc_host_server_submit(dxmc_server* server, const dxmc_host_world* world, const dxmc_source_spec* source, uint64_t seed,
    int32_t output_mode, uint64_t* request_id);
/* Wait up to timeout_seconds (< 0: forever) for the request; *state gets
 * its dxmc_request_state. */
int dxmc_host_server_wait(dxmc_server* server, uint64_t request_id, double timeout_seconds, int32_t* state);
/* Copy a finished request's result (any array may be NULL) and info, and
 * forget the request. */
int dxmc_host_server_collect(dxmc_server* server, uint64_t request_id, double* dose, uint32_t* n_events, double* variance,
    dxmc_request_info* info);
/* Hold (paused != 0) or resume admission; requests keep being prepared
 * while admission is held (draining a device, admitting a batch at once). */
int dxmc_host_server_pause(dxmc_server* server, int32_t paused);
/* Seconds on the server clock (the time base of dxmc_request_info). */
double dxmc_host_server_clock(const dxmc_server* server);

#ifdef __cplusplus
}
#endif
#endif
