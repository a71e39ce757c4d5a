/*
 * dxmc_host.h - C facade of the C++ host API (dxmclib_amd/include/dxmc).
 *
 * The host side builds what the device samples: the material tables of
 * AttenuationLut (reference attenuationlut.hpp:64-96), the exposure tables of
 * the sources (reference source.hpp getExposure/totalExposures) and runs the
 * complete Transport<T>::operator() (reference transport.hpp:135-202) on the
 * GPU backend of dxmc_hip.h. It exists for bindings that cannot use the C++
 * templates directly (ctypes, cgo, JNI); C++ users include dxmc/dxmc.hpp.
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
    DXMC_SOURCE_CT_SPIRAL = 6
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

#ifdef __cplusplus
}
#endif
#endif
