/*
 * dxmc_hip.h - C ABI of the MI355X (gfx950) photon transport backend.
 *
 * This is the drop-in seam for the reference's CPU transport call
 *   Transport<T>::operator()(const World<T>&, Source<T>*, ProgressBar<T>*, bool)
 *   (include/dxmc/transport.hpp:135-202)
 * whose replaceable core is the parallellRun<L>() dispatch
 *   (transport.hpp:167-173, 745-778) with the table generation before it
 *   (transport.hpp:157) and the normalisation after it (transport.hpp:187-200).
 * The reference has no FFI of its own; a C++ host (dxmclib_amd/include/dxmc/
 * transport.hpp) or any FFI (ctypes, cgo, JNI) binds these entry points.
 *
 * Conventions: plain pointers and sizes, no C++ or torch types; every call
 * returns 0 (DXMC_OK) or a status code, with a message available from
 * dxmc_hip_last_error(). Contexts are independent. Several host threads may
 * call into one context: each call holds the context for its duration, so
 * calls on one context run one after the other (use one context per thread
 * for concurrency, as the request server does). Device objects (world,
 * tables, tally) belong to the context that created them and stay resident
 * in HBM until destroyed.
 */
#ifndef DXMC_HIP_H
#define DXMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DXMC_HIP_ABI_VERSION 1

enum dxmc_status {
    DXMC_OK = 0,
    DXMC_ERR_ARGUMENT = 1, /* invalid or inconsistent argument */
    DXMC_ERR_DEVICE = 2, /* HIP runtime error */
    DXMC_ERR_MEMORY = 3, /* device allocation failed */
    DXMC_ERR_CANCELLED = 4, /* *cancel became non-zero; tallies hold a partial run */
    DXMC_ERR_ADMISSION = 5, /* request does not fit the memory budget */
    DXMC_ERR_STATE = 6 /* object used with the wrong context / destroyed */
};

enum dxmc_precision {
    DXMC_PRECISION_FP64 = 0, /* plain method: FP64 throughout, one thread per history */
    DXMC_PRECISION_FP32 = 1 /* FP32 transport, persistent kernel, FP64 tallies */
};

enum dxmc_output_mode {
    DXMC_OUTPUT_EV_PER_HISTORY = 0, /* transport.hpp:780-794 */
    DXMC_OUTPUT_DOSE = 1 /* transport.hpp:796-816 */
};

typedef struct dxmc_ctx dxmc_ctx;
typedef struct dxmc_world dxmc_world;
typedef struct dxmc_tables dxmc_tables;
typedef struct dxmc_tally dxmc_tally;

/* ---------------------------------------------------------------- context */

/* Create a context on HIP device 'device'. */
int dxmc_hip_ctx_create(int device, dxmc_ctx** out);
int dxmc_hip_ctx_destroy(dxmc_ctx* ctx);
/* Last error message of the calling thread's calls on the context ("" if
 * none failed), or of its last failed ctx_create or host_alloc/free when
 * ctx is NULL. The pointer stays valid until the thread's next call on ctx. */
const char* dxmc_hip_last_error(const dxmc_ctx* ctx);
int dxmc_hip_abi_version(void);
/* Device memory in bytes: free and total. */
int dxmc_hip_mem_info(dxmc_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* Wait for all work of the context. */
int dxmc_hip_synchronize(dxmc_ctx* ctx);
/* Page-locked host memory (no counterpart in the reference, whose results
 * never leave host memory). Copies between it and the device run at the
 * link rate without staging through a bounce buffer; a caller that reuses
 * result arrays (a request server) allocates them here once. */
int dxmc_hip_host_alloc(uint64_t bytes, void** out);
int dxmc_hip_host_free(void* p);

/* ------------------------------------------------------------------ world */
/* Voxel world, reference World<T> (world.hpp:37-227): x fastest layout
 * index = z*nx*ny + y*nx + x; extent_safe is matrixExtentSafe()
 * (world.hpp:55-68, nextafter-shrunk extents). */
typedef struct {
    uint64_t dims[3];
    double spacing[3]; /* mm */
    double extent_safe[6]; /* xmin, xmax, ymin, ymax, zmin, zmax [mm] */
    const uint8_t* material; /* material index per voxel */
    const float* density_f32; /* g/cm3 per voxel: exactly one of f32 / f64 */
    const double* density_f64;
    const uint8_t* measurement; /* forced-interaction flags; NULL = none set */
    uint32_t n_materials;
} dxmc_world_desc;

/* Upload a world; the host arrays may be freed afterwards. The passes over
 * the voxels (material indices inside the material map, as World::validate
 * checks them, world.hpp:182-227; per-material maximum density; the FP32
 * voxel records) run on the device. */
int dxmc_hip_world_upload(dxmc_ctx* ctx, const dxmc_world_desc* desc, dxmc_world** out);
int dxmc_hip_world_destroy(dxmc_ctx* ctx, dxmc_world* world);
/* Bytes the world occupies on the device. */
int dxmc_hip_world_bytes(const dxmc_world* world, uint64_t* bytes);
/* Largest density of each material over the world's voxels (0 for a material
 * no voxel uses), computed on the device at upload: the scan of
 * AttenuationLutInterpolator's majorant build (attenuationinterpolator.hpp
 * :52-58). n must equal the world's n_materials. */
int dxmc_hip_world_max_density(const dxmc_world* world, double* max_density, uint32_t n);

/* ----------------------------------------------------------------- tables */
/* One grouped electron shell: ElectronShellConfiguration (material.hpp:29-58). */
typedef struct {
    double binding_energy; /* keV */
    double number_electrons; /* normalised electron fraction */
    double hartree_fock_j0; /* Compton profile at pz = 0 */
    double photo_ionization_probability;
    double fluorescence_yield;
    double line_probability[3];
    double line_energy[3]; /* keV */
} dxmc_shell;

/* Heel filter grid: HeelFilter (beamfilters.hpp:435-569). */
typedef struct {
    double energy_start, energy_step, angle_start, angle_step;
    uint32_t energy_size, angle_size;
    uint32_t weight_offset; /* into heel_weight, energy_size*angle_size values */
    uint32_t reserved;
} dxmc_heel;

/* All host-built tables of AttenuationLut (attenuationlut.hpp:42-275) and
 * of the sources' samplers. */
typedef struct {
    uint32_t n_materials;
    /* log-log attenuation LUT, AttenuationLutInterpolator (attenuationinterpolator.hpp:30-250) */
    uint32_t lut_nodes; /* grid size after the first node is dropped (= intervals) */
    const double* lut_grid; /* [lut_nodes] log10(E/keV) */
    const double* lut_coefficients; /* [n_materials][lut_nodes][photo, compton, rayleigh][b, a] */
    const double* lut_majorant; /* [lut_nodes][b, a] of log10(1/mu_max) */
    uint32_t lut_linear_index;
    double lut_linear_step;
    double lut_linear_energy;
    /* Rayleigh form factor sampler RITA<T,56> (dxmcrandom.hpp:333-501) */
    const double* rita; /* [n_materials][x, e, a, b][56] */
    /* Compton scatter factor CubicSplineInterpolator<T,16> (interpolation.hpp:84-177) */
    const double* spline; /* [n_materials][60 coefficients, start, step, stop] */
    const dxmc_shell* shells; /* [n_materials][12] */
    /* spectra: SpecterDistribution alias tables (dxmcrandom.hpp:173-330) */
    uint32_t n_spectra;
    const uint32_t* spectrum_offset; /* [n_spectra + 1] */
    const double* spectrum_probability;
    const uint32_t* spectrum_alias;
    const double* spectrum_energy;
    /* heel filters */
    uint32_t n_heel;
    const dxmc_heel* heel;
    const double* heel_weight;
    /* bowtie filters: sorted |angle| -> normalised weight (beamfilters.hpp:63-198) */
    uint32_t n_bowtie;
    const uint32_t* bowtie_offset; /* [n_bowtie + 1] */
    const double* bowtie_angle;
    const double* bowtie_weight;
} dxmc_tables_desc;

int dxmc_hip_tables_upload(dxmc_ctx* ctx, const dxmc_tables_desc* desc, dxmc_tables** out);
int dxmc_hip_tables_destroy(dxmc_ctx* ctx, dxmc_tables* tables);

/* -------------------------------------------------------------- exposures */
/* One exposure, already aligned to the world basis
 * (Exposure::alignToDirectionCosines, exposure.hpp:268-278). Fields follow
 * Exposure<T> (exposure.hpp:36-49). */
typedef struct {
    double position[3]; /* mm */
    double direction_cosines[6];
    double beam_direction[3];
    double collimation_angles[4]; /* x0, x1, y0, y1 [rad] */
    double weight; /* beam intensity weight */
    double mono_energy; /* keV, used when spectrum < 0 */
    uint64_t n_histories;
    int32_t spectrum; /* index into the spectra of the tables, -1 = monoenergetic */
    int32_t heel; /* -1 = none */
    int32_t bowtie; /* -1 = none */
    int32_t reserved;
} dxmc_exposure;

/* ------------------------------------------------------------------- tally */
/* Result<T> (transport.hpp:42-65): dose (energy imparted) f64, event count
 * u32 and per-event variance sum f64 per voxel. Runs accumulate into the
 * tally's sums; normalisation writes its result arrays from them. */
int dxmc_hip_tally_create(dxmc_ctx* ctx, uint64_t n_voxels, dxmc_tally** out);
int dxmc_hip_tally_destroy(dxmc_ctx* ctx, dxmc_tally* tally);
int dxmc_hip_tally_reset(dxmc_ctx* ctx, dxmc_tally* tally);
/* Copy to host; any pointer may be NULL to skip that array. After a
 * normalisation (and no run since) the normalised arrays, else the sums. */
int dxmc_hip_tally_download(dxmc_ctx* ctx, const dxmc_tally* tally, double* dose, uint32_t* n_events, double* variance);
/* Maximum intensity projection of a tally's dose array along axis (0 x,
 * 1 y, 2 z) of the world's grid, laid out as the run preview of
 * dxmc_hip_run_monitored; image holds the product of the other two
 * dimensions. */
int dxmc_hip_tally_mip(dxmc_ctx* ctx, const dxmc_tally* tally, const dxmc_world* world, int32_t axis, float* image);
/* FP64 normalisation on the device, from the sums into the result arrays
 * (the sums stay, so normalising again gives the same arrays):
 * EV_PER_HISTORY: normalizeScoring (transport.hpp:780-794),
 * DOSE: energyImpartedToDose with a calibration factor (transport.hpp:796-816). */
int dxmc_hip_tally_normalize(dxmc_ctx* ctx, dxmc_tally* tally, const dxmc_world* world, int output_mode, uint64_t n_histories,
    double calibration);

/* -------------------------------------------------------------- transport */
typedef struct {
    int32_t precision; /* dxmc_precision */
    int32_t low_energy_correction; /* 0 NONE, 1 LIVERMORE, 2 IA (lowenergycorrectionmodel.hpp) */
    uint64_t seed; /* history h uses state1 = 2h+1, state0 = splitmix64(seed ^ splitmix64(h)) */
    uint64_t first_history; /* global index of the first history of exposure 0 */
    uint64_t histories_per_launch; /* 0: automatic; bounds cancel/progress latency; FP32 at most 2^32 */
    uint32_t block_size; /* 0: tuned default */
    uint32_t blocks_per_cu; /* FP32 persistent kernel occupancy, 0: tuned default */
    uint32_t regeneration_threshold; /* FP32: dead lanes per wave before they take queued new photons, 0: tuned */
    uint32_t flags; /* DXMC_RUN_* bits */
    uint32_t interaction_threshold; /* FP32: pending interactions per wave before they run, 0: tuned */
    uint32_t queue_depth; /* FP32: histories a wave claims per refill of its pool, 0: tuned */
    uint32_t waves_per_simd; /* FP32: kernel variant by register budget, 5, 6 or 7 resident waves per SIMD
                              * (7 only at L = 0, 1 on worlds of 1-byte codes or palette indices, else 6);
                              * 0: 7 on those worlds and 6 on others for L = 0, 1, 5 for L = 2 */
    uint32_t lds_tables; /* FP32: sampling tables in LDS, 1 none, 2 all but the attenuation coefficients, 3 all (0: tuned) */
    uint32_t block_bounds; /* FP32 on worlds with 1-byte codes: 1 settle Woodcock decisions from the code range of the
                            * voxel's 8x8x8 block before its own code, 2 from the voxel's code only (same results);
                            * 0: 1, except at L = 2 when the code array fits the 256 MB Infinity Cache */
    uint32_t reserved0; /* 0 */
} dxmc_run_cfg;

/* dxmc_run_cfg.flags */
#define DXMC_RUN_COUNTERS 1u /* collect step/event counters */
#define DXMC_RUN_AUTOTUNE 2u /* fill zero tuning fields from the autotuner (tuned once per world/tables/precision/L and kept in the context) */
#define DXMC_RUN_EXACT_STEPS 4u /* FP32: evaluate every Woodcock step from the voxel record and the full attenuation, without the bounds that settle most steps (same results, for checks) */

typedef struct {
    uint64_t histories; /* histories started */
    uint64_t steps; /* Woodcock steps ending inside the world */
    uint64_t events; /* scoring events (dose/nEvents/variance updates) */
    double kernel_ms; /* summed transport kernel time (HIP events) */
    uint64_t launches;
} dxmc_run_stats;

/* Transport all histories of n_exposures exposures and add their energy
 * imparted to the tally (not reset). cancel (may be NULL) is polled between
 * launches: a non-zero value stops the run with DXMC_ERR_CANCELLED.
 * exposures_done (may be NULL) receives the number of exposures finished.
 * Both are accessed with atomic acquire loads and release stores, so another
 * thread may set *cancel or read *exposures_done with atomic operations of
 * the same width (C11 _Atomic, std::atomic<int32_t> / <uint64_t>).
 * Replaces parallellRun<L> (transport.hpp:765-778) and its workers. */
int dxmc_hip_run(dxmc_ctx* ctx, const dxmc_world* world, const dxmc_tables* tables, const dxmc_exposure* exposures,
    uint64_t n_exposures, const dxmc_run_cfg* cfg, dxmc_tally* tally, const volatile int32_t* cancel,
    volatile uint64_t* exposures_done, dxmc_run_stats* stats);

/* Live view of a running transport: the dose preview of the reference's
 * ProgressBar::computeDoseProgressImage (progressbar.hpp:116-235), computed
 * on the device. 'preview' is called on the thread running
 * dxmc_hip_run_monitored, between kernel launches, at most every
 * preview_interval_seconds and once after the last launch, with the maximum
 * intensity projection of the tally's accumulated dose array along
 * preview_axis (0 x, 1 y, 2 z): width x height floats, first index fastest,
 * over (y, z) for x, (x, z) for y and (x, y) for z (dxmc_hip_tally_mip). */
typedef void (*dxmc_preview_fn)(void* user, const float* image, uint64_t width, uint64_t height);
typedef struct {
    const volatile int32_t* cancel; /* as dxmc_hip_run; may be NULL */
    volatile uint64_t* exposures_done; /* as dxmc_hip_run; may be NULL */
    dxmc_preview_fn preview; /* NULL: no preview */
    void* preview_user;
    int32_t preview_axis;
    float preview_interval_seconds;
} dxmc_run_monitor;

/* dxmc_hip_run with a monitor (cancel, progress, live preview); monitor may
 * be NULL. */
int dxmc_hip_run_monitored(dxmc_ctx* ctx, const dxmc_world* world, const dxmc_tables* tables, const dxmc_exposure* exposures,
    uint64_t n_exposures, const dxmc_run_cfg* cfg, dxmc_tally* tally, const dxmc_run_monitor* monitor, dxmc_run_stats* stats);

/* dxmc_hip_run without allocation. No counterpart in the reference (its CPU
 * workers allocate nothing per call either); for callers that cannot
 * allocate device or host memory during a run: code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc. The run's own device
 * memory (exposure records, history prefix sums, counters) lives in a fixed
 * scratch array of the library, one per GPU, held until the process ends
 * (about 3.5 MB for DXMC_NOALLOC_MAX_EXPOSURES); world, tables and tally are
 * the caller's, created beforehand. Limits: at most
 * DXMC_NOALLOC_MAX_EXPOSURES exposures (more is DXMC_ERR_ARGUMENT); no
 * cancel, progress or preview; DXMC_RUN_AUTOTUNE only reuses a configuration
 * tuned before for the same world, tables, precision and model. Calls on one
 * GPU share its array: the library runs them one after the other (the call
 * returns when the run is done), so only calls from separate processes run
 * at once. Same results as dxmc_hip_run for the same inputs. */
#define DXMC_NOALLOC_MAX_EXPOSURES 8192
int dxmc_hip_run_noalloc(dxmc_ctx* ctx, const dxmc_world* world, const dxmc_tables* tables, const dxmc_exposure* exposures, uint64_t n_exposures,
    const dxmc_run_cfg* cfg, dxmc_tally* tally, dxmc_run_stats* stats);

/* Autotuner of the transport kernels (no counterpart in the reference, whose
 * CPU threads need no tuning). Times trial launches of trial_histories
 * histories drawn evenly from the exposure table into a scratch tally and
 * searches, one parameter at a time from the defaults, over
 *   FP32: waves_per_simd, lds_tables, block_bounds, regeneration_threshold,
 *         interaction_threshold, queue_depth, blocks_per_cu;
 *   FP64: block_size.
 * Tuning fields that are non-zero in 'base' are kept fixed. 'best' receives
 * 'base' with the chosen values filled in; the choice is also kept in the
 * context for DXMC_RUN_AUTOTUNE runs on the same world and tables. The
 * tuning fields change speed only: per-history results do not depend on
 * them. */
typedef struct {
    uint32_t trials; /* trial launches timed */
    uint32_t reserved;
    double default_rate; /* histories/s with all tuning fields 0 */
    double best_rate; /* histories/s with 'best' */
} dxmc_tune_report;
int dxmc_hip_autotune(dxmc_ctx* ctx, const dxmc_world* world, const dxmc_tables* tables, const dxmc_exposure* exposures, uint64_t n_exposures,
    const dxmc_run_cfg* base, uint64_t trial_histories, dxmc_run_cfg* best, dxmc_tune_report* report);

/* ------------------------------------------------------ parity entry points */
/* PCG32 streams (RandomState::pcg32, dxmcrandom.hpp:156-165) of histories
 * first_history .. first_history+n-1 under 'seed': out[i*draws + k]. */
int dxmc_hip_rng_streams(dxmc_ctx* ctx, uint64_t seed, uint64_t first_history, uint64_t n, uint32_t draws, uint32_t* out);

/* LUT lookups (attenuationinterpolator.hpp:207-248) at n (material, energy)
 * pairs in the given precision: out_attenuation[i*3 + type] (photo,
 * Compton, Rayleigh, cm2/g) and out_majorant_inv[i] (cm). */
int dxmc_hip_lut_lookup(dxmc_ctx* ctx, const dxmc_tables* tables, int32_t precision, uint64_t n, const uint32_t* material,
    const double* energy, double* out_attenuation, double* out_majorant_inv);

/* One interaction per sample: kind 0 photoelectric (transport.hpp:216-263),
 * 1 Compton (:300-483), 2 Rayleigh (:264-298), 3 interaction dispatch
 * computeInteractions (:583-638), 4 forced dispatch computeInteractionsForced
 * (:523-581) with event probability 'forced_probability'. Sample i uses the
 * stream of history first_history+i. particles: [n][pos3, dir3, energy,
 * weight] in/out; energy_out[i]: energy imparted (kinds 0-2: the function's
 * return value; 3-4: the tallied value, 0 if none); alive_out[i] (kinds
 * 3-4): the function's return value. */
int dxmc_hip_interact(dxmc_ctx* ctx, const dxmc_tables* tables, int32_t precision, int32_t low_energy_correction,
    int32_t kind, uint64_t n, const uint32_t* material, double forced_probability, uint64_t seed, uint64_t first_history,
    double* particles, double* energy_out, int32_t* alive_out);

/* Exposure::sampleParticle (exposure.hpp:280-304) followed by
 * transportParticleToWorld (transport.hpp:702-728) for history
 * first_history+i of exposure 0: particles [n][8], inside[i]. */
int dxmc_hip_sample_particles(dxmc_ctx* ctx, const dxmc_world* world, const dxmc_tables* tables, int32_t precision,
    const dxmc_exposure* exposure, uint64_t seed, uint64_t first_history, uint64_t n, double* particles, int32_t* inside);

#ifdef __cplusplus
}
#endif
#endif
