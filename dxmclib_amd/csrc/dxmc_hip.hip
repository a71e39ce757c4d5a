// MI355X (gfx950) backend of the dxmc photon transport: implementation of
// the C ABI declared in include/dxmc_hip.h.
//
// Kernels in this file:
//   transportPlain<T,L>   one thread per history (the plain method for
//                         T = double; also instantiated for float)
//   rngStreams            PCG32 stream dump (parity)
//   lutLookup<T>          attenuation + majorant lookups (parity)
//   interact<T,L>         single interactions (parity)
//   sampleParticles<T>    source sampling + world entry (parity)
//   normalizeTally        FP64 normalisation (eV/history or dose)
// The persistent FP32 transport kernel lives in transport_fp32.hip.
//
// This file is compiled without FMA contraction (-ffp-contract=off) so the
// FP64 path rounds like the CPU oracle.
#include "dxmc_hip.h"
#include "internal.hpp"

#include <algorithm>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace dxmc_dev;
using namespace dxmc_backend;


namespace {

// ---------------------------------------------------------------- kernels

template <typename T, int L>
__global__ __launch_bounds__(256) void transportPlain(World<T> w, Tables<T> t, const Exposure<T>* exps, const uint64_t* expStart,
    uint32_t nExp, uint64_t hBegin, uint64_t hEnd, uint64_t seed, uint64_t firstHistory, Tally tally, Counters* ctr)
{
    const uint64_t h = hBegin + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    Scorer sc { tally, 0 };
    uint32_t steps = 0;
    if (h < hEnd) {
        const uint32_t e = upperBound(expStart, nExp + 1, h) - 1;
        Rng rng = Rng::forHistory(seed, firstHistory + h);
        Particle<T> p;
        sampleParticle(exps[e], t, rng, p);
        if (transportToWorld(w, p))
            woodcock<T, L>(w, t, p, rng, sc, steps);
    }
    if (ctr)
        accumulateCounters(ctr, steps, sc.events);
}

__global__ void rngStreams(uint64_t seed, uint64_t first, uint64_t n, uint32_t draws, uint32_t* out)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Rng rng = Rng::forHistory(seed, first + i);
    for (uint32_t k = 0; k < draws; ++k)
        out[i * draws + k] = rng.next();
}

template <typename T>
__global__ void lutLookup(Tables<T> t, uint64_t n, const uint32_t* mat, const double* energy, double* att, double* maj)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    T a[3];
    const T E = static_cast<T>(energy[i]);
    attenuation(t, mat[i], E, a);
    for (int k = 0; k < 3; ++k)
        att[i * 3 + k] = static_cast<double>(a[k]);
    maj[i] = static_cast<double>(majorantInverse(t, E));
}

template <typename T, int L>
__global__ void interact(Tables<T> t, int kind, uint64_t n, const uint32_t* mat, double forcedProbability, uint64_t seed, uint64_t first,
    double* particles, double* energyOut, int32_t* aliveOut, Tally tally)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Rng rng = Rng::forHistory(seed, first + i);
    Particle<T> p;
    double* q = particles + i * 8;
    for (int k = 0; k < 3; ++k) {
        p.pos[k] = static_cast<T>(q[k]);
        p.dir[k] = static_cast<T>(q[3 + k]);
    }
    p.energy = static_cast<T>(q[6]);
    p.weight = static_cast<T>(q[7]);
    T e = 0;
    int alive = 1;
    Scorer sc { tally, 0 };
    bool update = false;
    if (kind == 0)
        e = photoAbsorption<T, L>(t, p, mat[i], rng);
    else if (kind == 1)
        e = comptonScatter<T, L>(t, p, mat[i], rng);
    else if (kind == 2)
        rayleighScatter<T, L>(t, p, mat[i], rng);
    else {
        T att[3];
        attenuation(t, mat[i], p.energy, att);
        if (kind == 3)
            alive = computeInteractions<T, L>(t, att, p, mat[i], sc, i, rng, update) ? 1 : 0;
        else
            alive = computeInteractionsForced<T, L>(t, static_cast<T>(forcedProbability), att, p, mat[i], sc, i, rng, update) ? 1 : 0;
    }
    for (int k = 0; k < 3; ++k) {
        q[k] = static_cast<double>(p.pos[k]);
        q[3 + k] = static_cast<double>(p.dir[k]);
    }
    q[6] = static_cast<double>(p.energy);
    q[7] = static_cast<double>(p.weight);
    if (kind <= 2)
        energyOut[i] = static_cast<double>(e);
    aliveOut[i] = alive;
}

template <typename T>
__global__ void sampleParticles(World<T> w, Tables<T> t, Exposure<T> x, uint64_t seed, uint64_t first, uint64_t n, double* out, int32_t* inside)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Rng rng = Rng::forHistory(seed, first + i);
    Particle<T> p;
    sampleParticle(x, t, rng, p);
    inside[i] = transportToWorld(w, p) ? 1 : 0;
    double* q = out + i * 8;
    for (int k = 0; k < 3; ++k) {
        q[k] = static_cast<double>(p.pos[k]);
        q[3 + k] = static_cast<double>(p.dir[k]);
    }
    q[6] = static_cast<double>(p.energy);
    q[7] = static_cast<double>(p.weight);
}

// normalizeScoring (transport.hpp:780-794) / energyImpartedToDose (:796-816)
// from the voxel scores into the result arrays: mode < 0 copies the sums
__global__ void normalizeTally(const VoxelScore* __restrict__ scores, double* __restrict__ dose, uint32_t* __restrict__ events,
    double* __restrict__ variance, uint64_t n, int mode, uint64_t nHist, double voxelVolume, double calibration, const double* dens64)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const VoxelScore v = scores[i];
    events[i] = v.events;
    if (mode < 0) {
        dose[i] = v.dose;
        variance[i] = v.variance;
    } else if (mode == DXMC_OUTPUT_EV_PER_HISTORY) {
        const double hInv = 1.0 / static_cast<double>(nHist - 1);
        const double hd = 1e3 / static_cast<double>(nHist);
        const double hv = 1e6 / static_cast<double>(nHist);
        const double d = v.dose * hd;
        dose[i] = d;
        variance[i] = (v.variance * hv - d * d) * hInv;
    } else {
        const double de = dens64[i];
        const double mass = de * voxelVolume * 0.001;
        const double factor = calibration / mass;
        dose[i] = de > 0.0 ? v.dose * factor : 0.0;
        variance[i] = de > 0.0 ? v.variance * factor * factor : 0.0;
    }
}

// Maximum of the dose along one grid axis for every pixel of the other two
// (x fastest in the grid). Along y or z one thread per pixel walks the
// axis, reading consecutive x across the wave; along x one wave per pixel
// reduces the row.
// dose[i * stride] is voxel i's dose (stride 4 over the voxel scores, 1
// over the result array)
__global__ __launch_bounds__(256) void tallyMip(const double* __restrict__ dose, uint32_t stride, uint32_t nx, uint32_t ny, uint32_t nz, int axis,
    float* __restrict__ image)
{
    const uint64_t slice = uint64_t(nx) * ny;
    if (axis == 0) {
        const uint64_t pixel = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
        const uint32_t lane = threadIdx.x % 64;
        if (pixel >= uint64_t(ny) * nz)
            return;
        const double* row = dose + pixel * nx * stride; // pixel = y + ny * z
        double m = 0;
        for (uint32_t x = lane; x < nx; x += 64)
            m = fmax(m, row[uint64_t(x) * stride]);
        for (int off = 32; off > 0; off >>= 1)
            m = fmax(m, __shfl_xor(m, off));
        if (lane == 0)
            image[pixel] = static_cast<float>(m);
        return;
    }
    const uint64_t pixel = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (axis == 1) { // pixel = x + nx * z
        if (pixel >= uint64_t(nx) * nz)
            return;
        const uint64_t x = pixel % nx, z = pixel / nx;
        const double* p = dose + (z * slice + x) * stride;
        double m = 0;
        for (uint32_t y = 0; y < ny; ++y)
            m = fmax(m, p[uint64_t(y) * nx * stride]);
        image[pixel] = static_cast<float>(m);
        return;
    }
    if (pixel >= slice) // axis 2, pixel = x + nx * y
        return;
    const double* p = dose + pixel * stride;
    double m = 0;
    for (uint32_t z = 0; z < nz; ++z)
        m = fmax(m, p[uint64_t(z) * slice * stride]);
    image[pixel] = static_cast<float>(m);
}

__global__ void fillZero(VoxelScore* scores, uint64_t n)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        scores[i] = VoxelScore {};
}

uint32_t gridFor(uint64_t n, uint32_t block) { return static_cast<uint32_t>((n + block - 1) / block); }

template <typename T>
__host__ __device__ Exposure<T> toDevice(const dxmc_exposure& e)
{
    Exposure<T> x;
    for (int i = 0; i < 3; ++i) {
        x.pos[i] = static_cast<T>(e.position[i]);
        x.beamDir[i] = static_cast<T>(e.beam_direction[i]);
    }
    for (int i = 0; i < 6; ++i)
        x.cosines[i] = static_cast<T>(e.direction_cosines[i]);
    for (int i = 0; i < 4; ++i)
        x.collimation[i] = static_cast<T>(e.collimation_angles[i]);
    x.weight = static_cast<T>(e.weight);
    x.monoEnergy = static_cast<T>(e.mono_energy);
    x.spectrum = e.spectrum;
    x.heel = e.heel;
    x.bowtie = e.bowtie;
    x.pad = 0;
    return x;
}

template <typename T, int L>
void launchPlain(const World<T>& w, const Tables<T>& t, const Exposure<T>* exps, const uint64_t* expStart, uint32_t nExp, uint64_t hb, uint64_t he,
    uint64_t seed, uint64_t first, const Tally& tally, Counters* ctr, uint32_t block, hipStream_t s)
{
    const uint64_t n = he - hb;
    hipLaunchKernelGGL((transportPlain<T, L>), dim3(gridFor(n, block)), dim3(block), 0, s, w, t, exps, expStart, nExp, hb, he, seed, first, tally, ctr);
}

template <typename T>
void launchPlainL(int L, const World<T>& w, const Tables<T>& t, const Exposure<T>* exps, const uint64_t* expStart, uint32_t nExp, uint64_t hb,
    uint64_t he, uint64_t seed, uint64_t first, const Tally& tally, Counters* ctr, uint32_t block, hipStream_t s)
{
    if (L == 0)
        launchPlain<T, 0>(w, t, exps, expStart, nExp, hb, he, seed, first, tally, ctr, block, s);
    else if (L == 1)
        launchPlain<T, 1>(w, t, exps, expStart, nExp, hb, he, seed, first, tally, ctr, block, s);
    else
        launchPlain<T, 2>(w, t, exps, expStart, nExp, hb, he, seed, first, tally, ctr, block, s);
}

template <typename T>
void launchInteract(int L, const Tables<T>& t, int kind, uint64_t n, const uint32_t* mat, double fp, uint64_t seed, uint64_t first, double* part,
    double* eout, int32_t* aout, const Tally& tally, hipStream_t s)
{
    const dim3 g(gridFor(n, 128)), b(128);
    if (L == 0)
        hipLaunchKernelGGL((interact<T, 0>), g, b, 0, s, t, kind, n, mat, fp, seed, first, part, eout, aout, tally);
    else if (L == 1)
        hipLaunchKernelGGL((interact<T, 1>), g, b, 0, s, t, kind, n, mat, fp, seed, first, part, eout, aout, tally);
    else
        hipLaunchKernelGGL((interact<T, 2>), g, b, 0, s, t, kind, n, mat, fp, seed, first, part, eout, aout, tally);
}

} // namespace

// ============================================================== C ABI



extern "C" {

int dxmc_hip_abi_version(void) { return DXMC_HIP_ABI_VERSION; }

int dxmc_hip_ctx_create(int device, dxmc_ctx** out)
{
    if (!out)
        return setGlobalError("ctx_create: out is NULL"), DXMC_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return setGlobalError("ctx_create: no HIP device available"), DXMC_ERR_DEVICE;
    if (device < 0 || device >= count)
        return setGlobalError("ctx_create: device index out of range"), DXMC_ERR_ARGUMENT;
    auto* c = new (std::nothrow) dxmc_ctx();
    if (!c)
        return DXMC_ERR_MEMORY;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess
        || hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess
        || hipMalloc(&c->historyCounter, sizeof(unsigned long long)) != hipSuccess) {
        setGlobalError("ctx_create: HIP initialisation failed");
        delete c;
        return DXMC_ERR_DEVICE;
    }
    // the no-allocation run's scratch array: its address on this device,
    // taken with the device current (a failure only disables that entry point)
    if (noallocScratch(device, &c->scratch) != DXMC_OK)
        c->scratch = nullptr;
    // keep memory freed by hipFreeAsync in the device pool for reuse instead
    // of returning it at every synchronisation
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    hipGetLastError();
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->numCU = prop.multiProcessorCount;
        c->arch = prop.gcnArchName;
    }
    *out = c;
    return DXMC_OK;
}

int dxmc_hip_ctx_destroy(dxmc_ctx* c)
{
    if (!c)
        return DXMC_ERR_ARGUMENT;
    // wait for calls of other threads that still hold the context
    c->lock.lock();
    c->lock.unlock();
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    c->releaseStaging();
    hipFree(c->historyCounter);
    hipStreamDestroy(c->stream);
    threadErrors().erase(c);
    delete c;
    return DXMC_OK;
}

const char* dxmc_hip_last_error(const dxmc_ctx* c)
{
    if (!c)
        return globalError();
    const auto& errors = threadErrors();
    const auto it = errors.find(c);
    return it != errors.end() ? it->second.c_str() : "";
}

int dxmc_hip_mem_info(dxmc_ctx* c, uint64_t* freeBytes, uint64_t* totalBytes)
{
    if (!c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    size_t f = 0, t = 0;
    DXMC_CHECK(c, hipSetDevice(c->device));
    DXMC_CHECK(c, hipMemGetInfo(&f, &t));
    if (freeBytes)
        *freeBytes = f;
    if (totalBytes)
        *totalBytes = t;
    return DXMC_OK;
}

int dxmc_hip_host_alloc(uint64_t bytes, void** out)
{
    if (!out || bytes == 0)
        return setGlobalError("host_alloc: NULL argument or zero size"), DXMC_ERR_ARGUMENT;
    *out = nullptr;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        hipGetLastError();
        *out = nullptr;
        return setGlobalError("host_alloc: page-locked allocation of " + std::to_string(bytes) + " bytes failed"), DXMC_ERR_MEMORY;
    }
    registerPinned(*out, bytes);
    return DXMC_OK;
}

int dxmc_hip_host_free(void* p)
{
    if (!p)
        return DXMC_OK;
    if (!unregisterPinned(p))
        return setGlobalError("host_free: not from dxmc_hip_host_alloc"), DXMC_ERR_ARGUMENT;
    hipHostFree(p);
    return DXMC_OK;
}

int dxmc_hip_synchronize(dxmc_ctx* c)
{
    if (!c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    return DXMC_OK;
}

// World upload, destroy and queries: world.hip

// ----------------------------------------------------------------- tables

int dxmc_hip_tables_upload(dxmc_ctx* c, const dxmc_tables_desc* d, dxmc_tables** out)
{
    if (!c || !d || !out)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    *out = nullptr;
    std::string why;
    if (!validateTables(*d, why))
        return c->fail(DXMC_ERR_ARGUMENT, "tables_upload: " + why);
    DXMC_CHECK(c, hipSetDevice(c->device));
    auto* t = new dxmc_tables();
    t->ctx = c;
    int st = buildDeviceTables(c, *d, *t);
    if (st != DXMC_OK) {
        dxmc_hip_tables_destroy(c, t);
        return st;
    }
    *out = t;
    return DXMC_OK;
}

int dxmc_hip_tables_destroy(dxmc_ctx* c, dxmc_tables* t)
{
    if (!c || !t)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    if (t->ctx != c)
        return c->fail(DXMC_ERR_STATE, "tables_destroy: tables belong to another context");
    std::erase_if(c->tuned, [t](const auto& kv) { return std::get<1>(kv.first) == t; });
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (void* p : t->allocations)
        c->release(p);
    delete t;
    return DXMC_OK;
}

// ------------------------------------------------------------------ tally

int dxmc_hip_tally_create(dxmc_ctx* c, uint64_t n, dxmc_tally** out)
{
    if (!c || !out || n == 0)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    *out = nullptr;
    DXMC_CHECK(c, hipSetDevice(c->device));
    auto* t = new dxmc_tally();
    t->ctx = c;
    t->n = n;
    int st;
    if ((st = c->alloc(&t->scores, n * sizeof(VoxelScore))) || (st = c->alloc(&t->dose, n * sizeof(double)))
        || (st = c->alloc(&t->nEvents, n * sizeof(uint32_t))) || (st = c->alloc(&t->variance, n * sizeof(double)))) {
        dxmc_hip_tally_destroy(c, t);
        return st;
    }
    if (reinterpret_cast<uintptr_t>(t->scores) % alignof(VoxelScore) != 0) {
        dxmc_hip_tally_destroy(c, t);
        return c->fail(DXMC_ERR_MEMORY, "tally_create: voxel scores not 32-byte aligned");
    }
    if ((st = dxmc_hip_tally_reset(c, t))) {
        dxmc_hip_tally_destroy(c, t);
        return st;
    }
    *out = t;
    return DXMC_OK;
}

int dxmc_hip_tally_destroy(dxmc_ctx* c, dxmc_tally* t)
{
    if (!c || !t)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    if (t->ctx != c)
        return c->fail(DXMC_ERR_STATE, "tally_destroy: tally belongs to another context");
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    c->release(t->scores);
    c->release(t->dose);
    c->release(t->nEvents);
    c->release(t->variance);
    hipStreamSynchronize(c->stream);
    delete t;
    return DXMC_OK;
}

int dxmc_hip_tally_reset(dxmc_ctx* c, dxmc_tally* t)
{
    if (!c || !t || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    DXMC_CHECK(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(fillZero, dim3(gridFor(t->n, 256)), dim3(256), 0, c->stream, static_cast<VoxelScore*>(t->scores), t->n);
    DXMC_CHECK(c, hipGetLastError());
    t->arraysCurrent = false;
    return DXMC_OK;
}

int dxmc_hip_tally_download(dxmc_ctx* c, const dxmc_tally* t, double* dose, uint32_t* nEvents, double* variance)
{
    if (!c || !t || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    DXMC_CHECK(c, hipSetDevice(c->device));
    if (!t->arraysCurrent) { // the sums as they stand
        if (int st = launchNormalize(c, t, nullptr, -1, 0, 0.0))
            return st;
    }
    // through the context's page-locked staging chunks (a direct copy into
    // pageable memory ran at about 7 GB/s), or direct into page-locked memory
    int st = DXMC_OK;
    if (dose && (st = c->download(dose, t->dose, t->n * sizeof(double))))
        return st;
    if (nEvents && (st = c->download(nEvents, t->nEvents, t->n * sizeof(uint32_t))))
        return st;
    if (variance && (st = c->download(variance, t->variance, t->n * sizeof(double))))
        return st;
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    return DXMC_OK;
}

int dxmc_hip_tally_mip(dxmc_ctx* c, const dxmc_tally* t, const dxmc_world* w, int32_t axis, float* image)
{
    if (!c || !t || !w || !image || axis < 0 || axis > 2)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "tally_mip: NULL argument or axis not 0, 1 or 2") : DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    if (t->ctx != c || w->ctx != c)
        return c->fail(DXMC_ERR_STATE, "tally_mip: tally or world belongs to another context");
    if (t->n != w->n)
        return c->fail(DXMC_ERR_ARGUMENT, "tally_mip: tally and world sizes differ");
    DXMC_CHECK(c, hipSetDevice(c->device));
    const uint64_t pixels = w->n / w->dims[axis];
    DeviceBuffer img(c);
    uint64_t width = 0, height = 0;
    if (int st = img.alloc(pixels * sizeof(float)))
        return st;
    if (int st = launchTallyMip(c, t, w, axis, static_cast<float*>(img.p), width, height))
        return st;
    return c->download(image, img.p, pixels * sizeof(float));
}

int dxmc_hip_tally_normalize(dxmc_ctx* c, dxmc_tally* t, const dxmc_world* w, int mode, uint64_t nHist, double calibration)
{
    if (!c || !t || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    if (mode != DXMC_OUTPUT_EV_PER_HISTORY && mode != DXMC_OUTPUT_DOSE)
        return c->fail(DXMC_ERR_ARGUMENT, "tally_normalize: unknown output mode");
    if (mode == DXMC_OUTPUT_DOSE && (!w || w->n != t->n))
        return c->fail(DXMC_ERR_ARGUMENT, "tally_normalize: dose needs the world the tally was scored on");
    if (mode == DXMC_OUTPUT_EV_PER_HISTORY && nHist < 2)
        return c->fail(DXMC_ERR_ARGUMENT, "tally_normalize: need at least two histories");
    DXMC_CHECK(c, hipSetDevice(c->device));
    return launchNormalize(c, t, w, mode, nHist, calibration);
}

// -------------------------------------------------------------- transport

int dxmc_hip_run(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* exposures, uint64_t nExp, const dxmc_run_cfg* cfg,
    dxmc_tally* tally, const volatile int32_t* cancel, volatile uint64_t* exposuresDone, dxmc_run_stats* stats)
{
    const dxmc_run_monitor monitor { cancel, exposuresDone, nullptr, nullptr, 0, 0.0f };
    return dxmc_hip_run_monitored(c, w, t, exposures, nExp, cfg, tally, &monitor, stats);
}

int dxmc_hip_run_monitored(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* exposures, uint64_t nExp,
    const dxmc_run_cfg* cfg, dxmc_tally* tally, const dxmc_run_monitor* monitor, dxmc_run_stats* stats)
{
    if (!tally || !c)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "run: NULL argument") : DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    int st = validateRun(c, w, t, exposures, nExp, cfg, tally);
    if (st)
        return st;
    DXMC_CHECK(c, hipSetDevice(c->device));
    dxmc_run_cfg run = *cfg;
    if (run.flags & DXMC_RUN_AUTOTUNE) {
        const auto key = tuneKey(w, t, run);
        auto it = c->tuned.find(key);
        if (it == c->tuned.end()) {
            uint64_t total = 0;
            for (uint64_t i = 0; i < nExp; ++i)
                total += exposures[i].n_histories;
            dxmc_run_cfg best {};
            const uint64_t trial = std::clamp<uint64_t>(total / 8, uint64_t(1) << 20, uint64_t(1) << 23);
            if ((st = autotune(c, w, t, exposures, nExp, run, trial, best, nullptr)))
                return st;
            it = c->tuned.emplace(key, best).first;
        }
        const dxmc_run_cfg& b = it->second;
        for (auto f : { &dxmc_run_cfg::block_size, &dxmc_run_cfg::blocks_per_cu, &dxmc_run_cfg::regeneration_threshold,
                 &dxmc_run_cfg::interaction_threshold, &dxmc_run_cfg::queue_depth, &dxmc_run_cfg::waves_per_simd, &dxmc_run_cfg::lds_tables,
                 &dxmc_run_cfg::block_bounds })
            if (run.*f == 0)
                run.*f = b.*f;
    }
    if (monitor && monitor->preview && (monitor->preview_axis < 0 || monitor->preview_axis > 2))
        return c->fail(DXMC_ERR_ARGUMENT, "run: preview axis must be 0, 1 or 2");
    return runTransport(c, w, t, exposures, nExp, run, tally, monitor, stats);
}

} // extern "C"

namespace dxmc_backend {
int validateRun(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* exposures, uint64_t nExp, const dxmc_run_cfg* cfg,
    const dxmc_tally* tally)
{
    if (!c || !w || !t || !exposures || !cfg)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "run: NULL argument") : DXMC_ERR_ARGUMENT;
    if (w->ctx != c || t->ctx != c || (tally && tally->ctx != c))
        return c->fail(DXMC_ERR_STATE, "run: objects belong to another context");
    if (tally && tally->n != w->n)
        return c->fail(DXMC_ERR_ARGUMENT, "run: tally size differs from the world size");
    if (cfg->precision == DXMC_PRECISION_FP64 && cfg->block_size > 256)
        return c->fail(DXMC_ERR_ARGUMENT, "run: FP64 block size must be at most 256");
    if (cfg->block_bounds > 2)
        return c->fail(DXMC_ERR_ARGUMENT, "run: block_bounds must be 0, 1 or 2");
    if (t->nMaterials < w->nMaterials)
        return c->fail(DXMC_ERR_ARGUMENT, "run: tables hold fewer materials than the world uses");
    if (cfg->low_energy_correction < 0 || cfg->low_energy_correction > 2 || (cfg->precision != DXMC_PRECISION_FP64 && cfg->precision != DXMC_PRECISION_FP32))
        return c->fail(DXMC_ERR_ARGUMENT, "run: bad precision or low energy correction");
    if (nExp == 0 || nExp > 0xfffffffeULL)
        return c->fail(DXMC_ERR_ARGUMENT, "run: need 1 .. 2^32-2 exposures");
    for (uint64_t i = 0; i < nExp; ++i) {
        const auto& e = exposures[i];
        if (e.spectrum >= static_cast<int32_t>(t->nSpectra) || e.heel >= static_cast<int32_t>(t->nHeel) || e.bowtie >= static_cast<int32_t>(t->nBowtie))
            return c->fail(DXMC_ERR_ARGUMENT, "run: exposure refers to a missing spectrum, heel or bowtie table");
    }
    return DXMC_OK;
}
}

extern "C" {

// ----------------------------------------------------------------- parity

int dxmc_hip_rng_streams(dxmc_ctx* c, uint64_t seed, uint64_t first, uint64_t n, uint32_t draws, uint32_t* out)
{
    if (!c || !out || n == 0 || draws == 0)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    DXMC_CHECK(c, hipSetDevice(c->device));
    DeviceBuffer buf(c);
    int st = buf.alloc(n * draws * sizeof(uint32_t));
    if (st)
        return st;
    hipLaunchKernelGGL(rngStreams, dim3(gridFor(n, 256)), dim3(256), 0, c->stream, seed, first, n, draws, static_cast<uint32_t*>(buf.p));
    DXMC_CHECK(c, hipGetLastError());
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    if (int e = c->download(out, buf.p, n * draws * sizeof(uint32_t)))
        return e;
    return DXMC_OK;
}

int dxmc_hip_lut_lookup(dxmc_ctx* c, const dxmc_tables* t, int32_t precision, uint64_t n, const uint32_t* mat, const double* energy, double* outAtt,
    double* outMaj)
{
    if (!c || !t || !mat || !energy || !outAtt || !outMaj || n == 0 || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    for (uint64_t i = 0; i < n; ++i)
        if (mat[i] >= t->nMaterials || !(energy[i] > 0))
            return c->fail(DXMC_ERR_ARGUMENT, "lut_lookup: material out of range or energy not positive");
    DXMC_CHECK(c, hipSetDevice(c->device));
    DeviceBuffer dm(c), de(c), da(c), dj(c);
    int st;
    if ((st = dm.upload(mat, n * 4)) || (st = de.upload(energy, n * 8)) || (st = da.alloc(n * 24)) || (st = dj.alloc(n * 8)))
        return st;
    if (precision == DXMC_PRECISION_FP64)
        hipLaunchKernelGGL(lutLookup<double>, dim3(gridFor(n, 256)), dim3(256), 0, c->stream, t->f64, n, static_cast<const uint32_t*>(dm.p),
            static_cast<const double*>(de.p), static_cast<double*>(da.p), static_cast<double*>(dj.p));
    else
        hipLaunchKernelGGL(lutLookup<float>, dim3(gridFor(n, 256)), dim3(256), 0, c->stream, t->f32, n, static_cast<const uint32_t*>(dm.p),
            static_cast<const double*>(de.p), static_cast<double*>(da.p), static_cast<double*>(dj.p));
    DXMC_CHECK(c, hipGetLastError());
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    if (int e = c->download(outAtt, da.p, n * 24))
        return e;
    if (int e = c->download(outMaj, dj.p, n * 8))
        return e;
    return DXMC_OK;
}

int dxmc_hip_interact(dxmc_ctx* c, const dxmc_tables* t, int32_t precision, int32_t L, int32_t kind, uint64_t n, const uint32_t* mat, double fp,
    uint64_t seed, uint64_t first, double* particles, double* energyOut, int32_t* aliveOut)
{
    if (!c || !t || !mat || !particles || !energyOut || !aliveOut || n == 0 || t->ctx != c || L < 0 || L > 2 || kind < 0 || kind > 4)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "interact: bad argument") : DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    for (uint64_t i = 0; i < n; ++i)
        if (mat[i] >= t->nMaterials || !(particles[i * 8 + 6] > 0))
            return c->fail(DXMC_ERR_ARGUMENT, "interact: material out of range or energy not positive");
    DXMC_CHECK(c, hipSetDevice(c->device));
    DeviceBuffer dm(c), dp(c), de(c), da(c), ts(c);
    int st;
    if ((st = dm.upload(mat, n * 4)) || (st = dp.upload(particles, n * 64)) || (st = de.alloc(n * 8)) || (st = da.alloc(n * 4))
        || (st = ts.zeros(n * sizeof(VoxelScore))))
        return st;
    Tally tl { static_cast<VoxelScore*>(ts.p) };
    if (precision == DXMC_PRECISION_FP64)
        launchInteract<double>(L, t->f64, kind, n, static_cast<const uint32_t*>(dm.p), fp, seed, first, static_cast<double*>(dp.p),
            static_cast<double*>(de.p), static_cast<int32_t*>(da.p), tl, c->stream);
    else
        launchInteract<float>(L, t->f32, kind, n, static_cast<const uint32_t*>(dm.p), fp, seed, first, static_cast<double*>(dp.p),
            static_cast<double*>(de.p), static_cast<int32_t*>(da.p), tl, c->stream);
    DXMC_CHECK(c, hipGetLastError());
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    if (int e = c->download(particles, dp.p, n * 64))
        return e;
    if (int e = c->download(aliveOut, da.p, n * 4))
        return e;
    // kinds 0-2 return the energy deposited, 3-4 the tally the dispatch scored into
    if (kind <= 2)
        return c->download(energyOut, de.p, n * 8);
    std::vector<VoxelScore> scores(n);
    if (int e = c->download(scores.data(), ts.p, n * sizeof(VoxelScore)))
        return e;
    for (uint64_t i = 0; i < n; ++i)
        energyOut[i] = scores[i].dose;
    return DXMC_OK;
}

int dxmc_hip_sample_particles(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, int32_t precision, const dxmc_exposure* x, uint64_t seed,
    uint64_t first, uint64_t n, double* particles, int32_t* inside)
{
    if (!c || !w || !t || !x || !particles || !inside || n == 0 || w->ctx != c || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    if (x->spectrum >= static_cast<int32_t>(t->nSpectra) || x->heel >= static_cast<int32_t>(t->nHeel) || x->bowtie >= static_cast<int32_t>(t->nBowtie))
        return c->fail(DXMC_ERR_ARGUMENT, "sample_particles: exposure refers to a missing table");
    DXMC_CHECK(c, hipSetDevice(c->device));
    DeviceBuffer dp(c), di(c);
    int st;
    if ((st = dp.alloc(n * 64)) || (st = di.alloc(n * 4)))
        return st;
    if (precision == DXMC_PRECISION_FP64)
        hipLaunchKernelGGL(sampleParticles<double>, dim3(gridFor(n, 256)), dim3(256), 0, c->stream, deviceWorld<double>(w), t->f64, toDevice<double>(*x), seed,
            first, n, static_cast<double*>(dp.p), static_cast<int32_t*>(di.p));
    else
        hipLaunchKernelGGL(sampleParticles<float>, dim3(gridFor(n, 256)), dim3(256), 0, c->stream, deviceWorld<float>(w), t->f32, toDevice<float>(*x), seed,
            first, n, static_cast<double*>(dp.p), static_cast<int32_t*>(di.p));
    DXMC_CHECK(c, hipGetLastError());
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    if (int e = c->download(particles, dp.p, n * 64))
        return e;
    if (int e = c->download(inside, di.p, n * 4))
        return e;
    return DXMC_OK;
}

} // extern "C"

// ------------------------------------------------------------ run driver

namespace dxmc_backend {

// exposures -> device records of both precisions + history prefix sums
int uploadExposures(dxmc_ctx* c, const dxmc_exposure* x, uint64_t nExp, RunBuffers& rb)
{
    std::vector<Exposure<double>> e64(nExp);
    std::vector<Exposure<float>> e32(nExp);
    std::vector<uint64_t> start(nExp + 1, 0);
    for (uint64_t i = 0; i < nExp; ++i) {
        e64[i] = toDevice<double>(x[i]);
        e32[i] = toDevice<float>(x[i]);
        start[i + 1] = start[i] + x[i].n_histories;
    }
    rb.totalHistories = start[nExp];
    rb.start = start;
    int st;
    if ((st = rb.exp64.upload(e64.data(), nExp * sizeof(Exposure<double>))) || (st = rb.exp32.upload(e32.data(), nExp * sizeof(Exposure<float>)))
        || (st = rb.startDev.upload(start.data(), (nExp + 1) * sizeof(uint64_t))) || (st = rb.counters.zeros(sizeof(Counters))))
        return st;
    return DXMC_OK;
}

int launchTallyMip(dxmc_ctx* c, const dxmc_tally* tally, const dxmc_world* w, int axis, float* image, uint64_t& width, uint64_t& height)
{
    const uint32_t nx = w->dims[0], ny = w->dims[1], nz = w->dims[2];
    width = axis == 0 ? ny : nx;
    height = axis == 2 ? ny : nz;
    const uint64_t threads = axis == 0 ? width * height * 64 : width * height;
    // normalised arrays when the tally was normalised since its last run,
    // else the dose sums
    const bool arrays = tally->arraysCurrent;
    const double* dose = arrays ? static_cast<const double*>(tally->dose) : &static_cast<const VoxelScore*>(tally->scores)->dose;
    hipLaunchKernelGGL(tallyMip, dim3(gridFor(threads, 256)), dim3(256), 0, c->stream, dose, arrays ? 1u : 4u, nx, ny, nz, axis, image);
    return c->check(hipGetLastError());
}

int launchNormalize(dxmc_ctx* c, const dxmc_tally* t, const dxmc_world* w, int mode, uint64_t nHist, double calibration)
{
    const double vv = w ? w->spacing[0] * w->spacing[1] * w->spacing[2] / 1000.0 : 0.0;
    hipLaunchKernelGGL(normalizeTally, dim3(gridFor(t->n, 256)), dim3(256), 0, c->stream, static_cast<const VoxelScore*>(t->scores),
        static_cast<double*>(t->dose), static_cast<uint32_t*>(t->nEvents), static_cast<double*>(t->variance), t->n, mode, nHist, vv, calibration,
        w ? static_cast<const double*>(w->density) : nullptr);
    if (int e = c->check(hipGetLastError()))
        return e;
    t->arraysCurrent = true;
    return DXMC_OK;
}

int runLaunches(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, uint64_t nExp, const dxmc_run_cfg& cfg, dxmc_tally* tally,
    const RunInputs& in, const dxmc_run_monitor* monitor, double& kernelMs, uint64_t& launches)
{
    const volatile int32_t* cancel = monitor ? monitor->cancel : nullptr;
    volatile uint64_t* exposuresDone = monitor ? monitor->exposures_done : nullptr;
    const bool preview = monitor && monitor->preview;
    DeviceBuffer image(c);
    std::vector<float> hostImage;
    auto lastPreview = std::chrono::steady_clock::now();
    // the live preview: projection of the tally so far, to the host, to the callback
    auto sendPreview = [&]() -> int {
        uint64_t width = 0, height = 0;
        const uint64_t pixels = uint64_t(w->dims[0]) * w->dims[1] * w->dims[2] / w->dims[monitor->preview_axis];
        if (!image.p)
            if (int e = image.alloc(pixels * sizeof(float)))
                return e;
        hostImage.resize(pixels);
        if (int e = launchTallyMip(c, tally, w, monitor->preview_axis, static_cast<float*>(image.p), width, height))
            return e;
        if (int e = c->download(hostImage.data(), image.p, pixels * sizeof(float)))
            return e;
        monitor->preview(monitor->preview_user, hostImage.data(), width, height);
        lastPreview = std::chrono::steady_clock::now();
        return DXMC_OK;
    };
    const Tally tl { static_cast<VoxelScore*>(tally->scores) };
    tally->arraysCurrent = false;
    uint64_t chunk = cfg.histories_per_launch;
    if (chunk == 0)
        chunk = cfg.precision == DXMC_PRECISION_FP64 ? (uint64_t(1) << 22) : (uint64_t(1) << 27);
    // the FP32 kernel keeps a history's offset in its launch in 32 bits
    if (cfg.precision != DXMC_PRECISION_FP64)
        chunk = std::min<uint64_t>(chunk, uint64_t(1) << 32);
    const uint32_t block = cfg.block_size ? cfg.block_size : 256;
    if (block % 64 != 0 || block > 256)
        return c->fail(DXMC_ERR_ARGUMENT, "run: block size must be a multiple of 64 up to 256");
    kernelMs = 0;
    launches = 0;
    uint64_t done = 0;
    const World<double> w64 = deviceWorld<double>(w);
    const World<float> w32 = deviceWorld<float>(w);
    for (uint64_t hb = 0; hb < in.total; hb += chunk) {
        if (cancel && __atomic_load_n(cancel, __ATOMIC_ACQUIRE)) {
            hipStreamSynchronize(c->stream);
            return c->fail(DXMC_ERR_CANCELLED, "run: cancelled");
        }
        const uint64_t he = std::min(in.total, hb + chunk);
        DXMC_CHECK(c, hipEventRecord(c->ev0, c->stream));
        if (cfg.precision == DXMC_PRECISION_FP64)
            launchPlainL<double>(cfg.low_energy_correction, w64, t->f64, in.exp64, in.startDev, static_cast<uint32_t>(nExp), hb, he, cfg.seed,
                cfg.first_history, tl, in.ctr, block, c->stream);
        else
            launchFp32(c, cfg, w32, t->f32, in.exp32, in.startDev, static_cast<uint32_t>(nExp), hb, he, tl, in.ctr, in.historyCounter);
        DXMC_CHECK(c, hipGetLastError());
        DXMC_CHECK(c, hipEventRecord(c->ev1, c->stream));
        DXMC_CHECK(c, hipEventSynchronize(c->ev1));
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        kernelMs += ms;
        ++launches;
        // exposures whose last history is done
        if (exposuresDone && in.startHost) {
            while (done < nExp && in.startHost[done + 1] <= he)
                ++done;
            __atomic_store_n(exposuresDone, done, __ATOMIC_RELEASE);
        }
        const bool last = he >= in.total;
        if (preview
            && (last || std::chrono::duration<double>(std::chrono::steady_clock::now() - lastPreview).count() >= monitor->preview_interval_seconds))
            if (int e = sendPreview())
                return e;
    }
    return DXMC_OK;
}

int runTransport(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg& cfg,
    dxmc_tally* tally, const dxmc_run_monitor* monitor, dxmc_run_stats* stats)
{
    DXMC_CHECK(c, hipSetDevice(c->device));
    RunBuffers rb(c);
    int st = uploadExposures(c, x, nExp, rb);
    if (st)
        return st;
    const bool counters = (cfg.flags & 1u) != 0 || stats;
    const RunInputs in { static_cast<const Exposure<double>*>(rb.exp64.p), static_cast<const Exposure<float>*>(rb.exp32.p),
        static_cast<const uint64_t*>(rb.startDev.p), rb.start.data(), rb.totalHistories, counters ? static_cast<Counters*>(rb.counters.p) : nullptr,
        static_cast<unsigned long long*>(c->historyCounter) };
    double kernelMs = 0;
    uint64_t launches = 0;
    if ((st = runLaunches(c, w, t, nExp, cfg, tally, in, monitor, kernelMs, launches)))
        return st;
    if (stats) {
        Counters h {};
        if (int e = c->download(&h, rb.counters.p, sizeof(Counters)))
            return e;
        stats->histories = in.total;
        stats->steps = h.steps;
        stats->events = h.events;
        stats->kernel_ms = kernelMs;
        stats->launches = launches;
    }
    return DXMC_OK;
}
}
