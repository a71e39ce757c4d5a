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
__global__ void normalizeTally(Tally tally, uint64_t n, int mode, uint64_t nHist, double voxelVolume, double calibration, const double* dens64)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (mode == DXMC_OUTPUT_EV_PER_HISTORY) {
        const double hInv = 1.0 / static_cast<double>(nHist - 1);
        const double hd = 1e3 / static_cast<double>(nHist);
        const double hv = 1e6 / static_cast<double>(nHist);
        const double d = tally.dose[i] * hd;
        tally.dose[i] = d;
        tally.variance[i] = (tally.variance[i] * hv - d * d) * hInv;
    } else {
        const double de = dens64[i];
        const double mass = de * voxelVolume * 0.001;
        const double factor = calibration / mass;
        tally.dose[i] = de > 0.0 ? tally.dose[i] * factor : 0.0;
        tally.variance[i] = de > 0.0 ? tally.variance[i] * factor * factor : 0.0;
    }
}

__global__ void fillZero(double* a, uint32_t* b, double* c, uint64_t n)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = 0;
        b[i] = 0;
        c[i] = 0;
    }
}

__global__ void packVoxels(const uint8_t* mat, const double* dens, const uint8_t* meas, uint2* vox, uint64_t n)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        uint2 v;
        v.x = __float_as_uint(static_cast<float>(dens[i]));
        v.y = static_cast<uint32_t>(mat[i]) | (meas ? static_cast<uint32_t>(meas[i]) << 8 : 0u);
        vox[i] = v;
    }
}

// Distinct FP32 voxel records (density bits, material | measurement << 8)
// in order of first appearance, and the palette index of every voxel.
// Returns false (palette unusable) beyond 65536 distinct records.
bool buildPalette(const uint8_t* mat, const double* dens, const uint8_t* meas, uint64_t n, std::vector<uint2>& palette,
    std::vector<uint16_t>& index)
{
    constexpr uint32_t limit = 65536;
    constexpr uint32_t capacity = 1u << 18; // open addressing, load <= 1/4
    std::vector<uint64_t> keys(capacity, ~0ULL);
    std::vector<uint32_t> slots(capacity, 0);
    palette.clear();
    index.resize(n);
    uint64_t lastKey = ~0ULL;
    uint16_t lastIndex = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t bits = std::bit_cast<uint32_t>(static_cast<float>(dens[i]));
        const uint32_t tag = static_cast<uint32_t>(mat[i]) | (meas ? static_cast<uint32_t>(meas[i]) << 8 : 0u);
        const uint64_t key = static_cast<uint64_t>(tag) << 32 | bits;
        if (key != lastKey) {
            uint32_t h = static_cast<uint32_t>((key * 0x9E3779B97F4A7C15ULL) >> 46);
            while (keys[h] != key && keys[h] != ~0ULL)
                h = (h + 1) & (capacity - 1);
            if (keys[h] == ~0ULL) {
                if (palette.size() == limit) {
                    palette.clear();
                    index.clear();
                    return false;
                }
                keys[h] = key;
                slots[h] = static_cast<uint32_t>(palette.size());
                palette.push_back(make_uint2(bits, tag));
            }
            lastKey = key;
            lastIndex = static_cast<uint16_t>(slots[h]);
        }
        index[i] = lastIndex;
    }
    return true;
}

uint32_t gridFor(uint64_t n, uint32_t block) { return static_cast<uint32_t>((n + block - 1) / block); }

template <typename T>
Exposure<T> toDevice(const dxmc_exposure& e)
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
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipFree(c->historyCounter);
    hipStreamDestroy(c->stream);
    delete c;
    return DXMC_OK;
}

const char* dxmc_hip_last_error(const dxmc_ctx* c) { return c ? c->error.c_str() : globalError(); }

int dxmc_hip_mem_info(dxmc_ctx* c, uint64_t* freeBytes, uint64_t* totalBytes)
{
    if (!c)
        return DXMC_ERR_ARGUMENT;
    size_t f = 0, t = 0;
    DXMC_CHECK(c, hipSetDevice(c->device));
    DXMC_CHECK(c, hipMemGetInfo(&f, &t));
    if (freeBytes)
        *freeBytes = f;
    if (totalBytes)
        *totalBytes = t;
    return DXMC_OK;
}

int dxmc_hip_synchronize(dxmc_ctx* c)
{
    if (!c)
        return DXMC_ERR_ARGUMENT;
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    return DXMC_OK;
}

// ------------------------------------------------------------------ world

int dxmc_hip_world_upload(dxmc_ctx* c, const dxmc_world_desc* d, dxmc_world** out)
{
    if (!c || !d || !out)
        return DXMC_ERR_ARGUMENT;
    *out = nullptr;
    const uint64_t n = d->dims[0] * d->dims[1] * d->dims[2];
    if (n == 0 || !d->material || (!d->density_f32 == !d->density_f64) || d->n_materials == 0 || d->n_materials > 256)
        return c->fail(DXMC_ERR_ARGUMENT, "world_upload: need dims, material, exactly one density array and 1..256 materials");
    if (d->dims[0] > 0xffffffffULL || d->dims[1] > 0xffffffffULL || d->dims[2] > 0xffffffffULL)
        return c->fail(DXMC_ERR_ARGUMENT, "world_upload: dimension exceeds 2^32-1");
    for (int i = 0; i < 3; ++i)
        if (!(d->spacing[i] > 0) || !(d->extent_safe[2 * i] < d->extent_safe[2 * i + 1]))
            return c->fail(DXMC_ERR_ARGUMENT, "world_upload: invalid spacing or extent");
    // material indices must address the material table (reference World::validate)
    uint8_t maxMat = 0;
    for (uint64_t i = 0; i < n; ++i)
        maxMat = std::max(maxMat, d->material[i]);
    if (maxMat >= d->n_materials)
        return c->fail(DXMC_ERR_ARGUMENT, "world_upload: material index outside the material map");
    DXMC_CHECK(c, hipSetDevice(c->device));
    auto* w = new dxmc_world();
    w->ctx = c;
    w->n = n;
    for (int i = 0; i < 3; ++i) {
        w->dims[i] = static_cast<uint32_t>(d->dims[i]);
        w->spacing[i] = d->spacing[i];
    }
    for (int i = 0; i < 6; ++i)
        w->extent[i] = d->extent_safe[i];
    w->nMaterials = d->n_materials;
    w->hasMeasurement = false;
    if (d->measurement)
        for (uint64_t i = 0; i < n && !w->hasMeasurement; ++i)
            w->hasMeasurement = d->measurement[i] != 0;
    std::vector<double> dens64;
    const double* densSrc = d->density_f64;
    if (!densSrc) {
        dens64.resize(n);
        for (uint64_t i = 0; i < n; ++i)
            dens64[i] = static_cast<double>(d->density_f32[i]);
        densSrc = dens64.data();
    }
    w->maxDensity.assign(d->n_materials, 0.0);
    for (uint64_t i = 0; i < n; ++i)
        w->maxDensity[d->material[i]] = std::max(w->maxDensity[d->material[i]], densSrc[i]);
    // FP32 voxel records: palette of the distinct (density, material,
    // measurement) records and a 1- or 2-byte index per voxel when there
    // are at most 65536 of them (exact: the palette holds the same FP32
    // record the voxel would)
    std::vector<uint2> palette;
    std::vector<uint16_t> index;
    const bool indexed = buildPalette(d->material, densSrc, w->hasMeasurement ? d->measurement : nullptr, n, palette, index);
    w->voxelBytes = !indexed ? 8u : palette.size() <= 256 ? 1u : 2u;
    w->nPalette = indexed ? static_cast<uint32_t>(palette.size()) : 0u;
    int st = DXMC_OK;
    if ((st = c->alloc(&w->material, n)) || (st = c->alloc(&w->density, n * sizeof(double))) || (st = c->alloc(&w->voxels, n * w->voxelBytes))
        || (indexed && (st = c->alloc(&w->palette, palette.size() * sizeof(uint2)))) || (w->hasMeasurement && (st = c->alloc(&w->measurement, n)))) {
        dxmc_hip_world_destroy(c, w);
        return st;
    }
    if ((st = c->copy(w->material, d->material, n)) || (st = c->copy(w->density, densSrc, n * sizeof(double)))
        || (w->hasMeasurement && (st = c->copy(w->measurement, d->measurement, n)))
        || (indexed && (st = c->copy(w->palette, palette.data(), palette.size() * sizeof(uint2))))) {
        dxmc_hip_world_destroy(c, w);
        return st;
    }
    if (w->voxelBytes == 2) {
        st = c->copy(w->voxels, index.data(), n * sizeof(uint16_t));
    } else if (w->voxelBytes == 1) {
        std::vector<uint8_t> narrow(index.begin(), index.end());
        st = c->copy(w->voxels, narrow.data(), n);
    } else {
        hipLaunchKernelGGL(packVoxels, dim3(gridFor(n, 256)), dim3(256), 0, c->stream, static_cast<const uint8_t*>(w->material),
            static_cast<const double*>(w->density), static_cast<const uint8_t*>(w->measurement), static_cast<uint2*>(w->voxels), n);
        st = c->check(hipGetLastError());
    }
    if (st || c->check(hipStreamSynchronize(c->stream))) {
        dxmc_hip_world_destroy(c, w);
        return st ? st : DXMC_ERR_DEVICE;
    }
    w->bytes = n * (1 + sizeof(double) + w->voxelBytes + (w->hasMeasurement ? 1 : 0)) + w->nPalette * sizeof(uint2);
    *out = w;
    return DXMC_OK;
}

int dxmc_hip_world_destroy(dxmc_ctx* c, dxmc_world* w)
{
    if (!c || !w)
        return DXMC_ERR_ARGUMENT;
    if (w->ctx != c)
        return c->fail(DXMC_ERR_STATE, "world_destroy: world belongs to another context");
    std::erase_if(c->tuned, [w](const auto& kv) { return std::get<0>(kv.first) == w; });
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    c->release(w->material);
    c->release(w->density);
    c->release(w->measurement);
    c->release(w->voxels);
    c->release(w->palette);
    hipStreamSynchronize(c->stream);
    delete w;
    return DXMC_OK;
}

int dxmc_hip_world_bytes(const dxmc_world* w, uint64_t* bytes)
{
    if (!w || !bytes)
        return DXMC_ERR_ARGUMENT;
    *bytes = w->bytes;
    return DXMC_OK;
}

// ----------------------------------------------------------------- tables

int dxmc_hip_tables_upload(dxmc_ctx* c, const dxmc_tables_desc* d, dxmc_tables** out)
{
    if (!c || !d || !out)
        return DXMC_ERR_ARGUMENT;
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
    *out = nullptr;
    DXMC_CHECK(c, hipSetDevice(c->device));
    auto* t = new dxmc_tally();
    t->ctx = c;
    t->n = n;
    int st;
    if ((st = c->alloc(&t->dose, n * sizeof(double))) || (st = c->alloc(&t->nEvents, n * sizeof(uint32_t))) || (st = c->alloc(&t->variance, n * sizeof(double)))) {
        dxmc_hip_tally_destroy(c, t);
        return st;
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
    if (t->ctx != c)
        return c->fail(DXMC_ERR_STATE, "tally_destroy: tally belongs to another context");
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
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
    DXMC_CHECK(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(fillZero, dim3(gridFor(t->n, 256)), dim3(256), 0, c->stream, static_cast<double*>(t->dose), static_cast<uint32_t*>(t->nEvents),
        static_cast<double*>(t->variance), t->n);
    DXMC_CHECK(c, hipGetLastError());
    return DXMC_OK;
}

int dxmc_hip_tally_download(dxmc_ctx* c, const dxmc_tally* t, double* dose, uint32_t* nEvents, double* variance)
{
    if (!c || !t || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    DXMC_CHECK(c, hipSetDevice(c->device));
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    if (dose)
        DXMC_CHECK(c, hipMemcpyAsync(dose, t->dose, t->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (nEvents)
        DXMC_CHECK(c, hipMemcpyAsync(nEvents, t->nEvents, t->n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (variance)
        DXMC_CHECK(c, hipMemcpyAsync(variance, t->variance, t->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    return DXMC_OK;
}

int dxmc_hip_tally_normalize(dxmc_ctx* c, dxmc_tally* t, const dxmc_world* w, int mode, uint64_t nHist, double calibration)
{
    if (!c || !t || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
    if (mode != DXMC_OUTPUT_EV_PER_HISTORY && mode != DXMC_OUTPUT_DOSE)
        return c->fail(DXMC_ERR_ARGUMENT, "tally_normalize: unknown output mode");
    if (mode == DXMC_OUTPUT_DOSE && (!w || w->n != t->n))
        return c->fail(DXMC_ERR_ARGUMENT, "tally_normalize: dose needs the world the tally was scored on");
    if (mode == DXMC_OUTPUT_EV_PER_HISTORY && nHist < 2)
        return c->fail(DXMC_ERR_ARGUMENT, "tally_normalize: need at least two histories");
    DXMC_CHECK(c, hipSetDevice(c->device));
    const double vv = w ? w->spacing[0] * w->spacing[1] * w->spacing[2] / 1000.0 : 0.0;
    Tally tl { static_cast<double*>(t->dose), static_cast<uint32_t*>(t->nEvents), static_cast<double*>(t->variance) };
    hipLaunchKernelGGL(normalizeTally, dim3(gridFor(t->n, 256)), dim3(256), 0, c->stream, tl, t->n, mode, nHist, vv, calibration,
        w ? static_cast<const double*>(w->density) : nullptr);
    DXMC_CHECK(c, hipGetLastError());
    return DXMC_OK;
}

// -------------------------------------------------------------- transport

int dxmc_hip_run(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* exposures, uint64_t nExp, const dxmc_run_cfg* cfg,
    dxmc_tally* tally, const volatile int32_t* cancel, volatile uint64_t* exposuresDone, dxmc_run_stats* stats)
{
    if (!tally)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "run: NULL argument") : DXMC_ERR_ARGUMENT;
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
                 &dxmc_run_cfg::interaction_threshold, &dxmc_run_cfg::queue_depth, &dxmc_run_cfg::waves_per_simd, &dxmc_run_cfg::lds_tables })
            if (run.*f == 0)
                run.*f = b.*f;
    }
    return runTransport(c, w, t, exposures, nExp, run, tally, cancel, exposuresDone, stats);
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
    for (uint64_t i = 0; i < n; ++i)
        if (mat[i] >= t->nMaterials || !(particles[i * 8 + 6] > 0))
            return c->fail(DXMC_ERR_ARGUMENT, "interact: material out of range or energy not positive");
    DXMC_CHECK(c, hipSetDevice(c->device));
    DeviceBuffer dm(c), dp(c), de(c), da(c), td(c), tn(c), tv(c);
    int st;
    if ((st = dm.upload(mat, n * 4)) || (st = dp.upload(particles, n * 64)) || (st = de.alloc(n * 8)) || (st = da.alloc(n * 4)) || (st = td.zeros(n * 8))
        || (st = tn.zeros(n * 4)) || (st = tv.zeros(n * 8)))
        return st;
    Tally tl { static_cast<double*>(td.p), static_cast<uint32_t*>(tn.p), static_cast<double*>(tv.p) };
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
    return c->download(energyOut, kind <= 2 ? de.p : td.p, n * 8);
}

int dxmc_hip_sample_particles(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, int32_t precision, const dxmc_exposure* x, uint64_t seed,
    uint64_t first, uint64_t n, double* particles, int32_t* inside)
{
    if (!c || !w || !t || !x || !particles || !inside || n == 0 || w->ctx != c || t->ctx != c)
        return DXMC_ERR_ARGUMENT;
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

int runTransport(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg& cfg,
    dxmc_tally* tally, const volatile int32_t* cancel, volatile uint64_t* exposuresDone, dxmc_run_stats* stats)
{
    DXMC_CHECK(c, hipSetDevice(c->device));
    RunBuffers rb(c);
    int st = uploadExposures(c, x, nExp, rb);
    if (st)
        return st;
    const bool counters = (cfg.flags & 1u) != 0 || stats;
    Counters* ctr = counters ? static_cast<Counters*>(rb.counters.p) : nullptr;
    const Tally tl { static_cast<double*>(tally->dose), static_cast<uint32_t*>(tally->nEvents), static_cast<double*>(tally->variance) };
    const uint64_t total = rb.totalHistories;
    uint64_t chunk = cfg.histories_per_launch;
    if (chunk == 0)
        chunk = cfg.precision == DXMC_PRECISION_FP64 ? (uint64_t(1) << 22) : (uint64_t(1) << 27);
    const uint32_t block = cfg.block_size ? cfg.block_size : 256;
    if (block % 64 != 0 || block > 256)
        return c->fail(DXMC_ERR_ARGUMENT, "run: block size must be a multiple of 64 up to 256");
    double kernelMs = 0;
    uint64_t launches = 0;
    uint64_t done = 0;
    const World<double> w64 = deviceWorld<double>(w);
    const World<float> w32 = deviceWorld<float>(w);
    for (uint64_t hb = 0; hb < total; hb += chunk) {
        if (cancel && *cancel) {
            hipStreamSynchronize(c->stream);
            return c->fail(DXMC_ERR_CANCELLED, "run: cancelled");
        }
        const uint64_t he = std::min(total, hb + chunk);
        DXMC_CHECK(c, hipEventRecord(c->ev0, c->stream));
        if (cfg.precision == DXMC_PRECISION_FP64)
            launchPlainL<double>(cfg.low_energy_correction, w64, t->f64, static_cast<const Exposure<double>*>(rb.exp64.p),
                static_cast<const uint64_t*>(rb.startDev.p), static_cast<uint32_t>(nExp), hb, he, cfg.seed, cfg.first_history, tl, ctr, block,
                c->stream);
        else
            launchFp32(c, cfg, w32, t->f32, static_cast<const Exposure<float>*>(rb.exp32.p), static_cast<const uint64_t*>(rb.startDev.p),
                static_cast<uint32_t>(nExp), hb, he, tl, ctr);
        DXMC_CHECK(c, hipGetLastError());
        DXMC_CHECK(c, hipEventRecord(c->ev1, c->stream));
        DXMC_CHECK(c, hipEventSynchronize(c->ev1));
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        kernelMs += ms;
        ++launches;
        // exposures whose last history is done
        while (done < nExp && rb.start[done + 1] <= he)
            ++done;
        if (exposuresDone)
            *exposuresDone = done;
    }
    if (stats) {
        Counters h {};
        if (int e = c->download(&h, rb.counters.p, sizeof(Counters)))
        return e;
        stats->histories = total;
        stats->steps = h.steps;
        stats->events = h.events;
        stats->kernel_ms = kernelMs;
        stats->launches = launches;
    }
    return DXMC_OK;
}
}
