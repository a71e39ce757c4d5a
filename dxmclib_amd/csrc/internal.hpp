// Internal definitions shared by the backend's translation units: the opaque
// C ABI objects, device buffers, error handling and the table builder.
#pragma once
#include "device/physics.hpp"
#include "dxmc_hip.h"
#include "hostcopy.hpp"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <cmath>
#include <type_traits>
#include <vector>

namespace dxmc_backend {
// page-locked blocks from dxmc_hip_host_alloc; pinnedHost: whether
// [p, p + bytes) lies in one of them
void registerPinned(const void* p, size_t bytes);
bool unregisterPinned(const void* p);
bool pinnedHost(const void* p, size_t bytes);
}

struct dxmc_ctx;

namespace dxmc_backend {
// the calling thread's last failure message per context: with several
// threads on one context, each reads its own error
std::map<const dxmc_ctx*, std::string>& threadErrors();
}

struct dxmc_ctx {
    int device = 0;
    int numCU = 256;
    std::string arch;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* historyCounter = nullptr; // u64 work counter of the persistent kernel
    void* scratch = nullptr; // device address of the no-allocation run's scratch array (null: unavailable)
    // autotuner results per (world, tables, precision, low energy correction)
    std::map<std::tuple<const void*, const void*, int, int>, dxmc_run_cfg> tuned;
    // Held by every entry point for the whole call: the staging chunks, the
    // work counter, the autotuner map and the stream's order of a call's
    // steps are per context, so calls from several threads on one context
    // run one after the other (device work on the stream is ordered anyway).
    // Recursive: entry points call one another.
    std::recursive_mutex lock;

    int fail(int status, const std::string& msg)
    {
        dxmc_backend::threadErrors()[this] = msg;
        return status;
    }
    // records a HIP error, returns non-zero on failure
    int check(hipError_t e)
    {
        if (e == hipSuccess)
            return 0;
        return fail(DXMC_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
    }
    // Device memory is stream-ordered (hipMallocAsync / hipFreeAsync on the
    // context's stream, from the device's pool, which keeps freed blocks):
    // unlike hipFree, releasing memory then does not wait for the kernels
    // other contexts run on the same device.
    template <typename P>
    int alloc(P** p, size_t bytes)
    {
        void* q = nullptr;
        const hipError_t e = hipMallocAsync(&q, bytes ? bytes : 1, stream);
        if (e != hipSuccess) {
            hipGetLastError();
            return fail(DXMC_ERR_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
        *p = static_cast<P*>(q);
        return DXMC_OK;
    }
    void release(void* p)
    {
        if (p)
            hipFreeAsync(p, stream);
    }
    // Large host transfers go through two pinned chunks, so the DMA of one
    // chunk overlaps the host copy of the other; from pageable memory the
    // runtime's own staging reaches a fraction of the link rate. The chunks
    // are allocated at the context's first large transfer.
    static constexpr size_t STAGE_CHUNK = size_t { 8 } << 20;
    static constexpr size_t STAGE_MIN = size_t { 1 } << 20;
    void* stage[2] = { nullptr, nullptr };
    hipEvent_t stageDone[2] = { nullptr, nullptr };
    // copy a chunk into or out of the staging memory on several host threads
    dxmc_backend::CopyHelpers helpers;

    bool staging()
    {
        if (stage[0])
            return true;
        for (int i = 0; i < 2; ++i)
            if (hipHostMalloc(&stage[i], STAGE_CHUNK, hipHostMallocDefault) != hipSuccess
                || hipEventCreateWithFlags(&stageDone[i], hipEventDisableTiming) != hipSuccess) {
                hipGetLastError();
                releaseStaging();
                return false;
            }
        helpers.start(); // without helpers the copies run on the calling thread
        return true;
    }
    void releaseStaging()
    {
        helpers.stop();
        for (int i = 0; i < 2; ++i) {
            if (stage[i])
                hipHostFree(stage[i]);
            if (stageDone[i])
                hipEventDestroy(stageDone[i]);
            stage[i] = nullptr;
            stageDone[i] = nullptr;
        }
    }
    // host -> device, complete on return (the host array may be freed)
    int copy(void* dst, const void* src, size_t bytes)
    {
        if (bytes == 0)
            return DXMC_OK;
        if (bytes < STAGE_MIN || dxmc_backend::pinnedHost(src, bytes) || !staging()) {
            if (int st = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream)))
                return st;
            return check(hipStreamSynchronize(stream));
        }
        const auto* from = static_cast<const char*>(src);
        auto* to = static_cast<char*>(dst);
        for (size_t off = 0, i = 0; off < bytes; off += STAGE_CHUNK, ++i) {
            const int b = static_cast<int>(i & 1);
            const size_t len = std::min(STAGE_CHUNK, bytes - off);
            // the chunk's previous DMA (two chunks back) has to be done
            if (i >= 2)
                if (int st = check(hipEventSynchronize(stageDone[b])))
                    return st;
            helpers.copy(stage[b], from + off, len);
            if (int st = check(hipMemcpyAsync(to + off, stage[b], len, hipMemcpyHostToDevice, stream)))
                return st;
            if (int st = check(hipEventRecord(stageDone[b], stream)))
                return st;
        }
        return check(hipStreamSynchronize(stream));
    }
    // device -> host, complete on return
    int download(void* dst, const void* src, size_t bytes)
    {
        if (bytes == 0)
            return DXMC_OK;
        if (bytes < STAGE_MIN || dxmc_backend::pinnedHost(dst, bytes) || !staging()) {
            if (int st = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream)))
                return st;
            return check(hipStreamSynchronize(stream));
        }
        const auto* from = static_cast<const char*>(src);
        auto* to = static_cast<char*>(dst);
        size_t pendingOff[2] = { 0, 0 }, pendingLen[2] = { 0, 0 };
        auto drain = [&](int b) {
            if (int st = check(hipEventSynchronize(stageDone[b])))
                return st;
            helpers.copy(to + pendingOff[b], stage[b], pendingLen[b]);
            pendingLen[b] = 0;
            return 0;
        };
        for (size_t off = 0, i = 0; off < bytes; off += STAGE_CHUNK, ++i) {
            const int b = static_cast<int>(i & 1);
            const size_t len = std::min(STAGE_CHUNK, bytes - off);
            if (int st = check(hipMemcpyAsync(stage[b], from + off, len, hipMemcpyDeviceToHost, stream)))
                return st;
            if (int st = check(hipEventRecord(stageDone[b], stream)))
                return st;
            pendingOff[b] = off;
            pendingLen[b] = len;
            // copy out the previous chunk while this one is in flight
            if (pendingLen[1 - b])
                if (int st = drain(1 - b))
                    return st;
        }
        for (int b = 0; b < 2; ++b)
            if (pendingLen[b])
                if (int st = drain(b))
                    return st;
        return DXMC_OK;
    }
};

// serialises the calls on one context (see dxmc_ctx::lock)
#define DXMC_LOCK(ctx) std::lock_guard<std::recursive_mutex> dxmc_lock_((ctx)->lock)

#define DXMC_CHECK(ctx, expr)                      \
    do {                                           \
        if (int dxmc_st_ = (ctx)->check((expr))) \
            return dxmc_st_;                       \
    } while (0)

struct dxmc_world {
    dxmc_ctx* ctx = nullptr;
    uint64_t n = 0;
    uint32_t dims[3] = { 0, 0, 0 };
    double spacing[3] = { 0, 0, 0 };
    double extent[6] = { 0, 0, 0, 0, 0, 0 };
    uint32_t nMaterials = 0;
    bool hasMeasurement = false;
    std::vector<double> maxDensity; // per material, host copy
    void* material = nullptr; // u8[n]
    void* density = nullptr; // f64[n]
    void* measurement = nullptr; // u8[n] or null
    void* voxels = nullptr; // FP32 path: palette index (u8/u16) or uint2 record per voxel
    void* palette = nullptr; // uint2[nPalette] FP32 records, null for 8-byte voxels
    uint32_t voxelBytes = 8;
    uint32_t nPalette = 0;
    // 8-byte-record worlds: bounded-density code per voxel (u8/u16) and the
    // interval of each code (DensityBound[nCodes]); codeBytes 0: none
    void* codes = nullptr;
    void* bounds = nullptr;
    uint32_t codeBytes = 0;
    uint32_t nCodes = 0;
    // 1-byte code worlds: a 2-byte code per block of 8x8x8 voxels (smallest
    // | largest code << 8 of a one-material block, else BLOCK_MIXED)
    void* blocks = nullptr;
    uint32_t blockDims[3] = { 0, 0, 0 };
    uint32_t blockShift = 0; // block edge 2^blockShift voxels
    uint32_t codeMul = 0; // 1-byte codes: (code * codeMul) >> CODE_MUL_SHIFT is the code's material
    uint64_t bytes = 0;
};

struct dxmc_tables {
    dxmc_ctx* ctx = nullptr;
    uint32_t nMaterials = 0, nSpectra = 0, nHeel = 0, nBowtie = 0;
    dxmc_dev::Tables<double> f64 {};
    dxmc_dev::Tables<float> f32 {};
    std::vector<void*> allocations;
};

struct dxmc_tally {
    dxmc_ctx* ctx = nullptr;
    uint64_t n = 0;
    // the transport accumulates into scores (VoxelScore[n]); normalisation
    // and download fill the per-voxel arrays from them
    void* scores = nullptr;
    void* dose = nullptr;
    void* nEvents = nullptr;
    void* variance = nullptr;
    // the arrays hold the scores (raw or normalised) since the last run or
    // reset
    mutable bool arraysCurrent = false;
};

namespace dxmc_backend {

void setGlobalError(const std::string& s);
const char* globalError();

// RAII device buffer for temporaries
struct DeviceBuffer {
    dxmc_ctx* c;
    void* p = nullptr;
    explicit DeviceBuffer(dxmc_ctx* ctx)
        : c(ctx)
    {
    }
    ~DeviceBuffer() { c->release(p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    int alloc(size_t bytes)
    {
        c->release(p);
        p = nullptr;
        return c->alloc(&p, bytes);
    }
    int upload(const void* src, size_t bytes)
    {
        if (int st = alloc(bytes))
            return st;
        return c->copy(p, src, bytes);
    }
    int zeros(size_t bytes)
    {
        if (int st = alloc(bytes))
            return st;
        return c->check(hipMemsetAsync(p, 0, bytes, c->stream));
    }
};

struct RunBuffers {
    DeviceBuffer exp64, exp32, startDev, counters;
    std::vector<uint64_t> start;
    uint64_t totalHistories = 0;
    explicit RunBuffers(dxmc_ctx* c)
        : exp64(c)
        , exp32(c)
        , startDev(c)
        , counters(c)
    {
    }
};

template <typename T>
inline dxmc_dev::World<T> deviceWorld(const dxmc_world* w)
{
    dxmc_dev::World<T> d {};
    d.nx = w->dims[0];
    d.ny = w->dims[1];
    d.nz = w->dims[2];
    for (int i = 0; i < 3; ++i)
        d.spacing[i] = static_cast<T>(w->spacing[i]);
    if constexpr (std::is_same_v<T, double>) {
        for (int i = 0; i < 6; ++i)
            d.extent[i] = w->extent[i];
    } else {
        // the safe extent one ulp inside the grid, formed in T from the grid
        // bounds as World<T>::updateMatrixExtent does (world.hpp:55-67), so
        // that a position on a voxel face indexes the voxel below it
        for (int i = 0; i < 3; ++i) {
            const T lower = static_cast<T>(std::nextafter(w->extent[2 * i], -HUGE_VAL));
            const T upper = static_cast<T>(std::nextafter(w->extent[2 * i + 1], HUGE_VAL));
            d.extent[2 * i] = std::nextafter(lower, upper);
            d.extent[2 * i + 1] = std::nextafter(upper, lower);
        }
    }
    d.material = static_cast<const uint8_t*>(w->material);
    d.density = static_cast<const double*>(w->density);
    d.measurement = static_cast<const uint8_t*>(w->measurement);
    d.voxels = w->voxels;
    d.palette = static_cast<const uint2*>(w->palette);
    d.codes = w->codes;
    d.bounds = static_cast<const dxmc_dev::DensityBound*>(w->bounds);
    d.codeBytes = w->codeBytes;
    d.index24 = uint64_t(d.nx) * d.ny < (uint64_t(1) << 24) && d.nz < (uint32_t(1) << 24) && uint64_t(d.nx) * d.ny * d.nz < (uint64_t(1) << 32);
    d.nxy = d.index24 ? d.nx * d.ny : 0;
    d.blocks = static_cast<const uint16_t*>(w->blocks);
    d.bnx = w->blockDims[0];
    d.bnxy = w->blockDims[0] * w->blockDims[1];
    d.blockShift = w->blockShift;
    d.codeMul = w->codeMul;
    d.voxelBytes = w->voxelBytes;
    return d;
}

// block reduction of the step / event counters into the global counters
__device__ inline void accumulateCounters(dxmc_dev::Counters* ctr, uint32_t steps, uint32_t events)
{
    __shared__ unsigned long long s[2];
    if (threadIdx.x == 0) {
        s[0] = 0;
        s[1] = 0;
    }
    __syncthreads();
    if (steps)
        atomicAdd(&s[0], static_cast<unsigned long long>(steps));
    if (events)
        atomicAdd(&s[1], static_cast<unsigned long long>(events));
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s[0])
            atomicAdd(&ctr->steps, s[0]);
        if (s[1])
            atomicAdd(&ctr->events, s[1]);
    }
}

bool validateTables(const dxmc_tables_desc& d, std::string& why);
int buildDeviceTables(dxmc_ctx* c, const dxmc_tables_desc& d, dxmc_tables& t);
int uploadExposures(dxmc_ctx* c, const dxmc_exposure* x, uint64_t nExp, RunBuffers& rb);
// argument checks shared by dxmc_hip_run and dxmc_hip_autotune (tally may be null)
int validateRun(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg* cfg,
    const dxmc_tally* tally);
inline std::tuple<const void*, const void*, int, int> tuneKey(const dxmc_world* w, const dxmc_tables* t, const dxmc_run_cfg& cfg)
{
    return { w, t, cfg.precision, cfg.low_energy_correction };
}
// autotune.hip
int autotune(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg& base,
    uint64_t trialHistories, dxmc_run_cfg& best, dxmc_tune_report* report);
// resident blocks per CU of the FP32 kernel (transport_fp32.hip)
// resident blocks per CU of the FP32 kernel for a world kind (0: any world)
uint32_t residentBlocksFp32(int L, uint32_t wavesPerSimd, uint32_t ldsPerBlock, uint32_t worldKind = 0);
// the resident waves per SIMD a run uses (5, 6 or 7) for a requested value
// (0: the default) and world kind
uint32_t wavesFp32(uint32_t requested, int L, uint32_t worldKind);
int runTransport(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg& cfg,
    dxmc_tally* tally, const dxmc_run_monitor* monitor, dxmc_run_stats* stats);
// device inputs of a run's launches: exposure records of both precisions,
// history prefix sums (device; host copy for progress, may be null), total
// histories, step/event counters (may be null) and the FP32 kernel's work
// counter
struct RunInputs {
    const dxmc_dev::Exposure<double>* exp64;
    const dxmc_dev::Exposure<float>* exp32;
    const uint64_t* startDev;
    const uint64_t* startHost;
    uint64_t total;
    dxmc_dev::Counters* ctr;
    unsigned long long* historyCounter;
};
// the launch loop of a run on prepared inputs (cancel, progress and preview
// through the monitor, which may be null)
int runLaunches(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, uint64_t nExp, const dxmc_run_cfg& cfg, dxmc_tally* tally,
    const RunInputs& in, const dxmc_run_monitor* monitor, double& kernelMs, uint64_t& launches);
// maximum intensity projection of the tally's dose along axis into a device
// image (launch only, on the context's stream); width and height returned
int launchTallyMip(dxmc_ctx* c, const dxmc_tally* tally, const dxmc_world* w, int axis, float* image, uint64_t& width, uint64_t& height);
// fills the tally's result arrays from its voxel scores: normalised for an
// output mode, the sums as they stand for mode -1
int launchNormalize(dxmc_ctx* c, const dxmc_tally* t, const dxmc_world* w, int mode, uint64_t nHist, double calibration);
// device address of the no-allocation run's scratch array on the current
// device (noalloc.hip), taken once per device
int noallocScratch(int device, void** out);
// FP32 transport launch (transport_fp32.hip)
void launchFp32(dxmc_ctx* c, const dxmc_run_cfg& cfg, const dxmc_dev::World<float>& w, const dxmc_dev::Tables<float>& t,
    const dxmc_dev::Exposure<float>* exps, const uint64_t* expStart, uint32_t nExp, uint64_t hb, uint64_t he, const dxmc_dev::Tally& tally,
    dxmc_dev::Counters* ctr, unsigned long long* historyCounter);
}
