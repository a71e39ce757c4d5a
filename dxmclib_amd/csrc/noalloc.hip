// dxmc_hip_run_noalloc: dxmc_hip_run for callers that cannot allocate.
//
// The device memory a run needs beyond its inputs and outputs (the exposure
// records of both precisions, the history prefix sums, the step/event
// counters and the FP32 kernel's work counter) lives in one file-scope
// device array, `scratch`, of fixed size. Each GPU the library runs on holds
// its own copy until the process ends; its address on a device is taken once
// with hipGetSymbolAddress when the device's first context is created, and
// only that address is passed to kernels and copies. The exposures are
// copied in raw and converted and summed on the device, so the call does no
// device or host allocation. Calls on one device share the array: the
// library serialises them with a per-device lock and each call clears what
// must start at zero.
#include "dxmc_hip.h"
#include "internal.hpp"

#include <mutex>

using namespace dxmc_dev;
using namespace dxmc_backend;

namespace {

constexpr uint64_t MAX_EXPOSURES = DXMC_NOALLOC_MAX_EXPOSURES;
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t { 255 }; }
// layout: work counter + step/event counters | raw exposures | FP64 records |
// FP32 records | prefix sums
constexpr size_t OFF_RAW = 256;
constexpr size_t OFF_E64 = OFF_RAW + align256(MAX_EXPOSURES * sizeof(dxmc_exposure));
constexpr size_t OFF_E32 = OFF_E64 + align256(MAX_EXPOSURES * sizeof(Exposure<double>));
constexpr size_t OFF_START = OFF_E32 + align256(MAX_EXPOSURES * sizeof(Exposure<float>));
constexpr size_t SCRATCH_BYTES = OFF_START + align256((MAX_EXPOSURES + 1) * sizeof(uint64_t));
static_assert(sizeof(unsigned long long) + sizeof(Counters) <= OFF_RAW, "counters fit the first 256 bytes");

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex gSetupMutex;
void* gScratch[MAX_DEVICES] = {};
std::mutex gDeviceMutex[MAX_DEVICES];

template <typename T>
__device__ Exposure<T> record(const dxmc_exposure& e)
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

// raw exposures -> device records of both precisions; one thread forms the
// history prefix sums and clears the counters
__global__ void prepareRun(const dxmc_exposure* raw, uint64_t n, Exposure<double>* e64, Exposure<float>* e32, uint64_t* start,
    unsigned long long* historyCounter, Counters* ctr)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        e64[i] = record<double>(raw[i]);
        e32[i] = record<float>(raw[i]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint64_t s = 0;
        start[0] = 0;
        for (uint64_t i = 0; i < n; ++i) {
            s += raw[i].n_histories;
            start[i + 1] = s;
        }
        *historyCounter = 0;
        *ctr = Counters {};
    }
}

}

namespace dxmc_backend {

// the device address of `scratch` on the current device, taken once per
// device (the caller has made `device` current)
int noallocScratch(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES)
        return DXMC_ERR_DEVICE;
    std::lock_guard lock(gSetupMutex);
    if (!gScratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            hipGetLastError();
            return DXMC_ERR_DEVICE;
        }
        gScratch[device] = p;
    }
    *out = gScratch[device];
    return DXMC_OK;
}

}

extern "C" int dxmc_hip_run_noalloc(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* exposures, uint64_t nExp,
    const dxmc_run_cfg* cfg, dxmc_tally* tally, dxmc_run_stats* stats)
{
    if (!c || !tally)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "run_noalloc: NULL argument") : DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    int st = validateRun(c, w, t, exposures, nExp, cfg, tally);
    if (st)
        return st;
    if (nExp > MAX_EXPOSURES)
        return c->fail(DXMC_ERR_ARGUMENT,
            "run_noalloc: " + std::to_string(nExp) + " exposures exceed the maximum of " + std::to_string(MAX_EXPOSURES) + " (use dxmc_hip_run)");
    dxmc_run_cfg run = *cfg;
    if (run.flags & DXMC_RUN_AUTOTUNE) {
        // the autotuner allocates trial tallies: only a configuration tuned
        // before (dxmc_hip_autotune or a dxmc_hip_run with the flag) is used
        const auto it = c->tuned.find(tuneKey(w, t, run));
        if (it == c->tuned.end())
            return c->fail(DXMC_ERR_ARGUMENT, "run_noalloc: DXMC_RUN_AUTOTUNE needs a configuration tuned before (dxmc_hip_autotune)");
        const uint32_t flags = run.flags;
        const uint64_t seed = run.seed, first = run.first_history, perLaunch = run.histories_per_launch;
        run = it->second;
        run.flags = flags;
        run.seed = seed;
        run.first_history = first;
        run.histories_per_launch = perLaunch;
    }
    if (!c->scratch)
        return c->fail(DXMC_ERR_DEVICE, "run_noalloc: the scratch array of this device has no address (hipGetSymbolAddress failed)");
    std::lock_guard deviceLock(gDeviceMutex[c->device]);
    DXMC_CHECK(c, hipSetDevice(c->device));
    auto* base = static_cast<unsigned char*>(c->scratch);
    auto* historyCounter = reinterpret_cast<unsigned long long*>(base);
    auto* ctr = reinterpret_cast<Counters*>(base + 64);
    auto* raw = reinterpret_cast<dxmc_exposure*>(base + OFF_RAW);
    auto* e64 = reinterpret_cast<Exposure<double>*>(base + OFF_E64);
    auto* e32 = reinterpret_cast<Exposure<float>*>(base + OFF_E32);
    auto* start = reinterpret_cast<uint64_t*>(base + OFF_START);
    uint64_t total = 0;
    for (uint64_t i = 0; i < nExp; ++i)
        total += exposures[i].n_histories;
    DXMC_CHECK(c, hipMemcpyAsync(raw, exposures, nExp * sizeof(dxmc_exposure), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(prepareRun, dim3(static_cast<uint32_t>(std::max<uint64_t>(1, (nExp + 255) / 256))), dim3(256), 0, c->stream, raw, nExp, e64,
        e32, start, historyCounter, ctr);
    DXMC_CHECK(c, hipGetLastError());
    const RunInputs in { e64, e32, start, nullptr, total, ctr, historyCounter };
    double kernelMs = 0;
    uint64_t launches = 0;
    if ((st = runLaunches(c, w, t, nExp, run, tally, in, nullptr, kernelMs, launches)))
        return st;
    Counters h {};
    DXMC_CHECK(c, hipMemcpyAsync(&h, ctr, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    DXMC_CHECK(c, hipStreamSynchronize(c->stream));
    if (stats) {
        stats->histories = total;
        stats->steps = h.steps;
        stats->events = h.events;
        stats->kernel_ms = kernelMs;
        stats->launches = launches;
    }
    return DXMC_OK;
}
