// Request server over one GPU (dxmc_host_server_* of include/dxmc_host.h).
//
// Pipeline per request:
//   submit      copies the world arrays and the source spec (caller thread)
//   prepare     builds World, Source and Transport::prepare(): attenuation
//               tables and the exposure table (pool of host threads)
//   admit       in arrival order, to the first idle slot, once the request's
//               device memory fits in the budget left by admitted requests;
//               a request that does not fit lets later ones that fit go first
//   execute     Transport::execute() on the slot's own HIP context and
//               stream: upload, transport, normalise, download
// Slots run concurrently, so the kernels of admitted requests overlap on
// the device, and a request that arrives while others run is admitted as
// soon as a slot and memory free up. Every history keeps its own random
// stream (seed, history index), so a request's result does not depend on
// what else runs.
#include "dxmc_host.h"
#include "host_internal.hpp"
#include "hostpool.hpp"

#include "dxmc/transport.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace dxmc;
using dxmc_host_detail::VectorPool;

namespace {

struct Request {
    uint64_t id = 0;
    // inputs, owned
    dxmc_host_world world {};
    VectorPool<uint8_t>::Handle material, measurement;
    VectorPool<float>::Handle density;
    std::vector<std::string> names;
    std::vector<const char*> namePtrs;
    dxmc_source_spec spec {};
    std::vector<double> bowtieAngles, bowtieWeights, spectrumWeights, spectrumEnergies;
    uint64_t seed = 0;
    int32_t mode = DXMC_OUTPUT_EV_PER_HISTORY;
    // prepared
    World<double> w;
    std::unique_ptr<Source<double>> source;
    std::unique_ptr<Transport<double>> transport;
    Transport<double>::Prepared prep;
    uint64_t bytes = 0;
    // result: n values each of dose (or energy imparted) and variance
    // (f64), then event counts (u32), in one block
    uint64_t n = 0;
    dxmc_host_detail::BlockPool::Handle result;
    double* dose() const { return reinterpret_cast<double*>(result.get()); }
    double* variance() const { return dose() + n; }
    uint32_t* events() const { return reinterpret_cast<uint32_t*>(variance() + n); }
    dxmc_run_stats stats {};
    int32_t state = DXMC_REQUEST_QUEUED;
    int32_t status = DXMC_OK;
    int32_t slot = -1;
    std::string error;
    double tSubmit = 0, tPrepared = 0, tAdmitted = 0, tUploaded = 0, tTransported = 0, tDownloaded = 0, tReleased = 0, tFinished = 0;
    double tPrepStarted = 0, tWorldBuilt = 0, tSourceReady = 0;
};

// device memory a request holds while admitted: FP64 world (u8 material,
// f64 density, optional u8 measurement), FP32 voxel records (at most 8 B per
// voxel plus a palette), bounded-density codes (at most 2 B per voxel plus
// their intervals), tally (32-byte voxel scores and the f64 dose, variance
// and u32 count result arrays), tables, exposure buffers and the upload's
// temporaries (12 MB of palette hash)
uint64_t deviceBytes(const Request& r)
{
    const uint64_t n = r.world.dims[0] * r.world.dims[1] * r.world.dims[2];
    return n * (1 + 8 + 8 + 2 + (r.world.measurement ? 1 : 0)) + 65536 * (8 + 16) + n * 52 + (20u << 20) + r.prep.exposures.size() * 512;
}

}

struct dxmc_server {
    dxmc_server_cfg cfg {};
    // host arrays reused across requests (declared first: they outlive the
    // requests that hold them)
    VectorPool<uint8_t> bytes { size_t { 2 } << 30 };
    VectorPool<float> floats { size_t { 4 } << 30 };
    VectorPool<double> doubles { size_t { 8 } << 30 };
    dxmc_host_detail::BlockPool results { size_t { 8 } << 30 };
    dxmc_host_detail::WorldArrays worldArrays { [this](size_t n) { return doubles.take(n); }, [this](size_t n) { return bytes.take(n); } };
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::mutex m;
    std::condition_variable cvPrep, cvReady, cvDone;
    std::deque<std::shared_ptr<Request>> prepQueue;
    std::deque<std::shared_ptr<Request>> ready; // prepared, arrival order
    std::map<uint64_t, std::shared_ptr<Request>> requests;
    uint64_t nextId = 1;
    uint64_t budget = 0, inUse = 0;
    bool stopping = false;
    bool paused = false;
    std::vector<dxmc_ctx*> contexts;
    std::vector<std::thread> threads;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    void prepareLoop()
    {
        for (;;) {
            std::shared_ptr<Request> r;
            {
                std::unique_lock lock(m);
                cvPrep.wait(lock, [&] { return stopping || !prepQueue.empty(); });
                if (stopping)
                    return;
                r = prepQueue.front();
                prepQueue.pop_front();
            }
            std::string why;
            int status = DXMC_OK;
            r->tPrepStarted = now();
            const bool built = dxmc_host_detail::buildWorld(r->world, r->w, why, &worldArrays);
            r->tWorldBuilt = now();
            if (built && (r->source = dxmc_host_detail::makeSource(r->spec)))
                r->source->validate(); // spectrum and filters (prepare() finds them valid)
            r->tSourceReady = now();
            if (!built) {
                status = DXMC_ERR_ARGUMENT;
                r->error = why;
            } else if (!r->source) {
                status = DXMC_ERR_ARGUMENT;
                r->error = "unknown source type";
            } else {
                r->transport = std::make_unique<Transport<double>>();
                auto& t = *r->transport;
                t.setDevice(cfg.device);
                t.setDevicePrecision(cfg.precision);
                t.setLowEnergyCorrectionModel(static_cast<LOWENERGYCORRECTION>(cfg.low_energy_correction));
                t.setOutputMode(r->mode == DXMC_OUTPUT_DOSE ? Transport<double>::OUTPUTMODE::DOSE : Transport<double>::OUTPUTMODE::EV_PER_HISTORY);
                t.setSeed(r->seed);
                dxmc_run_cfg tuning = cfg.tuning;
                tuning.flags |= DXMC_RUN_COUNTERS;
                t.setRunConfiguration(tuning);
                r->prep = t.prepare(r->w, r->source.get());
                if (!r->prep.valid) {
                    status = DXMC_ERR_ARGUMENT;
                    r->error = "source is not valid for this world";
                }
            }
            // the raw copies are no longer needed
            r->material.reset();
            r->density.reset();
            r->measurement.reset();
            std::lock_guard lock(m);
            r->tPrepared = now();
            if (status == DXMC_OK) {
                r->bytes = deviceBytes(*r);
                if (r->bytes > budget) {
                    status = DXMC_ERR_ADMISSION;
                    r->error = "needs " + std::to_string(r->bytes) + " bytes of device memory, more than the budget of " + std::to_string(budget);
                }
            }
            if (status != DXMC_OK) {
                r->status = status;
                r->state = DXMC_REQUEST_DONE;
                r->tFinished = r->tPrepared;
                cvDone.notify_all();
                continue;
            }
            // keep arrival order among prepared requests
            auto it = std::find_if(ready.begin(), ready.end(), [&](const auto& q) { return q->id > r->id; });
            r->state = DXMC_REQUEST_READY;
            ready.insert(it, r);
            cvReady.notify_all();
        }
    }

    void slotLoop(int slot)
    {
        dxmc_ctx* ctx = contexts[slot];
        for (;;) {
            std::shared_ptr<Request> r;
            {
                std::unique_lock lock(m);
                cvReady.wait(lock, [&] {
                    if (stopping)
                        return true;
                    if (paused)
                        return false;
                    for (const auto& q : ready)
                        if (inUse + q->bytes <= budget)
                            return true;
                    return false;
                });
                if (stopping)
                    return;
                auto it = std::find_if(ready.begin(), ready.end(), [&](const auto& q) { return inUse + q->bytes <= budget; });
                r = *it;
                ready.erase(it);
                inUse += r->bytes;
                r->state = DXMC_REQUEST_RUNNING;
                r->slot = slot;
                r->tAdmitted = now();
            }
            r->n = r->w.size();
            r->result = results.take(r->n * (8 + 8 + 4));
            const double tStart = now();
            const bool ok = r->transport->executeInto(r->w, r->source.get(), r->prep, r->dose(), r->events(), r->variance(), nullptr, false, ctx) != 0;
            std::lock_guard lock(m);
            r->stats = r->transport->lastRunStats();
            const auto& timing = r->transport->lastDeviceTiming();
            r->tUploaded = tStart + timing.upload;
            r->tTransported = r->tUploaded + timing.transport;
            r->tDownloaded = r->tTransported + timing.download;
            r->tReleased = r->tDownloaded + timing.release;
            r->status = ok ? DXMC_OK : DXMC_ERR_DEVICE;
            r->error = r->transport->lastError();
            // keep the result only
            r->prep = {};
            r->w = World<double>();
            r->source.reset();
            if (!ok)
                r->result.reset();
            r->tFinished = now();
            r->state = DXMC_REQUEST_DONE;
            inUse -= r->bytes;
            cvDone.notify_all();
            cvReady.notify_all();
        }
    }
};

namespace {
int serverFail(int status, const std::string& msg) { return dxmc_host_detail::fail(status, msg); }
}

extern "C" {

int dxmc_host_server_create(const dxmc_server_cfg* cfg, dxmc_server** out)
{
    if (!cfg || !out)
        return serverFail(DXMC_ERR_ARGUMENT, "server_create: NULL argument");
    *out = nullptr;
    if (cfg->slots > 8 || cfg->prep_threads > 64 || cfg->low_energy_correction < 0 || cfg->low_energy_correction > 2
        || (cfg->precision != DXMC_PRECISION_FP64 && cfg->precision != DXMC_PRECISION_FP32))
        return serverFail(DXMC_ERR_ARGUMENT, "server_create: bad configuration");
    auto s = std::make_unique<dxmc_server>();
    s->cfg = *cfg;
    const uint32_t slots = cfg->slots ? cfg->slots : 4;
    const uint32_t prep = cfg->prep_threads ? cfg->prep_threads : 4;
    for (uint32_t i = 0; i < slots; ++i) {
        dxmc_ctx* c = nullptr;
        if (dxmc_hip_ctx_create(cfg->device, &c) != DXMC_OK) {
            for (auto* x : s->contexts)
                dxmc_hip_ctx_destroy(x);
            return serverFail(DXMC_ERR_DEVICE, std::string("server_create: ") + dxmc_hip_last_error(nullptr));
        }
        s->contexts.push_back(c);
    }
    uint64_t freeBytes = 0, totalBytes = 0;
    dxmc_hip_mem_info(s->contexts[0], &freeBytes, &totalBytes);
    s->budget = cfg->memory_budget ? cfg->memory_budget : freeBytes / 10 * 8;
    for (uint32_t i = 0; i < prep; ++i)
        s->threads.emplace_back([p = s.get()] { p->prepareLoop(); });
    for (uint32_t i = 0; i < slots; ++i)
        s->threads.emplace_back([p = s.get(), i] { p->slotLoop(static_cast<int>(i)); });
    *out = s.release();
    return DXMC_OK;
}

int dxmc_host_server_destroy(dxmc_server* s)
{
    if (!s)
        return DXMC_ERR_ARGUMENT;
    {
        std::lock_guard lock(s->m);
        s->stopping = true;
        s->cvPrep.notify_all();
        s->cvReady.notify_all();
    }
    for (auto& t : s->threads)
        t.join();
    for (auto* c : s->contexts)
        dxmc_hip_ctx_destroy(c);
    delete s;
    return DXMC_OK;
}

int dxmc_host_server_submit(dxmc_server* s, const dxmc_host_world* w, const dxmc_source_spec* spec, uint64_t seed, int32_t mode, uint64_t* id)
{
    if (!s || !w || !spec || !id)
        return serverFail(DXMC_ERR_ARGUMENT, "server_submit: NULL argument");
    if (mode != DXMC_OUTPUT_EV_PER_HISTORY && mode != DXMC_OUTPUT_DOSE)
        return serverFail(DXMC_ERR_ARGUMENT, "server_submit: unknown output mode");
    if (!w->material || !w->density || !w->material_names || w->n_materials == 0)
        return serverFail(DXMC_ERR_ARGUMENT, "server_submit: world arrays or material names missing");
    auto r = std::make_shared<Request>();
    const uint64_t n = w->dims[0] * w->dims[1] * w->dims[2];
    r->material = s->bytes.take(n);
    r->material->assign(w->material, w->material + n);
    r->density = s->floats.take(n);
    r->density->assign(w->density, w->density + n);
    if (w->measurement) {
        r->measurement = s->bytes.take(n);
        r->measurement->assign(w->measurement, w->measurement + n);
    }
    for (uint32_t i = 0; i < w->n_materials; ++i)
        r->names.emplace_back(w->material_names[i] ? w->material_names[i] : "");
    for (const auto& nm : r->names)
        r->namePtrs.push_back(nm.c_str());
    r->world = *w;
    r->world.material = r->material->data();
    r->world.density = r->density->data();
    r->world.measurement = w->measurement ? r->measurement->data() : nullptr;
    r->world.material_names = r->namePtrs.data();
    r->spec = *spec;
    if (spec->bowtie_n && spec->bowtie_angles && spec->bowtie_weights) {
        r->bowtieAngles.assign(spec->bowtie_angles, spec->bowtie_angles + spec->bowtie_n);
        r->bowtieWeights.assign(spec->bowtie_weights, spec->bowtie_weights + spec->bowtie_n);
        r->spec.bowtie_angles = r->bowtieAngles.data();
        r->spec.bowtie_weights = r->bowtieWeights.data();
    }
    if (spec->spectrum_n && spec->spectrum_weights && spec->spectrum_energies) {
        r->spectrumWeights.assign(spec->spectrum_weights, spec->spectrum_weights + spec->spectrum_n);
        r->spectrumEnergies.assign(spec->spectrum_energies, spec->spectrum_energies + spec->spectrum_n);
        r->spec.spectrum_weights = r->spectrumWeights.data();
        r->spec.spectrum_energies = r->spectrumEnergies.data();
    }
    r->seed = seed;
    r->mode = mode;
    // a world that can never fit the budget is refused at once
    const uint64_t minBytes = n * (1 + 8 + 1 + 20);
    std::lock_guard lock(s->m);
    if (s->stopping)
        return serverFail(DXMC_ERR_STATE, "server_submit: server is stopping");
    if (minBytes > s->budget)
        return serverFail(DXMC_ERR_ADMISSION, "server_submit: world exceeds the device memory budget");
    r->id = s->nextId++;
    r->tSubmit = s->now();
    s->requests[r->id] = r;
    s->prepQueue.push_back(r);
    s->cvPrep.notify_one();
    *id = r->id;
    return DXMC_OK;
}

int dxmc_host_server_wait(dxmc_server* s, uint64_t id, double timeout, int32_t* state)
{
    if (!s || !state)
        return DXMC_ERR_ARGUMENT;
    std::unique_lock lock(s->m);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        *state = DXMC_REQUEST_UNKNOWN;
        return DXMC_OK;
    }
    auto r = it->second;
    auto done = [&] { return r->state == DXMC_REQUEST_DONE; };
    if (timeout < 0)
        s->cvDone.wait(lock, done);
    else
        s->cvDone.wait_for(lock, std::chrono::duration<double>(timeout), done);
    *state = r->state;
    return DXMC_OK;
}

int dxmc_host_server_collect(dxmc_server* s, uint64_t id, double* dose, uint32_t* nEvents, double* variance, dxmc_request_info* info)
{
    if (!s)
        return DXMC_ERR_ARGUMENT;
    std::shared_ptr<Request> r;
    {
        std::lock_guard lock(s->m);
        auto it = s->requests.find(id);
        if (it == s->requests.end())
            return serverFail(DXMC_ERR_ARGUMENT, "server_collect: unknown request");
        if (it->second->state != DXMC_REQUEST_DONE)
            return serverFail(DXMC_ERR_STATE, "server_collect: request not finished");
        r = it->second;
        s->requests.erase(it);
    }
    if (info) {
        info->submitted = r->tSubmit;
        info->prepared = r->tPrepared;
        info->uploaded = r->tUploaded;
        info->transported = r->tTransported;
        info->downloaded = r->tDownloaded;
        info->released = r->tReleased;
        info->prep_started = r->tPrepStarted;
        info->world_built = r->tWorldBuilt;
        info->source_ready = r->tSourceReady;
        info->admitted = r->tAdmitted;
        info->finished = r->tFinished;
        info->histories = r->stats.histories;
        info->steps = r->stats.steps;
        info->events = r->stats.events;
        info->device_bytes = r->bytes;
        info->status = r->status;
        info->slot = r->slot;
    }
    if (r->status != DXMC_OK)
        return serverFail(r->status, "request " + std::to_string(id) + ": " + (r->error.empty() ? "failed" : r->error));
    {
        if (dose)
            std::copy_n(r->dose(), r->n, dose);
        if (nEvents)
            std::copy_n(r->events(), r->n, nEvents);
        if (variance)
            std::copy_n(r->variance(), r->n, variance);
    }
    return DXMC_OK;
}

int dxmc_host_server_pause(dxmc_server* s, int32_t paused)
{
    if (!s)
        return DXMC_ERR_ARGUMENT;
    std::lock_guard lock(s->m);
    s->paused = paused != 0;
    s->cvReady.notify_all();
    return DXMC_OK;
}

double dxmc_host_server_clock(const dxmc_server* s) { return s ? s->now() : 0.0; }
}
