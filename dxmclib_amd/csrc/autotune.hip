// Autotuner of the transport kernels.
//
// Times trial launches of the real path (runTransport) on a sample of the
// caller's exposures into a scratch tally and picks, one parameter at a
// time starting from the defaults, the value with the highest histories/s:
//   FP32 persistent kernel: register budget (resident waves per SIMD),
//                           LDS staging, block code ranges on or off,
//                           regeneration threshold, interaction threshold,
//                           queue depth (histories per pool refill), resident
//                           blocks per CU
//   FP64 plain kernel:      block size
// Each candidate is timed twice (best of two) from HIP events on the stream
// the kernel runs on. The tuning fields only change scheduling: every
// history keeps its own random stream, so results do not depend on them.
#include "dxmc_hip.h"
#include "internal.hpp"

#include <algorithm>
#include <vector>

using namespace dxmc_backend;

namespace {

struct Trial {
    dxmc_ctx* c;
    const dxmc_world* w;
    const dxmc_tables* t;
    const std::vector<dxmc_exposure>& x;
    dxmc_tally* tally;
    uint64_t histories;
    uint32_t trials = 0;

    // histories/s of one configuration (best of two launches), <0 on error
    double rate(dxmc_run_cfg cfg, int& status)
    {
        cfg.flags &= ~DXMC_RUN_AUTOTUNE;
        cfg.histories_per_launch = histories; // one launch per trial
        double best = -1;
        for (int rep = 0; rep < 2; ++rep) {
            dxmc_run_stats s {};
            status = runTransport(c, w, t, x.data(), x.size(), cfg, tally, nullptr, &s);
            if (status)
                return -1;
            ++trials;
            if (s.kernel_ms > 0)
                best = std::max(best, static_cast<double>(s.histories) / (s.kernel_ms * 1e-3));
        }
        return best;
    }
};

// A value replaces the incumbent only when it is faster by more than this
// fraction: trial timings scatter by about a percent, and taking every
// lucky measurement drifts away from the defaults for nothing.
constexpr double MIN_GAIN = 0.015;

// coordinate search over one field: keeps the value with the best rate
int searchField(Trial& tr, dxmc_run_cfg& cfg, uint32_t dxmc_run_cfg::*field, const std::vector<uint32_t>& values, double& bestRate)
{
    if (cfg.*field != 0)
        return DXMC_OK; // fixed by the caller
    uint32_t bestValue = 0;
    for (uint32_t v : values) {
        dxmc_run_cfg k = cfg;
        k.*field = v;
        int st = DXMC_OK;
        const double r = tr.rate(k, st);
        if (st)
            return st;
        if (r > bestRate * (1 + MIN_GAIN)) {
            bestRate = r;
            bestValue = v;
        }
    }
    cfg.*field = bestValue; // 0 keeps the default when it was best
    return DXMC_OK;
}

}

namespace dxmc_backend {

int autotune(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg& base,
    uint64_t trialHistories, dxmc_run_cfg& best, dxmc_tune_report* report)
{
    if (trialHistories < 64 * 1024)
        trialHistories = 64 * 1024;
    // up to 64 exposures spread over the table, sharing the trial histories
    const uint64_t m = std::min<uint64_t>(nExp, 64);
    std::vector<dxmc_exposure> sample(m);
    for (uint64_t i = 0; i < m; ++i) {
        sample[i] = x[i * nExp / m];
        sample[i].n_histories = std::max<uint64_t>(1, trialHistories / m);
    }
    dxmc_tally* scratch = nullptr;
    int st = dxmc_hip_tally_create(c, w->n, &scratch);
    if (st)
        return st;
    Trial tr { c, w, t, sample, scratch, trialHistories / m * m };
    dxmc_run_cfg cfg = base;
    cfg.flags &= ~DXMC_RUN_AUTOTUNE;
    dxmc_run_cfg defaults = cfg;
    defaults.block_size = defaults.blocks_per_cu = defaults.regeneration_threshold = defaults.interaction_threshold = defaults.queue_depth = 0;
    defaults.waves_per_simd = defaults.lds_tables = defaults.block_bounds = 0;
    double defaultRate = tr.rate(defaults, st);
    double bestRate = st ? -1 : tr.rate(cfg, st);
    if (!st) {
        if (cfg.precision == DXMC_PRECISION_FP32) {
            st = searchField(tr, cfg, &dxmc_run_cfg::waves_per_simd, { 5, 6, 7 }, bestRate);
            if (!st)
                st = searchField(tr, cfg, &dxmc_run_cfg::lds_tables, { 1, 2, 3 }, bestRate);
            if (!st && w->blocks)
                st = searchField(tr, cfg, &dxmc_run_cfg::block_bounds, { 1, 2 }, bestRate);
            const uint32_t resident = residentBlocksFp32(cfg.low_energy_correction, wavesFp32(cfg.waves_per_simd, cfg.low_energy_correction, 0), 0);
            std::vector<uint32_t> perCU;
            for (uint32_t b : { resident > 1 ? resident - 1 : 1u, std::max(1u, resident / 2) })
                if (b != resident && std::find(perCU.begin(), perCU.end(), b) == perCU.end())
                    perCU.push_back(b);
            if (!st && !(st = searchField(tr, cfg, &dxmc_run_cfg::regeneration_threshold, { 1, 2, 8, 16 }, bestRate))
                && !(st = searchField(tr, cfg, &dxmc_run_cfg::interaction_threshold, { 1, 8, 16, 24, 32, 40, 48 }, bestRate))
                && !(st = searchField(tr, cfg, &dxmc_run_cfg::queue_depth, { 128, 256, 512, 1024, 2048, 4096 }, bestRate)))
                st = searchField(tr, cfg, &dxmc_run_cfg::blocks_per_cu, perCU, bestRate);
        } else {
            st = searchField(tr, cfg, &dxmc_run_cfg::block_size, { 64, 128, 256 }, bestRate);
        }
    }
    dxmc_hip_tally_destroy(c, scratch);
    if (st)
        return st;
    best = cfg;
    best.flags = base.flags;
    if (report) {
        report->trials = tr.trials;
        report->reserved = 0;
        report->default_rate = defaultRate;
        report->best_rate = bestRate;
    }
    return DXMC_OK;
}
}

extern "C" int dxmc_hip_autotune(dxmc_ctx* c, const dxmc_world* w, const dxmc_tables* t, const dxmc_exposure* x, uint64_t nExp, const dxmc_run_cfg* base,
    uint64_t trialHistories, dxmc_run_cfg* best, dxmc_tune_report* report)
{
    if (!c || !best)
        return c ? c->fail(DXMC_ERR_ARGUMENT, "autotune: NULL argument") : DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    int st = validateRun(c, w, t, x, nExp, base, nullptr);
    if (st)
        return st;
    DXMC_CHECK(c, hipSetDevice(c->device));
    st = autotune(c, w, t, x, nExp, *base, trialHistories, *best, report);
    if (st)
        return st;
    c->tuned[tuneKey(w, t, *base)] = *best;
    return DXMC_OK;
}
