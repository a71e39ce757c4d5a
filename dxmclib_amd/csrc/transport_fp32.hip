// FP32 photon transport: persistent kernel with wave-level regeneration.
//
// Each lane carries one photon (position, direction, energy, weight,
// majorant, PCG state) in registers and advances it one Woodcock step per
// loop iteration. Lanes whose photon has died wait until at least
// 'regeneration_threshold' lanes of the wave are dead (or none is alive);
// then one lane claims that many history indices with a single atomic on a
// global counter and every dead lane starts the history at its rank among the
// dead lanes (ballot + mbcnt). Histories keep their own PCG stream (the stream
// depends only on the global history index), so results do not depend on
// which lane or wave ran a history, on the grid size or on the threshold.
// The grid is sized to the machine (CUs x blocks per CU) and drains when the
// counter passes the end of the launch range and no lane is alive.
//
// Arithmetic is FP32 (LUT in base 2, FP32 transcendentals); each scoring
// event is flushed to the FP64 tallies with hardware FP64 atomics.
#include "internal.hpp"

using namespace dxmc_dev;

namespace {

__device__ inline uint32_t laneId() { return __lane_id(); }

__device__ inline uint32_t rankBelow(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <int L>
__global__ __launch_bounds__(256) void transportPersistent(World<float> w, Tables<float> t, const Exposure<float>* exps, const uint64_t* expStart,
    uint32_t nExp, uint64_t hBegin, uint64_t hEnd, uint64_t seed, uint64_t firstHistory, Tally tally, Counters* ctr,
    unsigned long long* next, uint32_t threshold)
{
    Particle<float> p;
    Rng rng { 0, 1 };
    float majInv = 0;
    bool update = true;
    bool alive = false;
    bool exhausted = false;
    Scorer sc { tally, 0 };
    uint32_t steps = 0;
    uint32_t historySteps = 0;
    const uint32_t lane = laneId();

    for (;;) {
        const uint64_t dead = __ballot(!alive);
        const uint32_t nDead = __popcll(dead);
        if (!exhausted && nDead > 0 && (nDead >= threshold || nDead == 64u)) {
            // claim nDead histories for the dead lanes of this wave
            const uint32_t leader = __ffsll(static_cast<unsigned long long>(dead)) - 1;
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(next, static_cast<unsigned long long>(nDead));
            base = __shfl(base, static_cast<int>(leader));
            if (hBegin + base + nDead >= hEnd)
                exhausted = true;
            if (!alive) {
                const uint64_t h = hBegin + base + rankBelow(dead);
                if (h < hEnd) {
                    const uint32_t e = upperBound(expStart, nExp + 1, h) - 1;
                    rng = Rng::forHistory(seed, firstHistory + h);
                    sampleParticle(exps[e], t, rng, p);
                    alive = transportToWorld(w, p);
                    update = true;
                    historySteps = 0;
                }
            }
        }
        if (!__any(alive)) {
            if (exhausted)
                break;
            continue;
        }
        if (alive) {
            if (L == 0)
                alive = woodcockStep<float, 0>(w, t, p, majInv, update, rng, sc, steps);
            else if (L == 1)
                alive = woodcockStep<float, 1>(w, t, p, majInv, update, rng, sc, steps);
            else
                alive = woodcockStep<float, 2>(w, t, p, majInv, update, rng, sc, steps);
            if (++historySteps >= MAX_STEPS)
                alive = false;
        }
    }
    if (ctr)
        dxmc_backend::accumulateCounters(ctr, steps, sc.events);
}

}

namespace dxmc_backend {

void launchFp32(dxmc_ctx* c, const dxmc_run_cfg& cfg, const World<float>& w, const Tables<float>& t, const Exposure<float>* exps,
    const uint64_t* expStart, uint32_t nExp, uint64_t hb, uint64_t he, const Tally& tally, Counters* ctr)
{
    auto* nx = static_cast<unsigned long long*>(c->historyCounter);
    hipMemsetAsync(nx, 0, sizeof(unsigned long long), c->stream);
    const uint32_t block = 256;
    const uint32_t perCU = cfg.blocks_per_cu ? cfg.blocks_per_cu : 8;
    const uint32_t threshold = cfg.regeneration_threshold ? std::min<uint32_t>(cfg.regeneration_threshold, 64u) : 16u;
    uint64_t grid = static_cast<uint64_t>(c->numCU) * perCU;
    const uint64_t lanesNeeded = (he - hb + 63) / 64 * 64;
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, (lanesNeeded + block - 1) / block));
    if (cfg.low_energy_correction == 0)
        hipLaunchKernelGGL(transportPersistent<0>, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, c->stream, w, t, exps, expStart, nExp, hb, he,
            cfg.seed, cfg.first_history, tally, ctr, nx, threshold);
    else if (cfg.low_energy_correction == 1)
        hipLaunchKernelGGL(transportPersistent<1>, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, c->stream, w, t, exps, expStart, nExp, hb, he,
            cfg.seed, cfg.first_history, tally, ctr, nx, threshold);
    else
        hipLaunchKernelGGL(transportPersistent<2>, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, c->stream, w, t, exps, expStart, nExp, hb, he,
            cfg.seed, cfg.first_history, tally, ctr, nx, threshold);
}
}
