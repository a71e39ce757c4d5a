// FP32 photon transport: persistent kernel with wave-level regeneration and
// interaction compaction.
//
// Each lane carries one photon (position, direction, energy, weight,
// majorant, PCG state) in registers and is in one of three phases:
//   DEAD     no photon: waits for regeneration
//   FLYING   advances one Woodcock flight per loop iteration (distance under
//            the majorant, voxel gather, real-or-virtual decision, roulette
//            after virtual collisions)
//   PENDING  a real interaction is due in the voxel just reached
// Per iteration the wave does one of two things, decided from ballots:
//   - when at least 'interaction_threshold' lanes are PENDING (or none is
//     FLYING) it runs the interaction code for the pending lanes only,
//   - otherwise it advances the FLYING lanes.
// So the long, branchy interaction samplers (Compton/Rayleigh rejection
// loops, fluorescence) run with many lanes active instead of with the few
// lanes that happen to interact in a given flight, while the cheap flights
// run nearly full. New photons are sampled by the whole wave at once (up to
// the queue's 64 or 48 histories from a wave-private pool of history
// indices, which one lane refills with a single atomic on the global
// counter) into a per-wave queue in LDS; dead lanes take photons from that
// queue once at least
// 'regeneration_threshold' of them wait, each by its rank among the dead
// lanes (ballot + mbcnt). So the costly source sampling runs with all lanes
// active and a lane is refilled with a few LDS reads.
//
// Each history keeps its own PCG stream and consumes it in the reference's
// order (flight, decision, interaction, roulette), so delaying an interaction
// changes nothing but timing: results depend neither on the lane or wave that
// ran a history nor on the grid size or thresholds.
//
// Arithmetic is FP32 (base-2 anchored LUT, FP32 transcendentals, voxel index
// by multiplication with the reciprocal spacing, FP32 denormals flushed);
// each scoring event is added to the FP64 tallies with hardware FP64
// atomics. The kernel is compiled per low-energy model, register budget
// (5, 6 or 7 waves per SIMD), LDS staging level and world kind.
#include "internal.hpp"

using namespace dxmc_dev;

namespace {

enum : uint32_t { DEAD = 0, FLYING = 1, PENDING = 2 };

// World kinds the kernel is compiled for: the two common layouts get their
// code width, record width, 24-bit index and bounds table fixed at compile
// time, so a flight carries no branches on them
//   WORLD_CODES8    1-byte bounded-density codes (and their block codes
//                   when the run uses them), 8-byte records, the
//                   attenuation bounds table, 24-bit index (CT-sized worlds)
//   WORLD_PALETTE8  1-byte palette indices, no codes, 24-bit index
//   WORLD_ANY       every other world (branches at run time)
enum : uint32_t { WORLD_ANY = 0, WORLD_CODES8 = 1, WORLD_PALETTE8 = 2 };


__device__ inline uint32_t rankBelow(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// voxel index with the reciprocal spacing (FP32 path), kept inside the grid
// (the float-to-unsigned conversion clamps negative values to 0); with
// 24-bit products where the grid allows (the compiler turns a 24-bit
// multiply-add into a 64-bit one, so the three terms are summed instead)
__device__ inline uint3 voxelCoordsFast(const World<float>& w, const float* pos, const float* inv)
{
    return make_uint3(min(static_cast<uint32_t>((pos[0] - w.extent[0]) * inv[0]), w.nx - 1),
        min(static_cast<uint32_t>((pos[1] - w.extent[2]) * inv[1]), w.ny - 1), min(static_cast<uint32_t>((pos[2] - w.extent[4]) * inv[2]), w.nz - 1));
}

__device__ inline uint64_t voxelIndexFast(const World<float>& w, uint3 v, bool index24)
{
    if (index24)
        return __umul24(v.z, w.nxy) + __umul24(v.y, w.nx) + v.x;
    return (static_cast<uint64_t>(v.z) * w.ny + v.y) * w.nx + v.x;
}

// element 'index' of a table addressed by a 32-bit byte offset from its
// base (tables below 4 GiB): the load takes the base from scalar registers
// and the offset from one vector register, with no 64-bit address
// arithmetic per lane
template <typename T>
__device__ inline T loadAt(const T* base, uint32_t index)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const unsigned char*>(base) + index * static_cast<uint32_t>(sizeof(T)));
}

// block of a voxel (1-byte code worlds: 24-bit products hold)
__device__ inline uint32_t blockIndexFast(const World<float>& w, uint3 v)
{
    return __umul24(v.z >> w.blockShift, w.bnxy) + __umul24(v.y >> w.blockShift, w.bnx) + (v.x >> w.blockShift);
}

// ------------------------------------------------------ LDS-staged tables
// The tables the samplers read with data-dependent indices (RITA, spline,
// shells, spectrum alias tables, heel and bowtie filters, the majorant and
// the segment bounds, optionally the attenuation coefficients) are copied
// into LDS at kernel start, one copy per block: their lookups then cost LDS
// latency instead of L2 latency, also while the world gathers keep evicting
// them from L1. Staging level: 1 none, 2 all but the attenuation
// coefficients, 3 all. Of the interaction tables only those of the run's
// low-energy model are staged (RITA for L >= 1, the scatter-function spline
// for L = 1, the shells for L = 2). The total-attenuation bounds stay in
// global memory: in LDS they cost a resident block per CU.
enum : uint32_t { STAGE_NONE = 1, STAGE_SMALL = 2, STAGE_ALL = 3 };

__host__ __device__ constexpr uint32_t ldsAlign(uint64_t bytes) { return static_cast<uint32_t>((bytes + 15) & ~uint64_t(15)); }

// bytes of LDS a staging level needs for low-energy model L (the same walk
// as stageTables)
__host__ __device__ inline uint32_t ldsBytes(const Tables<float>& t, uint32_t level, int L)
{
    if (level <= STAGE_NONE)
        return 0;
    uint32_t b = 0;
    if (L >= 1)
        b += ldsAlign(uint64_t(t.nMaterials) * 4 * 56 * sizeof(float));
    if (L == 1)
        b += ldsAlign(uint64_t(t.nMaterials) * 63 * sizeof(float));
    if (L == 2)
        b += ldsAlign(uint64_t(t.nMaterials) * 12 * sizeof(Shell<float>));
    b += ldsAlign(uint64_t(t.lutNodes) * 2 * sizeof(float));
    b += ldsAlign(uint64_t(t.lutNodes) * sizeof(float));
    b += ldsAlign(uint64_t(t.lutNodes) * sizeof(float));
    b += ldsAlign(uint64_t(t.nSpectra + 1) * sizeof(uint32_t));
    b += ldsAlign(uint64_t(t.nSpectra ? t.nSpectra : 1) * sizeof(uint32_t));
    b += 2 * ldsAlign(uint64_t(t.specTotal) * sizeof(float)) + ldsAlign(uint64_t(t.specTotal) * sizeof(uint32_t));
    b += ldsAlign(uint64_t(t.nHeel) * sizeof(Heel<float>)) + ldsAlign(uint64_t(t.heelWeightTotal) * sizeof(float));
    b += ldsAlign(uint64_t(t.nBowtie + 1) * sizeof(uint32_t)) + 2 * ldsAlign(uint64_t(t.bowtieTotal) * sizeof(float));
    if (level >= STAGE_ALL)
        b += ldsAlign(uint64_t(t.nMaterials) * t.lutNodes * 6 * sizeof(float));
    return b;
}

template <typename T>
__device__ inline const T* stageArray(const T* src, uint64_t count, unsigned char*& cursor)
{
    static_assert(sizeof(T) % 4 == 0);
    T* dst = reinterpret_cast<T*>(cursor);
    const uint64_t words = count * sizeof(T) / 4;
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    for (uint64_t i = threadIdx.x; i < words; i += blockDim.x)
        d[i] = s[i];
    cursor += ldsAlign(count * sizeof(T));
    return dst;
}

// copies the tables into LDS and points t at the copies (all threads of the
// block call it; ends with a barrier). The level is a template parameter so
// that the compiler sees which pointers address LDS.
template <uint32_t level, int L>
__device__ inline void stageTables(Tables<float>& t, unsigned char* lds)
{
    if constexpr (level <= STAGE_NONE)
        return;
    unsigned char* c = lds;
    if constexpr (L >= 1)
        t.rita = stageArray(t.rita, uint64_t(t.nMaterials) * 4 * 56, c);
    if constexpr (L == 1)
        t.spline = stageArray(t.spline, uint64_t(t.nMaterials) * 63, c);
    if constexpr (L == 2)
        t.shells = stageArray(t.shells, uint64_t(t.nMaterials) * 12, c);
    t.lutMajorant = stageArray(t.lutMajorant, uint64_t(t.lutNodes) * 2, c);
    t.lutBoundUp = stageArray(t.lutBoundUp, t.lutNodes, c);
    t.lutAnchor = stageArray(t.lutAnchor, t.lutNodes, c);
    t.specOffset = stageArray(t.specOffset, t.nSpectra + 1, c);
    t.specThreshold = stageArray(t.specThreshold, t.nSpectra ? t.nSpectra : 1, c);
    t.specProb = stageArray(t.specProb, t.specTotal, c);
    t.specEnergy = stageArray(t.specEnergy, t.specTotal, c);
    t.specAlias = stageArray(t.specAlias, t.specTotal, c);
    t.heel = stageArray(t.heel, t.nHeel, c);
    t.heelWeight = stageArray(t.heelWeight, t.heelWeightTotal, c);
    t.bowtieOffset = stageArray(t.bowtieOffset, t.nBowtie + 1, c);
    t.bowtieAngle = stageArray(t.bowtieAngle, t.bowtieTotal, c);
    t.bowtieWeight = stageArray(t.bowtieWeight, t.bowtieTotal, c);
    if constexpr (level >= STAGE_ALL)
        t.lutCoef = stageArray(t.lutCoef, uint64_t(t.nMaterials) * t.lutNodes * 6, c);
    __syncthreads();
}

constexpr float LN2F = 0.693147180559945309f;

// per-wave queue of sampled photons in LDS: position, direction, energy,
// weight, PCG state, the history's offset in the launch (its PCG increment
// is 2 (first history + launch begin + offset) + 1) and the LUT segment of
// the energy, found while the whole wave samples (the segment search has
// loops; a lane that takes the photon only adds the segment offset and the
// majorant)
constexpr int QUEUE_FIELDS = 12;
// photons a wave's queue holds: 64 (the whole wave samples at once), or in
// the 7-wave kernels, whose LDS must fit 7 blocks per CU, a capacity the
// launch passes (48, or less where the staged tables are larger)
__host__ __device__ constexpr uint32_t queueBytes(uint32_t capacity) { return QUEUE_FIELDS * capacity * 4; }
constexpr uint32_t WAVES_PER_BLOCK = 4;

// inside the safe extent, without short-circuit branches
__device__ inline bool insideFast(const World<float>& w, const float* pos)
{
    return (pos[0] > w.extent[0]) & (pos[0] < w.extent[1]) & (pos[1] > w.extent[2]) & (pos[1] < w.extent[3]) & (pos[2] > w.extent[4])
        & (pos[2] < w.extent[5]);
}

// Russian roulette after a step that left the photon alive (transport.hpp:684-694)
__device__ inline bool roulette(Particle<float>& p, Rng& rng)
{
    if (p.energy * p.weight < 5.0f) {
        if (rng.uniform<float>() < 0.8f)
            return false;
        p.weight *= 1.0f / (1.0f - 0.8f);
    }
    return true;
}

// W: minimum resident waves per SIMD the register allocation must allow
// (1: unconstrained, 5 waves; 6: at most 80 VGPRs, spill-free at L = 0, 1;
// 7: at most 72 VGPRs, spill-free for the two fixed world kinds at L = 0, 1,
// with a 48-photon queue). S: LDS staging level. K: world kind. The
// autotuner picks among the register budgets and staging levels.
template <int L, int W, uint32_t S, uint32_t K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, 8))) void transportPersistent(World<float> w, Tables<float> tg, const Exposure<float>* exps, const uint64_t* expStart,
    uint32_t nExp, uint64_t hBegin, uint64_t hEnd, uint64_t seed, uint64_t firstHistory, Tally tally, Counters* ctr, unsigned long long* next,
    uint32_t regenThreshold, uint32_t interactThreshold, uint32_t poolSize, uint32_t queueCapacity)
{
    extern __shared__ __align__(16) unsigned char lds[];
    Tables<float> t = tg;
    stageTables<S, L>(t, lds);
    Particle<float> p;
    Rng rng { 0, 1 };
    float majInv = 0;
    bool update = true;
    uint32_t phase = DEAD;
    // the wave's pool of claimed, not yet started histories [poolNext, poolEnd)
    const uint64_t span = hEnd - hBegin;
    uint64_t poolNext = 0, poolEnd = 0;
    uint32_t poolExp = 0;
    bool claimedAll = span == 0;
    Scorer sc { tally, 0 };
    // in-world steps of the lane's finished histories, and of its current one
    // (added to 'steps' when the lane takes its next photon or at the end)
    uint32_t steps = 0;
    uint32_t historySteps = 0;
    // the pending interaction
    uint64_t idx = 0;
    // material | measurement flag << 24 of the voxel reached (one register;
    // the LUT row products are 24-bit and ignore the flag)
    uint32_t matMeas = 0;
    float density = 0;
    // the photon's place in the LUT; changes only with its energy
    LutPoint<float> lp { 0, 0.0f };
    const float inv[3] = { 1.0f / w.spacing[0], 1.0f / w.spacing[1], 1.0f / w.spacing[2] };
    // the world's layout, fixed by K where it can be
    const uint32_t codeBytes = K == WORLD_CODES8 ? 1u : K == WORLD_PALETTE8 ? 0u : w.codeBytes;
    const uint32_t voxelBytes = K == WORLD_PALETTE8 ? 1u : K == WORLD_CODES8 ? 8u : w.voxelBytes;
    const bool index24 = K != WORLD_ANY || w.index24 != 0;
    const bool sumBounds = K == WORLD_CODES8 || (K == WORLD_ANY && t.lutSumBound != nullptr);
    // global index of the launch's first history (PCG increments of queued photons)
    const uint64_t launchFirst = firstHistory + hBegin;
    const uint32_t lane = __lane_id();

    // the wave's queue of new photons, already in the world (LDS, SoA)
    // (fixed at 48 on palette worlds, where a run-time capacity cost DX 2 %)
    const uint32_t QUEUE_CAP = W < 7 ? 64u : K == WORLD_PALETTE8 ? 48u : queueCapacity;
    float* q = reinterpret_cast<float*>(lds + ldsBytes(tg, S, L) + (threadIdx.x / 64) * queueBytes(QUEUE_CAP));
    uint32_t queued = 0;

    for (;;) {
        uint64_t dead = __ballot(phase == DEAD);
        uint32_t nDead = __popcll(dead);
        const bool work = poolNext < poolEnd || !claimedAll;
        if (work && queued < nDead) {
            // sample new histories with the whole wave: lanes 0 .. want-1
            // take the next histories of the wave's pool (refilled from the
            // global counter as needed), photons that enter the world join
            // the queue
            const uint32_t want = QUEUE_CAP - queued;
            uint32_t assigned = 0;
            bool sampled = false;
            Particle<float> np;
            Rng nr { 0, 1 };
            uint32_t nOffset = 0;
            while (assigned < want) {
                if (poolNext >= poolEnd) {
                    if (claimedAll)
                        break;
                    unsigned long long base = 0;
                    if (lane == 0)
                        base = atomicAdd(next, static_cast<unsigned long long>(poolSize));
                    base = __shfl(base, 0);
                    // wave-uniform: the pool bounds live in scalar registers
                    base = (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base >> 32))) << 32)
                        | __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base));
                    if (base >= span) {
                        claimedAll = true;
                        break;
                    }
                    poolNext = hBegin + base;
                    poolEnd = hBegin + (base + poolSize < span ? base + poolSize : span);
                    if (poolEnd == hEnd)
                        claimedAll = true;
                    poolExp = upperBound(expStart, nExp + 1, poolNext) - 1;
                }
                const uint64_t avail = poolEnd - poolNext;
                const uint32_t take = avail < want - assigned ? static_cast<uint32_t>(avail) : want - assigned;
                if (lane >= assigned && lane < assigned + take) {
                    const uint64_t h = poolNext + (lane - assigned);
                    uint32_t e = poolExp;
                    while (h >= expStart[e + 1])
                        ++e;
                    nr = Rng::forHistory(seed, firstHistory + h);
                    nOffset = static_cast<uint32_t>(h - hBegin);
                    sampleParticle(exps[e], t, nr, np);
                    sampled = transportToWorld(w, np);
                }
                poolNext += take;
                assigned += take;
            }
            const uint64_t entered = __ballot(sampled);
            if (sampled) {
                const uint32_t slot = queued + rankBelow(entered);
                const float v[QUEUE_FIELDS] = { np.pos[0], np.pos[1], np.pos[2], np.dir[0], np.dir[1], np.dir[2], np.energy, np.weight,
                    __uint_as_float(static_cast<uint32_t>(nr.s)), __uint_as_float(static_cast<uint32_t>(nr.s >> 32)), __uint_as_float(nOffset),
                    __uint_as_float(lutPoint(t, np.energy).k) };
                for (int f = 0; f < QUEUE_FIELDS; ++f)
                    q[f * QUEUE_CAP + slot] = v[f];
            }
            queued += __popcll(entered);
        }
        // dead lanes take photons from the queue (last in, first out)
        if (nDead >= regenThreshold || (nDead > 0 && !work)) {
            const uint32_t take = nDead < queued ? nDead : queued;
            if (phase == DEAD) {
                const uint32_t r = rankBelow(dead);
                if (r < take) {
                    const uint32_t slot = queued - 1 - r;
                    float v[QUEUE_FIELDS];
                    for (int f = 0; f < QUEUE_FIELDS; ++f)
                        v[f] = q[f * QUEUE_CAP + slot];
                    for (int i = 0; i < 3; ++i) {
                        p.pos[i] = v[i];
                        p.dir[i] = v[3 + i];
                    }
                    p.energy = v[6];
                    p.weight = v[7];
                    rng.s = static_cast<uint64_t>(__float_as_uint(v[9])) << 32 | __float_as_uint(v[8]);
                    rng.inc = 2 * (launchFirst + __float_as_uint(v[10])) + 1;
                    // the LUT point lutPoint() gives, from the queued segment
                    lp.k = __float_as_uint(v[11]);
                    lp.x = __builtin_amdgcn_logf(p.energy) - t.lutAnchor[lp.k];
                    majInv = majorantInverseAt(t, lp);
                    update = false;
                    steps += historySteps;
                    historySteps = 0;
                    phase = FLYING;
                }
            }
            queued -= take;
        }
        const uint64_t pend = __ballot(phase == PENDING);
        const uint64_t fly = __ballot(phase == FLYING);
        if ((pend | fly) == 0) {
            if (poolNext >= poolEnd && claimedAll && queued == 0)
                break;
            continue;
        }
        const uint32_t nPend = __popcll(pend);
        if (nPend != 0 && (nPend >= interactThreshold || fly == 0)) {
            if (phase == PENDING) {
                const uint32_t mat = matMeas & 0xffu;
                float att[3];
                attenuationAt(t, mat, lp, att);
                bool alive;
                if (matMeas > 0xffffffu) { // forced event (measurement voxel)
                    const float prob = ((att[0] + att[1]) + att[2]) * density * majInv;
                    alive = computeInteractionsForced<float, L>(t, prob, att, p, mat, sc, idx, rng, update);
                } else {
                    alive = computeInteractions<float, L>(t, att, p, mat, sc, idx, rng, update);
                }
                if (alive)
                    alive = roulette(p, rng);
                if (alive && update) { // new energy: its LUT segment and majorant
                    lp = lutPoint(t, p.energy);
                    majInv = majorantInverseAt(t, lp);
                    update = false;
                }
                phase = alive ? FLYING : DEAD;
            }
        } else {
            // Woodcock flights of the FLYING lanes, one after the other
            // while no lane needs the outer loop: it ends when interactions
            // are due (interaction_threshold pending lanes, or none flying)
            // or dead lanes wait (for a refill of the queue while histories
            // remain, or for regeneration). Only the flight's own values are
            // live in this loop, so they stay in registers.
            for (;;) {
                if (phase == FLYING) {
                    // one Woodcock flight (woodcockParticleTracking, transport.hpp:
                    // 650-699), written without branches apart from the rare
                    // roulette: the random draws the reference makes only inside
                    // the world (decision) are drawn and kept by selection
                    const float u1 = rng.uniform<float>();
                    const float step = -__builtin_amdgcn_logf(u1) * (LN2F * 10.0f) * majInv; // cm -> mm
                    for (int i = 0; i < 3; ++i)
                        p.pos[i] += p.dir[i] * step;
                    const bool in = insideFast(w, p.pos);
                    const uint3 vc = voxelCoordsFast(w, p.pos, inv); // clamped: a valid voxel also outside
                    idx = voxelIndexFast(w, vc, index24);
                    Rng r2 = rng;
                    const float u2 = r2.uniform<float>();
                    // The acceptance probability (Σμ·ρ)·(1/μ̂) is monotone in Σμ and
                    // in ρ. Bounds of both settle most draws exactly as the values
                    // would: u2 < p(lower bounds) accepts, u2 >= p(upper bounds)
                    // rejects. Σμ is bounded over the photon's LUT segment; ρ by the
                    // voxel's block when the block holds one material, else by the
                    // voxel's code when the world has codes, else read exactly. Only
                    // a draw between the bounds evaluates the attenuation (and reads
                    // the record).
                    float prob = 0.0f;
                    bool settled = false;
                    // (the L = 2 kernel also runs without block codes: a run
                    // switch there costs no occupancy; at L = 0, 1 a world
                    // without them runs the K = any kernel)
                    if (K == WORLD_CODES8 && (L < 2 || w.blocks != nullptr)) {
                        const uint32_t bc = loadAt(w.blocks, blockIndexFast(w, vc));
                        // the material from the largest code by arithmetic, so
                        // the intervals and the segment bounds load together
                        const uint32_t bm = __umul24(bc >> 8, w.codeMul) >> CODE_MUL_SHIFT;
                        const float bLo = w.bounds[bc & 0xffu].lo;
                        const float bHi = w.bounds[bc >> 8].hi;
                        const float2 sb = loadAt(t.lutSumBound, __umul24(bm, t.lutNodes) + lp.k);
                        const float lo = sb.x * bLo * majInv;
                        const float hi = sb.y * bHi * majInv;
                        settled = (bc != BLOCK_MIXED) & ((u2 < lo) | (u2 >= hi));
                        matMeas = bm;
                        density = bLo;
                        prob = lo;
                    }
                    if (!settled) {
                        float dLo, dHi;
                        if (K == WORLD_CODES8) {
                            // material by arithmetic, interval and segment bounds
                            // loaded together (code worlds carry no measurement flags)
                            // (1-byte codes: below 2^32 voxels, a 32-bit offset)
                            const uint32_t code = loadAt(static_cast<const uint8_t*>(w.codes), static_cast<uint32_t>(idx));
                            matMeas = __umul24(code, w.codeMul) >> CODE_MUL_SHIFT;
                            const uint2 b = *reinterpret_cast<const uint2*>(w.bounds + code);
                            dLo = __uint_as_float(b.x);
                            dHi = __uint_as_float(b.y);
                            density = dLo;
                        } else if (codeBytes != 0) {
                            const uint32_t code = codeBytes == 1 ? static_cast<const uint8_t*>(w.codes)[idx] : static_cast<const uint16_t*>(w.codes)[idx];
                            // the interval as one 16-byte load (read field by field,
                            // the compiler splits it into two loads on the chain)
                            const uint4 b = *reinterpret_cast<const uint4*>(w.bounds + code);
                            matMeas = b.z; // material; code worlds carry no measurement flags
                            dLo = __uint_as_float(b.x);
                            dHi = __uint_as_float(b.y);
                            density = dLo;
                        } else {
                            readVoxelPacked(w, voxelBytes, idx, matMeas, density);
                            dLo = density;
                            dHi = density;
                        }
                        if (sumBounds) {
                            const float2 sb = loadAt(t.lutSumBound, __umul24(matMeas, t.lutNodes) + lp.k);
                            const float lo = sb.x * dLo * majInv;
                            const float hi = sb.y * dHi * majInv;
                            settled = (u2 < lo) | (u2 >= hi);
                            prob = lo;
                        }
                        if (!settled) {
                            if (codeBytes != 0)
                                density = __uint_as_float(static_cast<const uint2*>(w.voxels)[idx].x);
                            float att[3];
                            attenuationAt(t, matMeas, lp, att); // 24-bit row product: flag ignored
                            prob = ((att[0] + att[1]) + att[2]) * density * majInv;
                        }
                    }
                    const bool measured = matMeas > 0xffffffu;
                    const bool drawn = in & !measured;
                    rng.s = drawn ? r2.s : rng.s;
                    const bool real = measured | (u2 < prob);
                    uint32_t next = in ? (real ? PENDING : FLYING) : DEAD;
                    if (next == FLYING && p.energy * p.weight < 5.0f) { // Russian roulette
                        if (rng.uniform<float>() < 0.8f)
                            next = DEAD;
                        else
                            p.weight *= 1.0f / (1.0f - 0.8f);
                    }
                    historySteps += in ? 1u : 0u;
                    phase = (next != DEAD && historySteps >= MAX_STEPS) ? DEAD : next;
                }
                const uint64_t flyNow = __ballot(phase == FLYING);
                const uint32_t nPendNow = __popcll(__ballot(phase == PENDING));
                const uint32_t nDeadNow = __popcll(__ballot(phase == DEAD));
                if (flyNow == 0 || nPendNow >= interactThreshold || nDeadNow >= regenThreshold || (work && nDeadNow > queued)
                    || (!work && nDeadNow > 0 && queued > 0))
                    break;
            }
        }
    }
    if (ctr)
        dxmc_backend::accumulateCounters(ctr, steps + historySteps, sc.events);
}

}

namespace dxmc_backend {

namespace {

__global__ void clearCounter(unsigned long long* counter) { *counter = 0; }

using KernelFn = void (*)(World<float>, Tables<float>, const Exposure<float>*, const uint64_t*, uint32_t, uint64_t, uint64_t, uint64_t, uint64_t,
    Tally, Counters*, unsigned long long*, uint32_t, uint32_t, uint32_t, uint32_t);

// waves: resolved by wavesFp32 (5, 6 or 7); the 7-wave kernel (72 VGPRs)
// exists for the two fixed world kinds at L = 0, 1
template <int L, uint32_t S, uint32_t K>
KernelFn variant(uint32_t waves)
{
    if constexpr (K != WORLD_ANY && L < 2)
        if (waves >= 7)
            return transportPersistent<L, 7, S, K>;
    return waves >= 6 ? transportPersistent<L, 6, S, K> : transportPersistent<L, 1, S, K>;
}

template <int L, uint32_t K>
KernelFn variant(uint32_t waves, uint32_t stage)
{
    return stage >= STAGE_ALL ? variant<L, STAGE_ALL, K>(waves) : stage >= STAGE_SMALL ? variant<L, STAGE_SMALL, K>(waves) : variant<L, STAGE_NONE, K>(waves);
}

template <int L>
KernelFn variant(uint32_t waves, uint32_t stage, uint32_t kind)
{
    return kind == WORLD_CODES8 ? variant<L, WORLD_CODES8>(waves, stage)
        : kind == WORLD_PALETTE8 ? variant<L, WORLD_PALETTE8>(waves, stage) : variant<L, WORLD_ANY>(waves, stage);
}

KernelFn kernelFor(int L, uint32_t waves, uint32_t stage, uint32_t kind)
{
    return L == 0 ? variant<0>(waves, stage, kind) : L == 1 ? variant<1>(waves, stage, kind) : variant<2>(waves, stage, kind);
}

// the kind a world and its tables run as (after launchFp32 has dropped the
// codes and bounds where they do not pay)
uint32_t worldKind(const World<float>& w, const Tables<float>& t)
{
    if (!w.index24)
        return WORLD_ANY;
    if (w.codeBytes == 1 && w.voxelBytes == 8 && t.lutSumBound)
        return WORLD_CODES8;
    if (w.codeBytes == 0 && w.voxelBytes == 1 && !t.lutSumBound)
        return WORLD_PALETTE8;
    return WORLD_ANY;
}

}

// Resident waves per SIMD of a run: the requested 5, 6 or 7, or by default
// 6 where that fits 80 VGPRs without spills (L = 0, 1: CT spiral 3.55e9 vs
// 3.34e9, DX 4.13e9 vs 3.90e9 histories/s with 5), 7 for the two fixed
// world kinds (palette: DX 4.50e9 vs 4.39e9; 1-byte codes with block codes:
// CT spiral 3.95e9 vs 3.86e9, whole body 1.69e9 vs 1.59e9, their flights
// being bound by gather latency more than by VALU issue; the 72-VGPR code
// kernel keeps one lane counter in scratch, touched once per history);
// L = 2 would spill at 6 and keeps 5.
uint32_t wavesFp32(uint32_t requested, int L, uint32_t kind)
{
    if (requested == 0)
        return L == 2 ? 5u : kind != WORLD_ANY ? 7u : 6u;
    if (requested >= 7)
        return kind != WORLD_ANY && L < 2 ? 7u : 6u;
    return requested >= 6 ? 6u : 5u;
}

uint32_t residentBlocksFp32(int L, uint32_t waves, uint32_t ldsPerBlock, uint32_t kind)
{
    int resident = 0;
    const uint32_t stage = ldsPerBlock ? STAGE_SMALL : STAGE_NONE;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, reinterpret_cast<const void*>(kernelFor(L, waves, stage, kind)), 256, ldsPerBlock)
            != hipSuccess
        || resident <= 0) {
        hipGetLastError();
        resident = 1;
    }
    return static_cast<uint32_t>(resident);
}

// queue capacity of a run's kernel: 64, and in the 7-wave kernels the
// largest of 48, 40 and 32 photons with which 7 blocks stay resident on a
// CU beside the staged tables (48 when none does, and always 48 on palette
// worlds; whole body, 8 materials: 40, 1.87e9 -> 1.94e9 histories/s)
uint32_t queueCapFp32(int L, uint32_t waves, uint32_t tablesBytes, uint32_t kind)
{
    if (waves < 7)
        return 64;
    if (kind == WORLD_PALETTE8)
        return 48;
    for (uint32_t cap : { 48u, 40u, 32u })
        if (residentBlocksFp32(L, waves, tablesBytes + WAVES_PER_BLOCK * queueBytes(cap), kind) >= 7)
            return cap;
    return 48;
}

// staging level of a run: the requested one (lowered until it fits), or by
// default the attenuation coefficients too when they cost no resident block,
// else the sampling tables when they cost at most one, else none
uint32_t stageLevelFp32(const dxmc_run_cfg& cfg, const Tables<float>& t, uint32_t waves, uint32_t kind)
{
    // a block gets at most 64 KiB of dynamic LDS
    constexpr uint32_t ldsPerBlock = 64 * 1024;
    auto bytes = [&](uint32_t level) {
        const uint32_t tb = ldsBytes(t, level, cfg.low_energy_correction);
        return uint64_t(tb) + WAVES_PER_BLOCK * queueBytes(queueCapFp32(cfg.low_energy_correction, waves, tb, kind));
    };
    auto fits = [&](uint32_t level) { return bytes(level) <= ldsPerBlock; };
    if (cfg.lds_tables >= STAGE_NONE && cfg.lds_tables <= STAGE_ALL) {
        uint32_t level = cfg.lds_tables;
        while (level > STAGE_NONE && !fits(level))
            --level;
        return level;
    }
    // The sampling tables are worth a resident block (whole body, 8
    // materials: 1.04e9 histories/s with them and 4 blocks per CU against
    // 0.98e9 without and 5); the attenuation coefficients are not worth two
    // (CT spiral 2.84e9 vs 2.50e9, DX 4.07e9 vs 3.32e9 with 5 and 3 blocks)
    // and are staged only when they cost none.
    const uint32_t best = residentBlocksFp32(cfg.low_energy_correction, waves, bytes(STAGE_NONE), kind);
    auto resident = [&](uint32_t level) { return residentBlocksFp32(cfg.low_energy_correction, waves, static_cast<uint32_t>(bytes(level)), kind); };
    if (fits(STAGE_ALL) && resident(STAGE_ALL) >= best)
        return STAGE_ALL;
    if (fits(STAGE_SMALL) && resident(STAGE_SMALL) + 1 >= best)
        return STAGE_SMALL;
    return STAGE_NONE;
}

void launchFp32(dxmc_ctx* c, const dxmc_run_cfg& cfg, const World<float>& w, const Tables<float>& t, const Exposure<float>* exps,
    const uint64_t* expStart, uint32_t nExp, uint64_t hb, uint64_t he, const Tally& tally, Counters* ctr, unsigned long long* historyCounter)
{
    // the work counter is cleared by a kernel, not a memset, so the launch
    // sequence stays valid under stream capture
    auto* nx = historyCounter;
    hipLaunchKernelGGL(clearCounter, dim3(1), dim3(1), 0, c->stream, nx);
    const uint32_t block = 256;
    World<float> world = w;
    Tables<float> tables = t;
    // the attenuation bounds pay where they also spare the record read
    // (worlds with codes); elsewhere the extra lookup costs more than the
    // evaluations it skips in diverged waves
    if ((cfg.flags & DXMC_RUN_EXACT_STEPS) || world.codeBytes == 0) {
        world.codeBytes = 0;
        tables.lutSumBound = nullptr;
    }
    // the block code ranges spare most steps the code gather, and pay
    // wherever that gather misses the caches or 7 waves per SIMD hide the
    // block entry's loads (measured: CT spiral +3 % at L = 1, whole body
    // +39 %); at L = 2, with 5 waves per SIMD, the longer chain of a step
    // whose draw falls in the block's band costs more than a code gather
    // from the Infinity Cache saves (CT spiral -6.5 %, its 79 MB codes
    // cached; whole body +13 %, its 629 MB not)
    constexpr uint64_t INFINITY_CACHE_BYTES = uint64_t(256) << 20;
    const uint64_t codeArrayBytes = uint64_t(world.nx) * world.ny * world.nz * world.codeBytes;
    const bool blocksOn = world.codeBytes == 1 && world.blocks
        && (cfg.block_bounds == 1 || (cfg.block_bounds == 0 && !(cfg.low_energy_correction == 2 && codeArrayBytes <= INFINITY_CACHE_BYTES)));
    if (!blocksOn)
        world.blocks = nullptr;
    uint32_t kind = worldKind(world, tables);
    if (kind == WORLD_CODES8 && !world.blocks && cfg.low_energy_correction < 2)
        kind = WORLD_ANY;
    const uint32_t waves = wavesFp32(cfg.waves_per_simd, cfg.low_energy_correction, kind);
    const uint32_t stage = stageLevelFp32(cfg, tables, waves, kind);
    const uint32_t tablesBytes = ldsBytes(tables, stage, cfg.low_energy_correction);
    const uint32_t cap = queueCapFp32(cfg.low_energy_correction, waves, tablesBytes, kind);
    const uint32_t lds = tablesBytes + WAVES_PER_BLOCK * queueBytes(cap);
    // default: as many blocks as stay resident on a CU (registers and LDS)
    const uint32_t perCU = cfg.blocks_per_cu ? cfg.blocks_per_cu : residentBlocksFp32(cfg.low_energy_correction, waves, lds, kind);
    const uint32_t regen = cfg.regeneration_threshold ? std::min<uint32_t>(cfg.regeneration_threshold, 64u) : 4u;
    // pending lanes before an interaction pass: at L = 2 a pass costs more
    // (the impulse-approximation sampler), so it waits for more lanes (CT
    // spiral 2.39e9 at 20, 2.49e9 at 28, 2.52e9 at 32, 2.45e9 at 40; at
    // L = 1, 32 cost 3 % and 24 0.5 %)
    const uint32_t interact = cfg.interaction_threshold ? std::min<uint32_t>(cfg.interaction_threshold, 64u) : cfg.low_energy_correction == 2 ? 32u : 20u;
    // histories a wave claims per atomic on the global counter
    const uint32_t pool = cfg.queue_depth ? std::clamp<uint32_t>(cfg.queue_depth, 64u, 1u << 16) : 512u;
    uint64_t grid = static_cast<uint64_t>(c->numCU) * perCU;
    const uint64_t lanesNeeded = (he - hb + 63) / 64 * 64;
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, (lanesNeeded + block - 1) / block));
    hipLaunchKernelGGL(kernelFor(cfg.low_energy_correction, waves, stage, kind), dim3(static_cast<uint32_t>(grid)), dim3(block), lds, c->stream, world, tables,
        exps, expStart, nExp, hb, he, cfg.seed, cfg.first_history, tally, ctr, nx, regen, interact, pool, cap);
}
}
