// World upload with its voxel passes on the device.
//
// A world arrives as plain host arrays (material index, density, optional
// measurement flags). After one host->device copy of each array, kernels do
// every pass over the voxels:
//   scanWorld<D>    largest material index (World::validate, reference
//                   world.hpp:182-227), per-material maximum density (the
//                   majorant's input, reference attenuationinterpolator.hpp
//                   :52-58), whether any measurement flag is set, and the
//                   FP64 density copy the normalisation reads
//   paletteInsert   distinct FP32 voxel records (density bits, material |
//                   measurement << 24) into an open-addressing hash table
//   paletteCompact  palette index of each distinct record
//   paletteIndex<I> 1- or 2-byte palette index of every voxel
//   packVoxels      8-byte records when there are more than 65536
//                   distinct records
//   boundCodes<I>   for those worlds, a 1- or 2-byte code per voxel naming
//                   its material and a density interval (Woodcock decisions
//                   settled from the interval at most steps)
//   blockCodes      with 1-byte codes, the code range of every block of
//                   8^3 to 64^3 voxels that holds one material (decisions
//                   settled from the block's interval without reading the
//                   voxel's code)
// The host reads back a few hundred bytes (maxima, flags, distinct count),
// so the cost of an upload is the copy and a few HBM passes.
#include "dxmc_hip.h"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

using namespace dxmc_backend;

namespace {

constexpr uint32_t SCAN_BLOCK = 256; // one thread per material slot in scanWorld
constexpr uint32_t SCAN_GRID = 4096; // grid-stride blocks of the voxel passes
constexpr uint32_t INSERT_GRID = 1024; // fewer in flight: bounds the table's fill
constexpr uint32_t PALETTE_LIMIT = 65536;
constexpr uint32_t HASH_BITS = 20;
constexpr uint32_t HASH_SLOTS = 1u << HASH_BITS;
constexpr uint32_t HASH_PROBES = 256; // longer probe sequences fall back to 8-byte records
constexpr uint32_t COMPACT_BLOCK = 1024;
constexpr unsigned long long EMPTY = ~0ull; // not a record key: a tag is at most 0xff0000ff

uint32_t gridFor(uint64_t n, uint32_t block, uint32_t cap) { return static_cast<uint32_t>(std::min<uint64_t>((n + block - 1) / block, cap)); }

// Order-preserving image of a double: unsigned comparison of images is the
// numeric order. 0 is below the image of every number (only a negative NaN
// maps to 0, and NaNs are skipped).
__host__ __device__ inline unsigned long long orderKey(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

double fromOrderKey(unsigned long long k)
{
    if (k == 0)
        return 0.0;
    return __builtin_bit_cast(double, (k >> 63) ? (k & ~(1ull << 63)) : ~k);
}

__device__ inline unsigned long long recordKey(const uint8_t* mat, const double* dens, const uint8_t* meas, uint64_t i)
{
    const uint32_t tag = static_cast<uint32_t>(mat[i]) | (meas ? static_cast<uint32_t>(meas[i]) << 24 : 0u);
    return static_cast<unsigned long long>(tag) << 32 | __float_as_uint(static_cast<float>(dens[i]));
}

__device__ inline uint32_t hashSlot(unsigned long long k) { return static_cast<uint32_t>((k * 0x9E3779B97F4A7C15ull) >> (64 - HASH_BITS)); }

// flags[0]: largest material index; flags[1]: any measurement flag set.
// maxKey[m]: order key of material m's largest density, minKeyInv[m]: the
// complement of the order key of its smallest (0 for both: no voxel).
template <typename D>
__global__ __launch_bounds__(SCAN_BLOCK) void scanWorld(const uint8_t* __restrict__ mat, const D* __restrict__ dens, double* __restrict__ dens64,
    const uint8_t* __restrict__ meas, uint64_t n, unsigned long long* __restrict__ maxKey, unsigned long long* __restrict__ minKeyInv,
    unsigned int* __restrict__ flags)
{
    __shared__ unsigned long long key[SCAN_BLOCK], keyInv[SCAN_BLOCK];
    __shared__ unsigned int blockMat, blockMeas;
    key[threadIdx.x] = 0;
    keyInv[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        blockMat = 0;
        blockMeas = 0;
    }
    __syncthreads();
    uint32_t lastMat = SCAN_BLOCK, largestMat = 0, anyMeas = 0;
    unsigned long long lastKey = 0, lastKeyInv = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * SCAN_BLOCK;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK + threadIdx.x; i < n; i += stride) {
        const double v = static_cast<double>(dens[i]);
        if constexpr (!std::is_same_v<D, double>)
            dens64[i] = v;
        const uint32_t m = mat[i];
        largestMat = max(largestMat, m);
        if (meas)
            anyMeas |= meas[i];
        if (v == v) {
            const unsigned long long k = orderKey(v);
            // skip the LDS atomics when this thread already posted a larger
            // and a smaller value for the same material
            if (m != lastMat) {
                atomicMax(&key[m], k);
                atomicMax(&keyInv[m], ~k);
                lastMat = m;
                lastKey = k;
                lastKeyInv = ~k;
            } else {
                if (k > lastKey) {
                    atomicMax(&key[m], k);
                    lastKey = k;
                }
                if (~k > lastKeyInv) {
                    atomicMax(&keyInv[m], ~k);
                    lastKeyInv = ~k;
                }
            }
        }
    }
    atomicMax(&blockMat, largestMat);
    if (anyMeas)
        atomicOr(&blockMeas, 1u);
    __syncthreads();
    if (key[threadIdx.x]) {
        atomicMax(&maxKey[threadIdx.x], key[threadIdx.x]);
        atomicMax(&minKeyInv[threadIdx.x], keyInv[threadIdx.x]);
    }
    if (threadIdx.x == 0) {
        atomicMax(&flags[0], blockMat);
        if (blockMeas)
            atomicOr(&flags[1], 1u);
    }
}

// state[0]: distinct records inserted; state[1]: overflow (too many
// records, or a probe sequence longer than HASH_PROBES). Every thread stops
// at its next voxel once state[1] is set.
__global__ __launch_bounds__(SCAN_BLOCK) void paletteInsert(const uint8_t* __restrict__ mat, const double* __restrict__ dens,
    const uint8_t* __restrict__ meas, uint64_t n, unsigned long long* keys, unsigned int* state)
{
    unsigned long long last = EMPTY;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * SCAN_BLOCK;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK + threadIdx.x; i < n; i += stride) {
        if (__hip_atomic_load(&state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
        const unsigned long long k = recordKey(mat, dens, meas, i);
        if (k == last)
            continue;
        last = k;
        uint32_t h = hashSlot(k);
        bool placed = false;
        for (uint32_t p = 0; p < HASH_PROBES; ++p, h = (h + 1) & (HASH_SLOTS - 1)) {
            const unsigned long long cur = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == k) {
                placed = true;
                break;
            }
            if (cur != EMPTY)
                continue;
            const unsigned long long prev = atomicCAS(&keys[h], EMPTY, k);
            if (prev == EMPTY) {
                if (atomicAdd(&state[0], 1u) >= PALETTE_LIMIT)
                    atomicOr(&state[1], 1u);
                placed = true;
                break;
            }
            if (prev == k) {
                placed = true;
                break;
            }
        }
        if (!placed) {
            atomicOr(&state[1], 1u);
            return;
        }
    }
}

// One block: thread t owns slots t, t + 1024, ...; ranks follow (t, slot).
__global__ __launch_bounds__(COMPACT_BLOCK) void paletteCompact(const unsigned long long* __restrict__ keys, uint32_t* __restrict__ rank,
    uint2* __restrict__ palette)
{
    __shared__ uint32_t count[COMPACT_BLOCK];
    const uint32_t t = threadIdx.x;
    uint32_t c = 0;
    for (uint32_t j = t; j < HASH_SLOTS; j += COMPACT_BLOCK)
        c += keys[j] != EMPTY;
    count[t] = c;
    __syncthreads();
    for (uint32_t off = 1; off < COMPACT_BLOCK; off <<= 1) {
        const uint32_t v = t >= off ? count[t - off] : 0u;
        __syncthreads();
        count[t] += v;
        __syncthreads();
    }
    uint32_t r = count[t] - c;
    for (uint32_t j = t; j < HASH_SLOTS; j += COMPACT_BLOCK) {
        const unsigned long long k = keys[j];
        if (k != EMPTY) {
            rank[j] = r;
            palette[r] = make_uint2(static_cast<uint32_t>(k), static_cast<uint32_t>(k >> 32));
            ++r;
        }
    }
}

// Every key is in the table (no overflow), within HASH_PROBES of its slot.
// rank[] is zero-filled first, so even a miss would yield index 0.
template <typename I>
__global__ __launch_bounds__(SCAN_BLOCK) void paletteIndex(const uint8_t* __restrict__ mat, const double* __restrict__ dens,
    const uint8_t* __restrict__ meas, uint64_t n, const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ rank, I* __restrict__ out)
{
    unsigned long long last = EMPTY;
    I index = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * SCAN_BLOCK;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK + threadIdx.x; i < n; i += stride) {
        const unsigned long long k = recordKey(mat, dens, meas, i);
        if (k != last) {
            uint32_t h = hashSlot(k);
            for (uint32_t p = 1; p < HASH_PROBES && keys[h] != k; ++p)
                h = (h + 1) & (HASH_SLOTS - 1);
            index = static_cast<I>(rank[h]);
            last = k;
        }
        out[i] = index;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void packVoxels(const uint8_t* __restrict__ mat, const double* __restrict__ dens,
    const uint8_t* __restrict__ meas, uint2* __restrict__ vox, uint64_t n)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * SCAN_BLOCK;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK + threadIdx.x; i < n; i += stride) {
        uint2 v;
        v.x = __float_as_uint(static_cast<float>(dens[i]));
        v.y = static_cast<uint32_t>(mat[i]) | (meas ? static_cast<uint32_t>(meas[i]) << 24 : 0u);
        vox[i] = v;
    }
}

// Code of every voxel of an 8-byte-record world: the bin of its FP32
// density among the B bins of its material, base[m] + k, checked against
// the interval it names (bounds[code].lo <= density <= bounds[code].hi); a
// voxel for which that fails sets *failed and the codes are not used.
template <typename I>
__global__ __launch_bounds__(SCAN_BLOCK) void boundCodes(const uint2* __restrict__ records, uint64_t n, const dxmc_dev::DensityBound* __restrict__ bounds,
    const uint32_t* __restrict__ base, const double* __restrict__ lower, const double* __restrict__ scale, uint32_t bins, I* __restrict__ codes,
    unsigned int* failed)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * SCAN_BLOCK;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * SCAN_BLOCK + threadIdx.x; i < n; i += stride) {
        const uint2 r = records[i];
        const float d = __uint_as_float(r.x);
        const uint32_t m = r.y & 0xffu;
        const double x = (static_cast<double>(d) - lower[m]) * scale[m];
        int k = x > 0 ? static_cast<int>(min(x, static_cast<double>(bins - 1))) : 0;
        const dxmc_dev::DensityBound* b = bounds + base[m];
        for (uint32_t guard = 0; guard < bins && k > 0 && d < b[k].lo; ++guard)
            --k;
        for (uint32_t guard = 0; guard < bins && k + 1 < static_cast<int>(bins) && d > b[k].hi; ++guard)
            ++k;
        if (!(b[k].lo <= d && d <= b[k].hi && b[k].material == m))
            atomicOr(failed, 1u);
        codes[i] = static_cast<I>(base[m] + static_cast<uint32_t>(k));
    }
}

// Block code of every block of 2^shift voxels per edge of a 1-byte code
// world (the blocks of the last row, column and slice clipped to the grid),
// one workgroup per block: smallest code | largest code << 8 when every
// voxel of the block has the same material, else BLOCK_MIXED. A material's
// codes are its density bins in ascending order, so the interval [lo of the
// smallest, hi of the largest] holds every density of the block.
__global__ __launch_bounds__(SCAN_BLOCK) void blockCodes(const uint8_t* __restrict__ codes, const dxmc_dev::DensityBound* __restrict__ bounds,
    uint32_t nx, uint32_t ny, uint32_t nz, uint32_t shift, uint32_t bnx, uint32_t bny, uint64_t nBlocks, uint16_t* __restrict__ blocks)
{
    __shared__ unsigned int lo, hi, mixed;
    const uint32_t E = 1u << shift;
    for (uint64_t b = blockIdx.x; b < nBlocks; b += gridDim.x) {
        const uint32_t bx = static_cast<uint32_t>(b % bnx);
        const uint32_t by = static_cast<uint32_t>((b / bnx) % bny);
        const uint32_t bz = static_cast<uint32_t>(b / (static_cast<uint64_t>(bnx) * bny));
        const uint32_t x0 = bx * E, y0 = by * E, z0 = bz * E;
        const uint32_t w = min(x0 + E, nx) - x0, h = min(y0 + E, ny) - y0, d = min(z0 + E, nz) - z0;
        const uint32_t material = bounds[codes[(static_cast<uint64_t>(z0) * ny + y0) * nx + x0]].material;
        if (threadIdx.x == 0) {
            lo = 255u;
            hi = 0u;
            mixed = 0u;
        }
        __syncthreads();
        uint32_t l = 255u, u = 0u, m = 0u;
        const uint32_t plane = w * h;
        for (uint32_t i = threadIdx.x; i < plane * d; i += SCAN_BLOCK) {
            const uint32_t z = z0 + i / plane, r = i % plane;
            const uint32_t c = codes[(static_cast<uint64_t>(z) * ny + y0 + r / w) * nx + x0 + r % w];
            m |= bounds[c].material != material ? 1u : 0u;
            l = min(l, c);
            u = max(u, c);
        }
        atomicMin(&lo, l);
        atomicMax(&hi, u);
        if (m)
            atomicOr(&mixed, 1u);
        __syncthreads();
        if (threadIdx.x == 0)
            blocks[b] = mixed ? dxmc_dev::BLOCK_MIXED : static_cast<uint16_t>(lo | hi << 8);
        __syncthreads();
    }
}

// Scan results read back by the host.
struct ScanResult {
    unsigned long long maxKey[SCAN_BLOCK];
    unsigned long long minKeyInv[SCAN_BLOCK];
    unsigned int flags[2];
    unsigned int palette[2]; // distinct records, overflow
    unsigned int codesFailed, pad;
};

float floatBelow(double v)
{
    float f = static_cast<float>(v);
    if (static_cast<double>(f) > v)
        f = std::nextafter(f, -HUGE_VALF);
    return f;
}
float floatAbove(double v)
{
    float f = static_cast<float>(v);
    if (static_cast<double>(f) < v)
        f = std::nextafter(f, HUGE_VALF);
    return f;
}

// Bounded-density codes of an 8-byte-record world without measurement
// flags. With a material map of up to 64 entries the codes are 1 byte: the
// map's materials share codes 0-254 equally (code 255 is kept for mixed
// blocks), code = material * bins + bin, so a kernel finds the material of
// a code by a multiply and a shift (codeMul). Larger maps get 2-byte codes,
// the 65536 shared equally by the materials present. Each material's
// [min, max] density range is cut into its bins, each widened outwards to
// FP32 values. Leaves codeBytes 0 when a voxel's density does not fall in
// the bin computed for it. With 1-byte codes the block codes follow.
int buildBoundCodes(dxmc_ctx* c, dxmc_world* w, ScanResult* dev, const ScanResult& host)
{
    std::vector<uint32_t> present;
    for (uint32_t m = 0; m < w->nMaterials; ++m)
        if (host.maxKey[m])
            present.push_back(m);
    if (present.empty() || w->hasMeasurement)
        return DXMC_OK;
    const uint32_t codeBytes = w->nMaterials <= 64 ? 1u : 2u;
    const uint32_t nCodes = codeBytes == 1 ? 256u : 65536u;
    const uint32_t bins = codeBytes == 1 ? (nCodes - 1) / w->nMaterials : nCodes / static_cast<uint32_t>(present.size());
    std::vector<dxmc_dev::DensityBound> table(nCodes, dxmc_dev::DensityBound { 1.0f, 0.0f, 0xffffffffu, 0u });
    if (codeBytes == 1) // the smallest code of mixed blocks, which the kernels exclude before using it
        table[dxmc_dev::MIXED_CODE] = dxmc_dev::DensityBound { 0.0f, 0.0f, 0u, 0u };
    std::vector<uint32_t> base(SCAN_BLOCK, 0);
    std::vector<double> lower(SCAN_BLOCK, 0.0), scale(SCAN_BLOCK, 0.0);
    for (uint32_t j = 0; j < present.size(); ++j) {
        const uint32_t m = present[j];
        // FP32 records hold float(density): their range is the rounded range
        const float a = static_cast<float>(fromOrderKey(~host.minKeyInv[m]));
        const float b = static_cast<float>(fromOrderKey(host.maxKey[m]));
        base[m] = (codeBytes == 1 ? m : j) * bins;
        lower[m] = a;
        scale[m] = b > a ? bins / (static_cast<double>(b) - a) : 0.0;
        for (uint32_t k = 0; k < bins; ++k) {
            auto& e = table[base[m] + k];
            const double e0 = a + (static_cast<double>(b) - a) * k / bins;
            const double e1 = a + (static_cast<double>(b) - a) * (k + 1) / bins;
            e.lo = k == 0 ? a : floatBelow(e0);
            e.hi = k + 1 == bins ? b : floatAbove(e1);
            e.material = m;
        }
    }
    DeviceBuffer dBase(c), dLower(c), dScale(c);
    int st = DXMC_OK;
    if ((st = c->alloc(&w->bounds, nCodes * sizeof(dxmc_dev::DensityBound))) || (st = c->alloc(&w->codes, w->n * codeBytes))
        || (st = c->copy(w->bounds, table.data(), nCodes * sizeof(dxmc_dev::DensityBound))) || (st = dBase.upload(base.data(), base.size() * 4))
        || (st = dLower.upload(lower.data(), lower.size() * 8)) || (st = dScale.upload(scale.data(), scale.size() * 8)))
        return st;
    const auto* rec = static_cast<const uint2*>(w->voxels);
    const auto* bnd = static_cast<const dxmc_dev::DensityBound*>(w->bounds);
    const dim3 grid(gridFor(w->n, SCAN_BLOCK, SCAN_GRID)), block(SCAN_BLOCK);
    if (codeBytes == 1)
        hipLaunchKernelGGL(boundCodes<uint8_t>, grid, block, 0, c->stream, rec, w->n, bnd, static_cast<const uint32_t*>(dBase.p),
            static_cast<const double*>(dLower.p), static_cast<const double*>(dScale.p), bins, static_cast<uint8_t*>(w->codes), &dev->codesFailed);
    else
        hipLaunchKernelGGL(boundCodes<uint16_t>, grid, block, 0, c->stream, rec, w->n, bnd, static_cast<const uint32_t*>(dBase.p),
            static_cast<const double*>(dLower.p), static_cast<const double*>(dScale.p), bins, static_cast<uint16_t*>(w->codes), &dev->codesFailed);
    unsigned int failed = 1;
    if ((st = c->check(hipGetLastError())) || (st = c->download(&failed, &dev->codesFailed, sizeof(failed))))
        return st;
    if (failed) {
        c->release(w->codes);
        c->release(w->bounds);
        w->codes = w->bounds = nullptr;
        return DXMC_OK;
    }
    w->codeBytes = codeBytes;
    w->nCodes = nCodes;
    if (codeBytes != 1)
        return DXMC_OK;
    // (code * codeMul) >> CODE_MUL_SHIFT = code / bins for every code below
    // 256: codeMul exceeds 2^20 / bins by less than 1, which adds less than
    // 256 / 2^20 to a quotient whose fraction is at most 1 - 1/bins
    w->codeMul = ((uint32_t(1) << dxmc_dev::CODE_MUL_SHIFT) + bins - 1) / bins;
    // block edge: the smallest of 8, 16, 32 and 64 voxels whose table fits
    // 64 KiB, so that it stays in the caches (CT spiral, 512x512x300: 16,
    // 39 KB, 4.27e9 -> 4.44e9 histories/s against 8; whole body,
    // 1024x1024x600: 32, 39 KB, 1.95e9 -> 2.09e9; a larger block is more
    // often mixed, but its entry is far more often cached)
    uint32_t shift = dxmc_dev::BLOCK_SHIFT_MIN, bd[3];
    for (;; ++shift) {
        const uint32_t E = 1u << shift;
        for (int i = 0; i < 3; ++i)
            bd[i] = (w->dims[i] + E - 1) / E;
        if (uint64_t(bd[0]) * bd[1] * bd[2] * sizeof(uint16_t) <= dxmc_dev::BLOCK_TABLE_BYTES || shift == dxmc_dev::BLOCK_SHIFT_MAX)
            break;
    }
    const uint64_t nBlocks = uint64_t(bd[0]) * bd[1] * bd[2];
    if ((st = c->alloc(&w->blocks, nBlocks * sizeof(uint16_t))))
        return st;
    hipLaunchKernelGGL(blockCodes, dim3(gridFor(nBlocks, 1, SCAN_GRID * 4)), block, 0, c->stream, static_cast<const uint8_t*>(w->codes), bnd,
        w->dims[0], w->dims[1], w->dims[2], shift, bd[0], bd[1], nBlocks, static_cast<uint16_t*>(w->blocks));
    if ((st = c->check(hipGetLastError())))
        return st;
    for (int i = 0; i < 3; ++i)
        w->blockDims[i] = bd[i];
    w->blockShift = shift;
    return DXMC_OK;
}

int fillVoxelRecords(dxmc_ctx* c, dxmc_world* w, ScanResult* dev, ScanResult& host)
{
    const uint64_t n = w->n;
    const auto* mat = static_cast<const uint8_t*>(w->material);
    const auto* dens = static_cast<const double*>(w->density);
    const auto* meas = static_cast<const uint8_t*>(w->measurement);
    DeviceBuffer keys(c), rank(c);
    int st = DXMC_OK;
    if ((st = keys.alloc(HASH_SLOTS * sizeof(unsigned long long))) || (st = rank.zeros(HASH_SLOTS * sizeof(uint32_t)))
        || (st = c->check(hipMemsetAsync(keys.p, 0xff, HASH_SLOTS * sizeof(unsigned long long), c->stream))))
        return st;
    auto* k = static_cast<unsigned long long*>(keys.p);
    hipLaunchKernelGGL(paletteInsert, dim3(gridFor(n, SCAN_BLOCK, INSERT_GRID)), dim3(SCAN_BLOCK), 0, c->stream, mat, dens, meas, n, k, dev->palette);
    if ((st = c->check(hipGetLastError())) || (st = c->download(host.palette, dev->palette, sizeof(host.palette))))
        return st;
    const uint32_t distinct = host.palette[0];
    if (host.palette[1] == 0 && distinct <= PALETTE_LIMIT) {
        w->voxelBytes = distinct <= 256 ? 1u : 2u;
        w->nPalette = distinct;
        if ((st = c->alloc(&w->palette, distinct * sizeof(uint2))) || (st = c->alloc(&w->voxels, n * w->voxelBytes)))
            return st;
        auto* r = static_cast<uint32_t*>(rank.p);
        hipLaunchKernelGGL(paletteCompact, dim3(1), dim3(COMPACT_BLOCK), 0, c->stream, k, r, static_cast<uint2*>(w->palette));
        if (w->voxelBytes == 1)
            hipLaunchKernelGGL(paletteIndex<uint8_t>, dim3(gridFor(n, SCAN_BLOCK, SCAN_GRID)), dim3(SCAN_BLOCK), 0, c->stream, mat, dens, meas, n, k, r,
                static_cast<uint8_t*>(w->voxels));
        else
            hipLaunchKernelGGL(paletteIndex<uint16_t>, dim3(gridFor(n, SCAN_BLOCK, SCAN_GRID)), dim3(SCAN_BLOCK), 0, c->stream, mat, dens, meas, n, k, r,
                static_cast<uint16_t*>(w->voxels));
    } else {
        w->voxelBytes = 8;
        w->nPalette = 0;
        if ((st = c->alloc(&w->voxels, n * sizeof(uint2))))
            return st;
        hipLaunchKernelGGL(packVoxels, dim3(gridFor(n, SCAN_BLOCK, SCAN_GRID)), dim3(SCAN_BLOCK), 0, c->stream, mat, dens, meas,
            static_cast<uint2*>(w->voxels), n);
        if ((st = c->check(hipGetLastError())) || (st = buildBoundCodes(c, w, dev, host)))
            return st;
    }
    if ((st = c->check(hipGetLastError())))
        return st;
    // the temporaries are released on the stream after these kernels
    return c->check(hipStreamSynchronize(c->stream));
}

} // namespace

extern "C" {

int dxmc_hip_world_upload(dxmc_ctx* c, const dxmc_world_desc* d, dxmc_world** out)
{
    if (!c || !d || !out)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
    *out = nullptr;
    const uint64_t n = d->dims[0] * d->dims[1] * d->dims[2];
    if (n == 0 || !d->material || (!d->density_f32 == !d->density_f64) || d->n_materials == 0 || d->n_materials > 256)
        return c->fail(DXMC_ERR_ARGUMENT, "world_upload: need dims, material, exactly one density array and 1..256 materials");
    if (d->dims[0] > 0xffffffffULL || d->dims[1] > 0xffffffffULL || d->dims[2] > 0xffffffffULL)
        return c->fail(DXMC_ERR_ARGUMENT, "world_upload: dimension exceeds 2^32-1");
    for (int i = 0; i < 3; ++i)
        if (!(d->spacing[i] > 0) || !(d->extent_safe[2 * i] < d->extent_safe[2 * i + 1]))
            return c->fail(DXMC_ERR_ARGUMENT, "world_upload: invalid spacing or extent");
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
    auto fail = [&](int st) {
        dxmc_hip_world_destroy(c, w);
        return st;
    };
    DeviceBuffer dens32(c), scan(c);
    int st = DXMC_OK;
    if ((st = c->alloc(&w->material, n)) || (st = c->alloc(&w->density, n * sizeof(double))) || (d->measurement && (st = c->alloc(&w->measurement, n)))
        || (st = scan.zeros(sizeof(ScanResult))))
        return fail(st);
    if ((st = c->copy(w->material, d->material, n)) || (d->measurement && (st = c->copy(w->measurement, d->measurement, n))))
        return fail(st);
    auto* res = static_cast<ScanResult*>(scan.p);
    const auto* mat = static_cast<const uint8_t*>(w->material);
    const auto* meas = static_cast<const uint8_t*>(w->measurement);
    auto* dens64 = static_cast<double*>(w->density);
    const dim3 grid(gridFor(n, SCAN_BLOCK, SCAN_GRID)), block(SCAN_BLOCK);
    if (d->density_f64) {
        if ((st = c->copy(w->density, d->density_f64, n * sizeof(double))))
            return fail(st);
        hipLaunchKernelGGL(scanWorld<double>, grid, block, 0, c->stream, mat, dens64, dens64, meas, n, res->maxKey, res->minKeyInv, res->flags);
    } else {
        if ((st = dens32.upload(d->density_f32, n * sizeof(float))))
            return fail(st);
        hipLaunchKernelGGL(scanWorld<float>, grid, block, 0, c->stream, mat, static_cast<const float*>(dens32.p), dens64, meas, n, res->maxKey,
            res->minKeyInv, res->flags);
    }
    ScanResult host {};
    if ((st = c->check(hipGetLastError())) || (st = c->download(&host, res, sizeof(ScanResult))))
        return fail(st);
    if (host.flags[0] >= d->n_materials)
        return fail(c->fail(DXMC_ERR_ARGUMENT, "world_upload: material index outside the material map"));
    w->hasMeasurement = host.flags[1] != 0;
    if (!w->hasMeasurement && w->measurement) {
        c->release(w->measurement);
        w->measurement = nullptr;
    }
    // the reference starts each maximum at 0 (attenuationinterpolator.hpp:52-58)
    w->maxDensity.resize(d->n_materials);
    for (uint32_t m = 0; m < d->n_materials; ++m)
        w->maxDensity[m] = std::max(0.0, fromOrderKey(host.maxKey[m]));
    if ((st = fillVoxelRecords(c, w, res, host)))
        return fail(st);
    w->bytes = n * (1 + sizeof(double) + w->voxelBytes + (w->hasMeasurement ? 1 : 0) + w->codeBytes) + w->nPalette * sizeof(uint2)
        + w->nCodes * sizeof(dxmc_dev::DensityBound) + uint64_t(w->blockDims[0]) * w->blockDims[1] * w->blockDims[2] * sizeof(uint16_t);
    *out = w;
    return DXMC_OK;
}

int dxmc_hip_world_destroy(dxmc_ctx* c, dxmc_world* w)
{
    if (!c || !w)
        return DXMC_ERR_ARGUMENT;
    DXMC_LOCK(c);
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
    c->release(w->codes);
    c->release(w->bounds);
    c->release(w->blocks);
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

int dxmc_hip_world_max_density(const dxmc_world* w, double* out, uint32_t n)
{
    if (!w || !out || n != w->nMaterials)
        return DXMC_ERR_ARGUMENT;
    std::copy(w->maxDensity.begin(), w->maxDensity.end(), out);
    return DXMC_OK;
}

} // extern "C"
