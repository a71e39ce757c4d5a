// Device-side photon transport physics, templated on the arithmetic type T.
//
// Restates the reference hot path (include/dxmc/transport.hpp) for the GPU:
//   sampling a photon from an exposure      exposure.hpp:280-304
//   spectrum alias sampling                 dxmcrandom.hpp:211-216, 322-326
//   bowtie / heel weights                   beamfilters.hpp:150-190, 514-538
//   ray to world entry                      transport.hpp:702-728
//   attenuation LUT and majorant            attenuationinterpolator.hpp:207-248
//   photoelectric absorption                transport.hpp:216-263
//   Rayleigh scatter (Thomson / RITA)       transport.hpp:264-298, dxmcrandom.hpp:483-500
//   Compton scatter (KN, scatter factor)    transport.hpp:300-340, interpolation.hpp:169-176
//   Compton impulse approximation           transport.hpp:342-483
//   interaction dispatch, forced dispatch   transport.hpp:523-638
//   Woodcock delta tracking                 transport.hpp:640-700
//   direction deflection                    vectormath.hpp:83-99, 197-221
// For T = double (the plain method) every expression keeps the reference's
// operation order so the CPU oracle and this code round identically except
// where the math library differs (log, pow, sin, cos, acos, exp); the file
// including this header is compiled without FMA contraction. For T = float
// the attenuation tables are evaluated in base 2 (exp2/log2 with host
// re-based coefficients) instead of pow(10, .)/log10.
//
// Bounds: indices that the reference could push past an array end (voxel
// index from a rounded position, the unclamped majorant index, a RITA draw
// equal to the last CDF node) are clamped to the last valid element.
#pragma once
#include "rng.hpp"
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dxmc_dev {

template <typename T>
struct Consts;
// the values of std::numbers::{pi, sqrt2, sqrt3, inv_sqrtpi, e}_v<T>
template <>
struct Consts<double> {
    static constexpr double PI = 3.141592653589793238462643383279502884;
    static constexpr double SQRT2 = 1.414213562373095048801688724209698079;
    static constexpr double SQRT3 = 1.732050807568877293527446341505872367;
    static constexpr double INV_SQRTPI = 0.564189583547756286948079451560772586;
    static constexpr double E = 2.718281828459045235360287471352662498;
};
template <>
struct Consts<float> {
    static constexpr float PI = 3.141592653589793238462643383279502884f;
    static constexpr float SQRT2 = 1.414213562373095048801688724209698079f;
    static constexpr float SQRT3 = 1.732050807568877293527446341505872367f;
    static constexpr float INV_SQRTPI = 0.564189583547756286948079451560772586f;
    static constexpr float E = 2.718281828459045235360287471352662498f;
};

template <typename T>
struct Particle {
    T pos[3];
    T dir[3];
    T energy;
    T weight;
};

template <typename T>
struct Shell {
    T binding, electrons, j0, photoProb, yield;
    T lineProb[3];
    T lineEnergy[3];
};

template <typename T>
struct Heel {
    T energyStart, energyStep, angleStart, angleStep;
    uint32_t energySize, angleSize, weightOffset, pad;
};

// Flat device tables (one copy per precision). The LUT grid and linear
// step/energy are log10 E in both copies; for T = float each coefficient pair
// is (log2 mu at the segment's upper node, slope).
template <typename T>
struct Tables {
    uint32_t nMaterials;
    uint32_t lutNodes;
    uint32_t lutLinearIndex;
    uint32_t nSpectra;
    // element counts of the variable-length tables
    uint32_t specTotal, nHeel, heelWeightTotal, nBowtie, bowtieTotal, pad;
    double lutLinearStep; // log10 units, both precisions
    double lutLinearEnergy;
    const double* lutGrid; // log10 E of the nodes, both precisions
    const double* lutBound; // FP32: lowest energy of each segment on the FP64 path
    // FP32: the same rounded up to float; for a float E, E >= b and E < b
    // decide exactly as against b itself
    const float* lutBoundUp;
    const float* lutAnchor; // FP32: log2 E of the nodes
    float lutLinearEnergy2, lutLinearStepInv2; // FP32: linear index start and 1/step, log2 units
    const T* lutCoef; // [mat][node][type][b,a]
    // FP32: [mat][node] lower and upper bound of the total attenuation the
    // FP32 evaluation gives at any energy its segment serves (null: none)
    const float2* lutSumBound;
    const T* lutMajorant; // [node][b,a]
    const T* rita; // [mat][x,e,a,b][56]
    const T* spline; // [mat][60 + start, step, stop]
    const Shell<T>* shells; // [mat][12]
    const uint32_t* specOffset;
    const uint32_t* specThreshold; // 32-bit truncation of 2^64 mod n per spectrum
    const T* specProb;
    const uint32_t* specAlias;
    const T* specEnergy;
    const Heel<T>* heel;
    const T* heelWeight;
    const uint32_t* bowtieOffset;
    const T* bowtieAngle;
    const T* bowtieWeight;
};

template <typename T>
struct Exposure {
    T pos[3];
    T cosines[6];
    T beamDir[3];
    T collimation[4];
    T weight;
    T monoEnergy;
    int32_t spectrum, heel, bowtie, pad;
};

// Voxel world. FP64 plain path: separate material / density / measurement
// arrays as the reference. FP32 path: a record per distinct voxel content
// (x = FP32 density bits, y = material | measurement << 24) in 'palette' and
// a 1- or 2-byte palette index per voxel in 'voxels' (voxelBytes 1 or 2);
// worlds with more than 65536 distinct records keep the 8-byte record itself
// per voxel (voxelBytes 8, palette unused). Such a world may also carry a
// 1- or 2-byte code per voxel (codeBytes 1 or 2, else 0) that names the
// voxel's material and an interval holding its density ('bounds'): the
// Woodcock decision is settled from the interval at most steps, and the
// 8-byte record is read only when the draw falls inside the interval's
// acceptance band. A world with 1-byte codes also carries 'blocks': per
// block of 2^blockShift voxels per edge (8 to 64, the smallest whose table
// fits 64 KiB; clipped at the grid's upper faces) the smallest
// code | the largest code << 8 when all its voxels share one material (the
// codes of a material are ordered by density), else BLOCK_MIXED, which
// settles nothing (its smallest code MIXED_CODE is no voxel's; its largest,
// 0, names a valid material row). A 1-byte code is material * bins + bin:
// (code * codeMul) >> CODE_MUL_SHIFT is its material. The block's interval
// settles most decisions without the voxel's code; the code array is read
// only for the draws that fall inside the block's acceptance band.
struct DensityBound {
    float lo, hi; // lo <= FP32 density <= hi for every voxel with this code
    uint32_t material, pad;
};

template <typename T>
struct World {
    uint32_t nx, ny, nz, voxelBytes;
    T spacing[3];
    T extent[6];
    const uint8_t* material;
    const double* density;
    const uint8_t* measurement; // may be null
    const void* voxels;
    const uint2* palette;
    const void* codes;
    const DensityBound* bounds;
    uint32_t codeBytes;
    // 1 when the linear voxel index fits 24-bit products: nx * ny (the
    // slice stride 'nxy') and nz below 2^24 and the voxel count below 2^32
    uint32_t index24, nxy;
    // block codes (null without 1-byte codes, or when a run does not use
    // them), blocks per row and per slice
    const uint16_t* blocks;
    uint32_t bnx, bnxy, blockShift;
    uint32_t codeMul; // 1-byte codes: material of a code by multiply and shift
};

constexpr uint32_t BLOCK_SHIFT_MIN = 3, BLOCK_SHIFT_MAX = 6; // 8 to 64 voxels per block edge
constexpr uint64_t BLOCK_TABLE_BYTES = 64 * 1024; // a block-code table above this size takes the next block edge
constexpr uint32_t MIXED_CODE = 255; // the 1-byte code kept for mixed blocks
constexpr uint16_t BLOCK_MIXED = MIXED_CODE; // smallest code 255, largest 0
constexpr uint32_t CODE_MUL_SHIFT = 20;

// One voxel's accumulators in 32 aligned bytes: the three atomics of a
// scoring event land in one cache line (separate arrays cost three random
// lines per event). Normalisation and download write the per-voxel arrays
// of Result from these.
struct alignas(32) VoxelScore {
    double dose; // energy imparted, eV x weight
    double variance; // sum of squares of the same
    uint32_t events;
    uint32_t pad0;
    uint64_t pad1;
};
static_assert(sizeof(VoxelScore) == 32, "one voxel score is 32 bytes");

struct Tally {
    VoxelScore* voxels;
};

// tally plus this thread's event counter
struct Scorer {
    Tally tally;
    uint32_t events;
};

struct Counters {
    unsigned long long steps;
    unsigned long long events;
};

// Hard bounds that make every loop finish (the reference loops until
// acceptance): a rejection loop that has not accepted after MAX_TRIALS
// trials keeps its last trial, a photon still alive after MAX_STEPS Woodcock
// steps is dropped without scoring. Neither bound is reached by any
// physical case (acceptance rates are >= 1e-2, histories take < 1e4 steps);
// the oracle applies the same bounds.
constexpr uint32_t MAX_TRIALS = 1000;
constexpr uint32_t MAX_STEPS = 1u << 20;

// ------------------------------------------------------------------ helpers

template <typename T>
__device__ inline T dmin(T a, T b) { return b < a ? b : a; } // std::min
template <typename T>
__device__ inline T dmax(T a, T b) { return a < b ? b : a; } // std::max

// first index i with a[i] > v, n if none (std::upper_bound)
template <typename T>
__device__ inline uint32_t upperBound(const T* a, uint32_t n, T v)
{
    uint32_t lo = 0, count = n;
    while (count > 0) {
        const uint32_t step = count / 2;
        const uint32_t mid = lo + step;
        if (!(v < a[mid])) {
            lo = mid + 1;
            count -= step + 1;
        } else
            count = step;
    }
    return lo;
}

template <typename T>
__device__ inline T dot3(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <typename T>
__device__ inline void rotate(T* v, const T* axis, T angle)
{
    const T s = sin(angle);
    const T c = cos(angle);
    const T m = (T(1) - c) * dot3(v, axis);
    const T r0 = c * v[0] + m * axis[0] + s * (axis[1] * v[2] - axis[2] * v[1]);
    const T r1 = c * v[1] + m * axis[1] + s * (-axis[0] * v[2] + axis[2] * v[0]);
    const T r2 = c * v[2] + m * axis[2] + s * (axis[0] * v[1] - axis[1] * v[0]);
    v[0] = r0;
    v[1] = r1;
    v[2] = r2;
}

// cosine and sine of a polar scattering angle
template <typename T>
struct Polar {
    T c, s;
};

template <typename T>
__device__ inline Polar<T> polarOfAngle(T theta) { return { cos(theta), sin(theta) }; }

// The polar angle of a scatter from its sampled cosine. FP64 goes through
// the angle as the reference does (acos, then cos and sin of it) and rounds
// like it; FP32 keeps the cosine and takes the sine as sqrt(1 - c^2) (the
// same values up to rounding: theta = acos(c) lies in [0, pi]).
__device__ inline Polar<double> polarOfCos(double c) { return polarOfAngle(acos(c)); }
__device__ inline Polar<float> polarOfCos(float c) { return { c, sqrt(fmaxf(0.0f, 1.0f - c * c)) }; }

template <typename T>
__device__ inline void peturb(T* dir, Polar<T> polar, T phi)
{
    const T ax = fabs(dir[0]), ay = fabs(dir[1]), az = fabs(dir[2]);
    const int k = ax <= ay ? (ax <= az ? 0 : 2) : (ay <= az ? 1 : 2);
    T e[3] = { T(0), T(0), T(0) };
    e[k] = T(1);
    T o[3];
    o[0] = dir[1] * e[2] - dir[2] * e[1];
    o[1] = dir[2] * e[0] - dir[0] * e[2];
    o[2] = dir[0] * e[1] - dir[1] * e[0];
    rotate(o, dir, phi);
    const T st = polar.s;
    const T ct = polar.c;
    dir[0] = dir[0] * ct + o[0] * st;
    dir[1] = dir[1] * ct + o[1] * st;
    dir[2] = dir[2] * ct + o[2] * st;
}

template <typename T>
__device__ inline void peturb(T* dir, T theta, T phi)
{
    peturb(dir, polarOfAngle(theta), phi);
}

// ------------------------------------------------------ attenuation tables

// Segment search in log10 E on the FP64 grid for both precisions
// (attenuationinterpolator.hpp:207-229, linear index with its lag above the
// last edge, binary search below): the FP32 path then uses the same segment
// as the FP64 path, also for energies that sit on a grid node, where float
// rounding of log E would otherwise pick the neighbouring segment.
template <typename T>
__device__ inline uint32_t lutInterval(const Tables<T>& t, double le)
{
    if (le > t.lutLinearEnergy) {
        const double f = (le - t.lutLinearEnergy) / t.lutLinearStep;
        if (!(f < double(t.lutNodes))) // beyond the table (or NaN): last interval
            return t.lutNodes - 1;
        const uint32_t k = static_cast<uint32_t>(f) + t.lutLinearIndex;
        return k < t.lutNodes - 1 ? k : t.lutNodes - 1;
    }
    const uint32_t up = upperBound(t.lutGrid, t.lutNodes, le);
    return up != t.lutNodes ? up : t.lutNodes - 1;
}

// An energy's place in the LUT: segment k and
//   FP64: x = log10 E, mu = 10^(b + a x) exactly as the reference;
//   FP32: x = log2 E - X_k, the offset from the segment's anchor node X_k
//         (log2 units), mu = 2^(y_k + a x) with y_k = log2 mu(X_k) stored by
//         the host. This avoids the cancellation of b + a log E in float for
//         steep segments.
template <typename T>
struct LutPoint {
    uint32_t k;
    T x;
};

__device__ inline LutPoint<double> lutPoint(const Tables<double>& t, double E)
{
    const double le = log10(E);
    return { lutInterval(t, le), le };
}
// FP32: a first segment from log2f, then corrected against the segment
// bounds (lutBound[k] = smallest FP64 energy the FP64 path places in segment
// k or above, found on the device with the FP64 code when the tables are
// uploaded): energies on or next to a node get the FP64 path's segment
// without an FP64 logarithm.
__device__ inline LutPoint<float> lutPoint(const Tables<float>& t, float E)
{
    const float le2 = __builtin_amdgcn_logf(E); // log2 (v_log_f32); E >= 1 keV
    // first guess in float (the linear index, or a search of the anchors)
    uint32_t k;
    if (le2 > t.lutLinearEnergy2) {
        // f > 0; integer clamps (no float copy of lutNodes held in a
        // register): from f >= lutNodes on, the last segment
        const float f = (le2 - t.lutLinearEnergy2) * t.lutLinearStepInv2;
        k = min(min(static_cast<uint32_t>(f), t.lutNodes) + t.lutLinearIndex, t.lutNodes - 1);
    } else {
        k = upperBound(t.lutAnchor, t.lutNodes, le2);
        k = k != t.lutNodes ? k : t.lutNodes - 1;
    }
    while (k + 1 < t.lutNodes && E >= t.lutBoundUp[k + 1])
        ++k;
    while (k > 0 && E < t.lutBoundUp[k])
        --k;
    return { k, le2 - t.lutAnchor[k] };
}
__device__ inline double lutValue(const double* c, const LutPoint<double>& p) { return pow(10.0, c[0] + c[1] * p.x); }
// v_exp_f32 directly: the arguments (log2 of attenuations) are far from the
// denormal range the library exp2f guards against
__device__ inline float lutValue(const float* c, const LutPoint<float>& p) { return __builtin_amdgcn_exp2f(c[0] + c[1] * p.x); }

// photo, Compton, Rayleigh mass attenuation [cm2/g] of a material at a LUT point
template <typename T>
__device__ inline void attenuationAt(const Tables<T>& t, uint32_t mat, const LutPoint<T>& p, T att[3])
{
    // row below 256 x 2^20: 24-bit product, 32-bit offset
    const T* c = t.lutCoef + (__umul24(mat, t.lutNodes) + p.k) * 6u;
    att[0] = lutValue(c, p);
    att[1] = lutValue(c + 2, p);
    att[2] = lutValue(c + 4, p);
}

template <typename T>
__device__ inline void attenuation(const Tables<T>& t, uint32_t mat, T E, T att[3])
{
    attenuationAt(t, mat, lutPoint(t, E), att);
}

// 1 / majorant (maxAttenuationInverse, attenuationinterpolator.hpp:231-248)
template <typename T>
__device__ inline T majorantInverseAt(const Tables<T>& t, const LutPoint<T>& p)
{
    return lutValue(t.lutMajorant + p.k * 2, p);
}

template <typename T>
__device__ inline T majorantInverse(const Tables<T>& t, T E)
{
    return majorantInverseAt(t, lutPoint(t, E));
}

// RITA sample of q^2 below qsqrMax (dxmcrandom.hpp:483-500)
template <typename T>
__device__ inline T ritaSample(const Tables<T>& t, uint32_t mat, Rng& rng, T maxValue)
{
    constexpr uint32_t N = 56;
    const T* x = t.rita + static_cast<size_t>(mat) * 4 * N;
    const T* e = x + N;
    const T* a = x + 2 * N;
    const T* b = x + 3 * N;
    const uint32_t im = upperBound(x, N, maxValue);
    const T cdfMax = im != N ? e[im] : T(1);
    T res;
    uint32_t trials = 0;
    do {
        const T r = rng.uniform<T>(cdfMax);
        uint32_t i = upperBound(e, N, r);
        i = i > 0 ? i - 1 : 0;
        i = i < N - 2 ? i : N - 2;
        const T v = r - e[i];
        const T d = e[i + 1] - e[i];
        res = x[i] + (1 + a[i] + b[i]) * d * v / (d * d + a[i] * d * v + b[i] * v * v) * (x[i + 1] - x[i]);
    } while (res > maxValue && ++trials < MAX_TRIALS);
    return res;
}

// Compton scatter factor spline S(q)/Z (interpolation.hpp:169-176)
template <typename T>
__device__ inline T scatterFactor(const Tables<T>& t, uint32_t mat, T q)
{
    const T* s = t.spline + static_cast<size_t>(mat) * 63;
    const T start = s[60], step = s[61], stop = s[62];
    const T x = q < start ? start : (stop < q ? stop : q); // std::clamp
    const uint32_t k = x > start ? static_cast<uint32_t>((x - start) / step) : 0;
    const uint32_t o = k < 15 ? k * 4 : 14 * 4;
    return s[o] + s[o + 1] * x + s[o + 2] * x * x + s[o + 3] * x * x * x;
}

// --------------------------------------------------------------- sources

template <typename T>
__device__ inline T sampleSpectrum(const Tables<T>& t, int32_t spec, Rng& rng)
{
    const uint32_t off = t.specOffset[spec];
    const uint32_t n = t.specOffset[spec + 1] - off;
    const T r = rng.uniform<T>();
    const uint32_t k = rng.index(n, t.specThreshold[spec]);
    const uint32_t bin = r < t.specProb[off + k] ? k : t.specAlias[off + k];
    return bin < n - 1 ? rng.uniform<T>(t.specEnergy[off + bin], t.specEnergy[off + bin + 1]) : t.specEnergy[off + bin];
}

template <typename T>
__device__ inline T bowtieWeight(const Tables<T>& t, int32_t id, T signedAngle)
{
    const T angle = fabs(signedAngle);
    const uint32_t off = t.bowtieOffset[id];
    const T* xa = t.bowtieAngle + off;
    const T* ya = t.bowtieWeight + off;
    const uint32_t n = t.bowtieOffset[id + 1] - off;
    uint32_t lo = 0, hi = n - 1;
    uint32_t it = (hi - lo) / 2;
    while (it != lo) {
        if (angle < xa[it]) {
            hi = it;
            it = lo;
        } else
            lo = it;
        it += (hi - lo) / 2;
    }
    if (angle < xa[lo])
        return ya[0];
    if (angle > xa[hi])
        return ya[n - 1];
    return ya[lo] + (angle - xa[lo]) * (ya[hi] - ya[lo]) / (xa[hi] - xa[lo]);
}

template <typename T>
__device__ inline T heelWeight(const Tables<T>& t, int32_t id, T angle, T energy)
{
    const Heel<T> h = t.heel[id];
    uint32_t e = energy < h.energyStart ? 0 : static_cast<uint32_t>((energy - h.energyStart + T(0.5) * h.energyStep) / h.energyStep);
    if (e >= h.energySize)
        e = h.energySize - 1;
    uint32_t a = angle < h.angleStart ? 0 : static_cast<uint32_t>((angle - h.angleStart) / h.angleStep);
    if (a >= h.angleSize)
        a = h.angleSize - 1;
    const T* w = t.heelWeight + h.weightOffset + e * h.angleSize + a;
    if (a < h.angleSize - 1) {
        const T a0 = h.angleStart + h.angleStep * a;
        const T a1 = h.angleStart + h.angleStep * (a + 1);
        return w[0] + (w[1] - w[0]) * (angle - a0) / (a1 - a0);
    }
    return w[0];
}

template <typename T>
__device__ inline void sampleParticle(const Exposure<T>& x, const Tables<T>& t, Rng& rng, Particle<T>& p)
{
    const T theta = rng.uniform<T>(x.collimation[0], x.collimation[1]);
    const T phi = rng.uniform<T>(x.collimation[2], x.collimation[3]);
    for (int i = 0; i < 3; ++i) {
        p.pos[i] = x.pos[i];
        p.dir[i] = x.beamDir[i];
    }
    p.weight = x.weight;
    rotate(p.dir, &x.cosines[3], theta);
    rotate(p.dir, &x.cosines[0], phi);
    p.energy = x.spectrum >= 0 ? sampleSpectrum(t, x.spectrum, rng) : x.monoEnergy;
    if (x.bowtie >= 0)
        p.weight *= bowtieWeight(t, x.bowtie, theta);
    if (x.heel >= 0)
        p.weight *= heelWeight(t, x.heel, phi, p.energy);
}

// ------------------------------------------------------------------ world

template <typename T>
__device__ inline bool inside(const World<T>& w, const T* pos)
{
    return (pos[0] > w.extent[0] && pos[0] < w.extent[1]) && (pos[1] > w.extent[2] && pos[1] < w.extent[3]) && (pos[2] > w.extent[4] && pos[2] < w.extent[5]);
}

template <typename T>
__device__ inline bool transportToWorld(const World<T>& w, Particle<T>& p)
{
    if (inside(w, p.pos))
        return true;
    // numeric_limits<T>::lowest() / max()
    T amin = sizeof(T) == 8 ? T(-1.7976931348623157e308) : T(-3.40282347e38f);
    T amax = sizeof(T) == 8 ? T(1.7976931348623157e308) : T(3.40282347e38f);
    for (int i = 0; i < 3; ++i) {
        if (fabs(p.dir[i]) > T(1.0e-9)) {
            const T a0 = (w.extent[i * 2] - p.pos[i]) / p.dir[i];
            const T an = (w.extent[i * 2 + 1] - p.pos[i]) / p.dir[i];
            amin = dmax(amin, dmin(a0, an));
            amax = dmin(amax, dmax(a0, an));
        }
    }
    if (amin < amax && amin > 0) {
        for (int i = 0; i < 3; ++i)
            p.pos[i] += amin * p.dir[i];
        return true;
    }
    return false;
}

template <typename T>
__device__ inline uint64_t voxelIndex(const World<T>& w, const T* pos)
{
    uint32_t ix = static_cast<uint32_t>((pos[0] - w.extent[0]) / w.spacing[0]);
    uint32_t iy = static_cast<uint32_t>((pos[1] - w.extent[2]) / w.spacing[1]);
    uint32_t iz = static_cast<uint32_t>((pos[2] - w.extent[4]) / w.spacing[2]);
    ix = ix < w.nx ? ix : w.nx - 1;
    iy = iy < w.ny ? iy : w.ny - 1;
    iz = iz < w.nz ? iz : w.nz - 1;
    return (static_cast<uint64_t>(iz) * w.ny + iy) * w.nx + ix;
}

// safeValueAdd x3 (transport.hpp:208-214): hardware FP64 / u32 atomics
__device__ inline void score(Scorer& sc, uint64_t idx, double e)
{
    VoxelScore* v = sc.tally.voxels + idx;
    unsafeAtomicAdd(&v->dose, e);
    atomicAdd(&v->events, 1u);
    unsafeAtomicAdd(&v->variance, e * e);
    ++sc.events;
}

// -------------------------------------------------------------- interactions

constexpr double ELECTRON_REST_MASS = 510.9989461;
constexpr double KEV_TO_ANGSTROM = 12.398520;

template <typename T, int L>
__device__ inline T photoAbsorption(const Tables<T>& t, Particle<T>& p, uint32_t mat, Rng& rng)
{
    const T E = p.energy;
    p.energy = 0;
    if constexpr (L < 2) {
        return E;
    } else {
        const Shell<T>* sh = t.shells + mat * 12;
        T cum[12];
        T acc = 0;
        for (int i = 0; i < 12; ++i) {
            acc = i == 0 ? (E > sh[0].binding ? sh[0].photoProb : T(0)) : acc + (E > sh[i].binding ? sh[i].photoProb : T(0));
            cum[i] = acc;
        }
        int idx = 0;
        const T sample = cum[11] * rng.uniform<T>();
        while (cum[idx] < sample && idx < 11)
            ++idx;
        if (sh[idx].binding > E || sh[idx].binding < T(1))
            return E;
        const T r1 = rng.uniform<T>();
        if (r1 <= sh[idx].yield) {
            int line = 0;
            T r3 = rng.uniform<T>() - sh[idx].lineProb[0];
            while (r3 > 0 && line < 2) {
                ++line;
                r3 -= sh[idx].lineProb[line];
            }
            p.energy = sh[idx].lineEnergy[line];
            const T theta = rng.uniform<T>(Consts<T>::PI);
            const T phi = rng.uniform<T>(Consts<T>::PI + Consts<T>::PI);
            peturb(p.dir, theta, phi);
            return E - p.energy;
        }
        return E;
    }
}

// The Compton and Rayleigh samplers as a set-up (values that depend on the
// photon only, and for the impulse approximation the shell drawn first) and
// one attempt of their rejection loops (its draws and the acceptance test).
// scatterEvent() below runs attempts until one is accepted or MAX_TRIALS are
// spent, as the reference's loops do, and ends both event types with one
// copy of the azimuth draw and the turn of the direction.
template <typename T>
struct ScatterSetup {
    T k, emin, gmaxInv; // Compton: E / mc^2, smallest kept energy fraction, 1 / largest KN weight
    T U, pb, J0; // impulse approximation: binding / mc^2 of the shell, its momentum bound, J0
    T qmax2; // Rayleigh, L >= 1: largest q^2
    bool turns; // false: an impulse-approximation shell bound more tightly than E (no scatter)
};

// shell of an impulse-approximation Compton event (one draw;
// comptonscatterIA, interactions.hpp)
template <typename T>
__device__ inline uint32_t impulseShell(const Tables<T>& t, uint32_t mat, T E, Rng& rng)
{
    const Shell<T>* sh = t.shells + mat * 12;
    T cum[12];
    T acc = 0;
    for (int i = 0; i < 12; ++i) {
        const T w = (sh[i].binding < E && sh[i].j0 > 0) ? sh[i].electrons : T(0);
        acc = i == 0 ? w : acc + w;
        cum[i] = acc;
    }
    uint32_t si = 0;
    const T sample = cum[11] * rng.uniform<T>();
    while (cum[si] < sample && si < 11)
        ++si;
    return si;
}

template <typename T, int L>
__device__ inline ScatterSetup<T> comptonSetup(const Tables<T>& t, uint32_t mat, T E, uint32_t shell)
{
    ScatterSetup<T> s {};
    s.turns = true;
    if constexpr (L == 2) {
        const Shell<T>& sh = t.shells[mat * 12 + shell];
        s.U = sh.binding / T(ELECTRON_REST_MASS);
        s.pb = sqrt(2 * s.U + s.U * s.U);
        s.J0 = sh.j0;
    }
    s.k = E / T(ELECTRON_REST_MASS);
    if constexpr (L == 2)
        s.turns = !(s.U > s.k);
    s.emin = 1 / (1 + 2 * s.k);
    s.gmaxInv = 1 / (1 / s.emin + s.emin);
    return s;
}

template <typename T, int L>
__device__ inline ScatterSetup<T> rayleighSetup(T E)
{
    ScatterSetup<T> s {};
    s.turns = true;
    if constexpr (L >= 1) {
        constexpr T kinv = T(1) / T(KEV_TO_ANGSTROM);
        const T qmax = E * kinv;
        s.qmax2 = qmax * qmax;
    }
    return s;
}

// One attempt of the Compton sampler: Klein-Nishina (L = 0), times the
// scatter function (L = 1), the impulse approximation (L = 2). Sets the
// kept energy fraction e and the cosine of the polar angle; true when
// accepted.
template <typename T, int L>
__device__ inline bool comptonAttempt(const Tables<T>& t, uint32_t mat, T E, const ScatterSetup<T>& s, Rng& rng, T& e, T& cosAngle)
{
    const T r1 = rng.uniform<T>();
    e = r1 + (1 - r1) * s.emin;
    const T tt = (1 - e) / (s.k * e);
    const T sin2 = tt * (2 - tt);
    cosAngle = 1 - tt;
    const T g = (1 / e + e - sin2) * s.gmaxInv;
    const T r2 = rng.uniform<T>();
    if constexpr (L == 0) {
        return !(r2 > g);
    } else if constexpr (L == 1) {
        constexpr T kinv = T(1) / T(KEV_TO_ANGSTROM);
        const T q = E * kinv * sqrt(T(0.5) - cosAngle * T(0.5));
        return !(r2 > g * scatterFactor(t, mat, q));
    } else {
        if (r2 > g)
            return false;
        const T k = s.k, U = s.U, pb = s.pb, J0 = s.J0;
        const T pi = (k * (k - U) * (1 - cosAngle) - U) / sqrt(2 * k * (k - U) * (1 - cosAngle) + U * U);
        const T kc = k * e;
        const T qc = sqrt(k * k + kc * kc - 2 * k * kc * cosAngle);
        const T alpha = qc * (1 + kc * (kc - k * cosAngle) / (qc * qc)) / k;
        const T bpart = 1 + 2 * J0 * fabs(pi);
        const T b = (bpart * bpart + 1) / 2;
        const T expb = exp(-b);
        T S;
        if (pi <= -pb) {
            S = (1 - alpha * pb) * expb / 2;
        } else if (pi < pb) {
            constexpr T sp2 = 1 / (Consts<T>::INV_SQRTPI / Consts<T>::SQRT2);
            const T pabs = fabs(pb);
            const T piabs = fabs(pi);
            const T part1 = alpha * sp2 / (4 * J0);
            constexpr T a1 = T(0.34802);
            constexpr T a2 = T(-0.0958798);
            constexpr T a3 = T(0.7478556);
            const T sqrte = sqrt(Consts<T>::E);
            const T tp = 1 / (1 + T(0.332673) * (1 + 2 * J0 * pabs));
            const T tpi = 1 / (1 + T(0.332673) * (1 + 2 * J0 * piabs));
            const T part2p = sqrte - expb * tp * (a1 + a2 * tp + a3 * tp * tp);
            const T part2pi = sqrte - expb * tpi * (a1 + a2 * tpi + a3 * tpi * tpi);
            if (pi <= 0)
                S = (1 - alpha * pi) * expb / 2 - part1 * (part2p - part2pi);
            else
                S = 1 - (1 - alpha * pi) * expb / 2 - part1 * (part2p - part2pi);
        } else {
            S = 1 - (1 - alpha * pb) * expb / 2;
        }
        const T r3 = rng.uniform<T>();
        if (r3 > S)
            return false;
        T Fmax;
        if (pi <= -pb)
            Fmax = 1 - alpha * pb;
        else if (pi >= pb)
            Fmax = 1 + alpha * pb;
        else
            Fmax = 1 + alpha * pi;
        const T r4 = rng.uniform<T>();
        const T rbar2 = 2 * r4 * expb;
        T pz;
        if (rbar2 < 1) {
            const T part = sqrt(1 - 2 * log(rbar2));
            pz = (1 - part) / (2 * J0) / T(ELECTRON_REST_MASS);
        } else {
            const T part = sqrt(1 - 2 * log(2 - rbar2));
            pz = (part - 1) / (2 * J0) / T(ELECTRON_REST_MASS);
        }
        T Fpz;
        if (pz <= -pb)
            Fpz = 1 - alpha * pb;
        else if (pz >= pb)
            Fpz = 1 + alpha * pb;
        else
            Fpz = 1 + alpha * pz;
        const T r5 = rng.uniform<T>();
        if (r5 > Fpz / Fmax)
            return false;
        const T part = sqrt(1 - 2 * e * cosAngle + e * e * (1 - pz * pz * sin2));
        const T kbar = kc / (1 - pz * pz * e * e) * (1 - pz * pz * e * cosAngle + pz * part);
        e = kbar / k;
        return true;
    }
}

// One attempt of the Rayleigh sampler: L = 0 the polar angle itself from
// (2 - sin^2) sin, L >= 1 the cosine from the form factor's RITA table and
// the Thomson factor. 'angle' is the angle (L = 0) or its cosine.
template <typename T, int L>
__device__ inline bool rayleighAttempt(const Tables<T>& t, uint32_t mat, T E, const ScatterSetup<T>& s, Rng& rng, T& angle)
{
    if constexpr (L == 0) {
        constexpr T extreme = (4 * Consts<T>::SQRT2) / (3 * Consts<T>::SQRT3);
        const T r1 = rng.uniform<T>(T(0), extreme);
        angle = rng.uniform<T>(T(0), Consts<T>::PI);
        const T sn = sin(angle);
        return !(r1 > ((2 - sn * sn) * sn));
    } else {
        const T q2 = ritaSample(t, mat, rng, s.qmax2);
        const T invE = T(KEV_TO_ANGSTROM) / E;
        angle = 1 - 2 * q2 * invE * invE;
        return !((1 + angle * angle) * T(0.5) < rng.uniform<T>());
    }
}

// The end of a Compton or Rayleigh event from its last attempt: the
// azimuth draw and the turn of the direction (none for an impulse shell
// that does not scatter), the kept energy fraction applied. Returns the
// energy given to the medium (0 for Rayleigh).
template <typename T, int L>
__device__ inline T finishScatter(Particle<T>& p, Rng& rng, bool compton, bool turns, T e, T angle)
{
    if (turns) {
        const Polar<T> polar = (L == 0 && !compton) ? polarOfAngle(angle) : polarOfCos(angle);
        const T phi = rng.uniform<T>(Consts<T>::PI + Consts<T>::PI);
        peturb(p.dir, polar, phi);
    }
    const T E = p.energy;
    p.energy *= compton ? e : T(1);
    return E - p.energy;
}

// A Compton (compton = true) or Rayleigh event after the type draw, with
// its rejection loop run to the end (at most MAX_TRIALS attempts).
template <typename T, int L>
__device__ inline T scatterEvent(const Tables<T>& t, Particle<T>& p, uint32_t mat, Rng& rng, bool compton)
{
    T e = 1, angle = 1;
    bool turns = true;
    if (compton) {
        const uint32_t shell = L == 2 ? impulseShell(t, mat, p.energy, rng) : 0u;
        const ScatterSetup<T> s = comptonSetup<T, L>(t, mat, p.energy, shell);
        turns = s.turns;
        uint32_t trials = 0;
        if (turns)
            while (!comptonAttempt<T, L>(t, mat, p.energy, s, rng, e, angle) && ++trials < MAX_TRIALS) { }
    } else {
        const ScatterSetup<T> s = rayleighSetup<T, L>(p.energy);
        uint32_t trials = 0;
        while (!rayleighAttempt<T, L>(t, mat, p.energy, s, rng, angle) && ++trials < MAX_TRIALS) { }
    }
    return finishScatter<T, L>(p, rng, compton, turns, e, angle);
}

template <typename T, int L>
__device__ inline T comptonScatter(const Tables<T>& t, Particle<T>& p, uint32_t mat, Rng& rng)
{
    return scatterEvent<T, L>(t, p, mat, rng, true);
}

template <typename T, int L>
__device__ inline void rayleighScatter(const Tables<T>& t, Particle<T>& p, uint32_t mat, Rng& rng)
{
    scatterEvent<T, L>(t, p, mat, rng, false);
}

// Energy imparted by an event, with the 1 keV cutoff: below it the photon is
// absorbed and its remaining energy is added (transport.hpp:596-611).
template <typename T>
__device__ inline bool depositAfterEvent(Scorer& sc, uint64_t idx, Particle<T>& p, T e, bool& updateMajorant)
{
    if (p.energy < T(1)) {
        const T ei = (e + p.energy) * p.weight;
        score(sc, idx, static_cast<double>(ei));
        p.energy = 0;
        return false;
    }
    const T ei = e * p.weight;
    score(sc, idx, static_cast<double>(ei));
    updateMajorant = true;
    return true;
}

template <typename T, int L>
__device__ inline bool computeInteractions(const Tables<T>& t, const T att[3], Particle<T>& p, uint32_t mat, Scorer& sc, uint64_t idx,
    Rng& rng, bool& updateMajorant)
{
    const T total = (att[0] + att[1]) + att[2];
    const T r3 = rng.uniform<T>(total);
    // one scoring call site for photoabsorption and Compton (a wave whose
    // lanes hold both runs the atomics once)
    const bool photo = r3 < att[0];
    const bool compton = !photo && r3 < (att[0] + att[1]);
    const T e = photo ? photoAbsorption<T, L>(t, p, mat, rng) : scatterEvent<T, L>(t, p, mat, rng, compton);
    return (photo || compton) ? depositAfterEvent(sc, idx, p, e, updateMajorant) : true;
}

template <typename T, int L>
__device__ inline bool computeInteractionsForced(const Tables<T>& t, T eventProbability, const T att[3], Particle<T>& p, uint32_t mat,
    Scorer& sc, uint64_t idx, Rng& rng, bool& updateMajorant)
{
    const T total = (att[0] + att[1]) + att[2];
    const T photoProbability = att[0] / total;
    const T weightCorrection = eventProbability * photoProbability;
    {
        Particle<T> pf = p; // virtual photoabsorption; fluorescence is dropped unless below the cutoff
        const T ef = photoAbsorption<T, L>(t, pf, mat, rng);
        T ei;
        if (pf.energy < T(1))
            ei = (ef + pf.energy) * pf.weight * weightCorrection;
        else
            ei = ef * pf.weight * weightCorrection;
        score(sc, idx, static_cast<double>(ei));
    }
    const T r1 = rng.uniform<T>();
    if (r1 < eventProbability * (1 - photoProbability)) {
        const T r2 = rng.uniform<T>(att[1] + att[2]);
        const bool compton = r2 < att[1];
        const T e = scatterEvent<T, L>(t, p, mat, rng, compton);
        if (compton && !depositAfterEvent(sc, idx, p, e, updateMajorant))
            return false;
    }
    p.weight *= (1 - weightCorrection);
    return true;
}

template <typename T>
__device__ inline void readVoxel(const World<T>& w, uint64_t idx, uint32_t& mat, T& density, uint32_t& meas);

template <>
__device__ inline void readVoxel<double>(const World<double>& w, uint64_t idx, uint32_t& mat, double& density, uint32_t& meas)
{
    mat = w.material[idx];
    density = w.density[idx];
    meas = w.measurement ? w.measurement[idx] : 0u;
}
template <>
__device__ inline void readVoxel<float>(const World<float>& w, uint64_t idx, uint32_t& mat, float& density, uint32_t& meas)
{
    uint2 v;
    if (w.voxelBytes == 1)
        v = w.palette[static_cast<const uint8_t*>(w.voxels)[idx]];
    else if (w.voxelBytes == 2)
        v = w.palette[static_cast<const uint16_t*>(w.voxels)[idx]];
    else
        v = static_cast<const uint2*>(w.voxels)[idx];
    density = __uint_as_float(v.x);
    mat = v.y & 0xffu;
    meas = v.y >> 24;
}

// The FP32 record's material and measurement flag as they are stored,
// material | measurement << 24 (the FP32 kernel keeps them in one register:
// 24-bit products with the material, as the LUT row index, ignore the flag)
// (the record is read as one 8-byte integer: read as a uint2 whose halves
// are used apart, the compiler issues two 4-byte loads)
__device__ inline void readVoxelPacked(const World<float>& w, uint32_t voxelBytes, uint64_t idx, uint32_t& matMeas, float& density)
{
    const uint64_t* rec = reinterpret_cast<const uint64_t*>(w.palette);
    uint64_t v;
    if (voxelBytes == 1)
        v = rec[static_cast<const uint8_t*>(w.voxels)[idx]];
    else if (voxelBytes == 2)
        v = rec[static_cast<const uint16_t*>(w.voxels)[idx]];
    else
        v = static_cast<const uint64_t*>(w.voxels)[idx];
    density = __uint_as_float(static_cast<uint32_t>(v));
    matMeas = static_cast<uint32_t>(v >> 32);
}

// One Woodcock delta-tracking step (body of the loop in transport.hpp:650-699):
// refresh the majorant after an energy change, fly an exponential distance
// under the majorant, then a real or virtual interaction in the voxel
// reached, then Russian roulette. Returns false when the photon is gone.
template <typename T, int L>
__device__ inline bool woodcockStep(const World<T>& w, const Tables<T>& t, Particle<T>& p, T& majInv, bool& update, Rng& rng, Scorer& sc,
    uint32_t& steps)
{
    if (update) {
        majInv = majorantInverse(t, p.energy);
        update = false;
    }
    const T r1 = rng.uniform<T>();
    const T step = -log(r1) * majInv * T(10); // cm -> mm
    for (int i = 0; i < 3; ++i)
        p.pos[i] += p.dir[i] * step;
    if (!inside(w, p.pos))
        return false;
    ++steps;
    const uint64_t idx = voxelIndex(w, p.pos);
    uint32_t mat, meas;
    T density;
    readVoxel(w, idx, mat, density, meas);
    T att[3];
    attenuation(t, mat, p.energy, att);
    const T prob = ((att[0] + att[1]) + att[2]) * density * majInv;
    bool alive = true;
    if (meas == 0) {
        const T r2 = rng.uniform<T>();
        if (r2 < prob)
            alive = computeInteractions<T, L>(t, att, p, mat, sc, idx, rng, update);
    } else {
        alive = computeInteractionsForced<T, L>(t, prob, att, p, mat, sc, idx, rng, update);
    }
    if (alive && p.energy * p.weight < T(5)) { // Russian roulette
        const T r4 = rng.uniform<T>();
        if (r4 < T(0.8))
            alive = false;
        else
            p.weight *= T(1) / (T(1) - T(0.8));
    }
    return alive;
}

// Woodcock tracking of one photon until it leaves the world, is absorbed or
// is killed by Russian roulette (transport.hpp:640-700).
template <typename T, int L>
__device__ inline void woodcock(const World<T>& w, const Tables<T>& t, Particle<T>& p, Rng& rng, Scorer& sc, uint32_t& steps)
{
    T majInv = 0;
    bool update = true;
    for (uint32_t n = 0; n < MAX_STEPS && woodcockStep<T, L>(w, t, p, majInv, update, rng, sc, steps); ++n) { }
}
}
