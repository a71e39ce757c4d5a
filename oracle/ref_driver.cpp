// ORACLE PINNING DRIVER - test infrastructure.
//
// Compiled against the reference's own headers (include/dxmc/*.hpp of
// DXMClib, read from the reference checkout at build time, never copied) into
// oracle/_ref/libdxmc_ref.so. It exposes the parts of the reference hot path
// that build without the reference's Material implementation (which needs
// xraylib, absent here): PCG32 and its uniforms, alias tables and spectrum
// sampling, RITA construction and sampling, the cubic spline, vector
// rotations, world entry and voxel indexing, exposure sampling without heel
// or bowtie, and the L=0 (free electron) Compton / Rayleigh samplers.
// tests/test_oracle_vs_reference.py checks the oracle against these.
#include "dxmc/dxmcrandom.hpp"
#include "dxmc/exposure.hpp"
#include "dxmc/interpolation.hpp"
#include "dxmc/transport.hpp"
#include "dxmc/vectormath.hpp"
#include "dxmc/world.hpp"

#include <cstdint>
#include <vector>

using namespace dxmc;

namespace {
// per-history stream seeding of this project (the reference seeds its
// workers from std::random_device): state1 = 2h+1, state0 = splitmix(seed ^ splitmix(h))
std::uint64_t splitmix(std::uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <typename T>
struct Probe : public Transport<T> {
    T compton0(Particle<T>& p, RandomState& s) { return this->template comptonScatter<0>(p, 0, s); }
    void rayleigh0(Particle<T>& p, RandomState& s) { this->template rayleightScatter<0>(p, 0, s); }
    T photo0(Particle<T>& p, RandomState& s) { return this->template photoAbsorption<0>(p, 0, s); }
    bool enter(const World<T>& w, Particle<T>& p) { return this->transportParticleToWorld(w, p); }
    std::size_t index(const World<T>& w, const Particle<T>& p) { return this->indexFromPosition(p, w); }
};

World<double> makeWorld(const std::uint64_t dims[3], const double spacing[3], const double origin[3])
{
    World<double> w;
    w.setDimensions({ dims[0], dims[1], dims[2] });
    w.setSpacing({ spacing[0], spacing[1], spacing[2] });
    w.setOrigin({ origin[0], origin[1], origin[2] });
    return w;
}
}

extern "C" {

void ref_pcg32(std::uint64_t s0, std::uint64_t s1, std::uint64_t n, std::uint32_t* out)
{
    std::uint64_t st[2] = { s0, s1 };
    RandomState r(st);
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = r.pcg32();
}

void ref_uniform(std::uint64_t s0, std::uint64_t s1, std::uint64_t n, int f32, double* out)
{
    std::uint64_t st[2] = { s0, s1 };
    RandomState r(st);
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = f32 ? static_cast<double>(r.randomUniform<float>()) : r.randomUniform<double>();
}

void ref_random_index(std::uint64_t s0, std::uint64_t s1, std::uint64_t max, std::uint64_t n, std::uint64_t* out)
{
    std::uint64_t st[2] = { s0, s1 };
    RandomState r(st);
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = r.randomUniform<std::size_t>(static_cast<std::size_t>(max));
}

void ref_alias_table(const double* w, std::uint64_t n, double* prob, std::uint64_t* alias)
{
    RandomDistribution<double> d(std::vector<double>(w, w + n));
    for (std::uint64_t i = 0; i < n; ++i) {
        prob[i] = d.probabilityData()[i];
        alias[i] = d.aliasingData()[i];
    }
}

void ref_spectrum_sample(const double* w, const double* e, std::uint64_t n, std::uint64_t s0, std::uint64_t s1, std::uint64_t count, double* out)
{
    SpecterDistribution<double> d(std::vector<double>(w, w + n), std::vector<double>(e, e + n));
    std::uint64_t st[2] = { s0, s1 };
    RandomState r(st);
    for (std::uint64_t i = 0; i < count; ++i)
        out[i] = d.sampleValue(r);
}

// RITA<double,56> over [lo, hi] of a callback pdf, sampled below maxValue
void ref_rita_sample(double lo, double hi, double (*pdf)(double), std::uint64_t s0, std::uint64_t s1, double maxValue, std::uint64_t n, double* out)
{
    RITA<double, 56> rita(lo, hi, [=](double x) -> double { return pdf(x); });
    std::uint64_t st[2] = { s0, s1 };
    RandomState r(st);
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = rita(r, maxValue);
}

void ref_spline_eval(double start, double stop, double (*f)(double), std::uint64_t n, const double* q, double* out)
{
    CubicSplineInterpolator<double, 16> s(start, stop, [=](double x) -> double { return f(x); });
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = s(q[i]);
}

void ref_rotate(double* v, const double* axis, double angle) { vectormath::rotate(v, axis, angle); }
void ref_peturb(double* d, double theta, double phi) { vectormath::peturb(d, theta, phi); }

// particles [n][8]: pos, dir, energy, weight
void ref_enter_world(const std::uint64_t dims[3], const double spacing[3], const double origin[3], std::uint64_t n, double* particles, int* inside,
    std::uint64_t* index, double* extentSafe)
{
    auto w = makeWorld(dims, spacing, origin);
    for (int i = 0; i < 6; ++i)
        extentSafe[i] = w.matrixExtentSafe()[i];
    Probe<double> probe;
    for (std::uint64_t i = 0; i < n; ++i) {
        double* q = particles + 8 * i;
        Particle<double> p { { q[0], q[1], q[2] }, { q[3], q[4], q[5] }, q[6], q[7] };
        inside[i] = probe.enter(w, p) ? 1 : 0;
        index[i] = inside[i] ? probe.index(w, p) : 0;
        for (int k = 0; k < 3; ++k) {
            q[k] = p.pos[k];
            q[3 + k] = p.dir[k];
        }
    }
}

// kind 0 photo, 1 Compton, 2 Rayleigh with L = 0; stream (s0, s1 + 2 i) per sample
void ref_interact_l0(int kind, std::uint64_t n, std::uint64_t seed, std::uint64_t first, double* particles, double* energy)
{
    Probe<double> probe;
    for (std::uint64_t i = 0; i < n; ++i) {
        const std::uint64_t h = first + i;
        std::uint64_t st[2] = { splitmix(seed ^ splitmix(h)), 2 * h + 1 };
        RandomState r(st);
        double* q = particles + 8 * i;
        Particle<double> p { { q[0], q[1], q[2] }, { q[3], q[4], q[5] }, q[6], q[7] };
        double e = 0;
        if (kind == 0)
            e = probe.photo0(p, r);
        else if (kind == 1)
            e = probe.compton0(p, r);
        else
            probe.rayleigh0(p, r);
        energy[i] = e;
        for (int k = 0; k < 3; ++k) {
            q[k] = p.pos[k];
            q[3 + k] = p.dir[k];
        }
        q[6] = p.energy;
        q[7] = p.weight;
    }
}

// Exposure::sampleParticle after alignToDirectionCosines, mono energetic
// (spectrum_n == 0) or with a spectrum; no heel or bowtie filter
void ref_sample_exposure(const double position[3], const double cosines[6], const double collimation[4], double weight, double monoEnergy,
    const double* specW, const double* specE, std::uint64_t specN, const double worldCosines[6], std::uint64_t seed, std::uint64_t first, std::uint64_t n,
    double* out, double* exposureRecord)
{
    std::unique_ptr<SpecterDistribution<double>> spec;
    if (specN)
        spec = std::make_unique<SpecterDistribution<double>>(std::vector<double>(specW, specW + specN), std::vector<double>(specE, specE + specN));
    Exposure<double> x({ position[0], position[1], position[2] }, { cosines[0], cosines[1], cosines[2], cosines[3], cosines[4], cosines[5] },
        std::array<double, 4> { collimation[0], collimation[1], collimation[2], collimation[3] }, n, weight, spec.get());
    x.setMonoenergeticPhotonEnergy(monoEnergy);
    x.alignToDirectionCosines({ worldCosines[0], worldCosines[1], worldCosines[2], worldCosines[3], worldCosines[4], worldCosines[5] });
    for (int k = 0; k < 3; ++k) {
        exposureRecord[k] = x.position()[k];
        exposureRecord[9 + k] = x.beamDirection()[k];
    }
    for (int k = 0; k < 6; ++k)
        exposureRecord[3 + k] = x.directionCosines()[k];
    for (std::uint64_t i = 0; i < n; ++i) {
        const std::uint64_t h = first + i;
        std::uint64_t st[2] = { splitmix(seed ^ splitmix(h)), 2 * h + 1 };
        RandomState r(st);
        const auto p = x.sampleParticle(r);
        for (int k = 0; k < 3; ++k) {
            out[8 * i + k] = p.pos[k];
            out[8 * i + 3 + k] = p.dir[k];
        }
        out[8 * i + 6] = p.energy;
        out[8 * i + 7] = p.weight;
    }
}

double ref_xcare_weight(double filterAngle, double span, double ramp, double low, double angle)
{
    XCareFilter<double> f;
    f.setFilterAngle(filterAngle);
    f.setSpanAngle(span);
    f.setRampAngle(ramp);
    f.setLowWeight(low);
    return f.sampleIntensityWeight(angle);
}
}
