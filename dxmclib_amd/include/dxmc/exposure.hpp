// Exposure: one emitter position/orientation with its collimation, weight,
// spectrum and filters (reference exposure.hpp:36-305). The device samples
// photons from the flattened record (dxmc_exposure); sampleParticle() here is
// the host restatement used by host-side tools.
#pragma once

#include "dxmc/beamfilters.hpp"
#include "dxmc/dxmcrandom.hpp"
#include "dxmc/floating.hpp"
#include "dxmc/particle.hpp"
#include "dxmc/vectormath.hpp"

#include <array>
#include <cstdint>

namespace dxmc {

template <Floating T = double>
class Exposure {
public:
    Exposure(const std::array<T, 3>& position, const std::array<T, 6>& cosines, const std::array<T, 4>& collimation, std::uint64_t nHistories = 1000,
        T weight = 1, const SpecterDistribution<T>* spectrum = nullptr, const HeelFilter<T>* heel = nullptr, const BeamFilter<T>* filter = nullptr)
        : m_position(position)
        , m_directionCosines(cosines)
        , m_collimationAngles(collimation)
        , m_beamIntensityWeight(weight)
        , m_beamFilter(filter)
        , m_specterDistribution(spectrum)
        , m_heelFilter(heel)
        , m_nHistories(nHistories)
    {
        normalizeDirectionCosines();
    }
    Exposure(const std::array<T, 3>& position, const std::array<T, 6>& cosines, const std::array<T, 2>& collimation, std::uint64_t nHistories = 1000,
        T weight = 1, const SpecterDistribution<T>* spectrum = nullptr, const HeelFilter<T>* heel = nullptr, const BeamFilter<T>* filter = nullptr)
        : m_position(position)
        , m_directionCosines(cosines)
        , m_beamIntensityWeight(weight)
        , m_beamFilter(filter)
        , m_specterDistribution(spectrum)
        , m_heelFilter(heel)
        , m_nHistories(nHistories)
    {
        setCollimationAngles(collimation);
        normalizeDirectionCosines();
    }

    void setPosition(T x, T y, T z) { m_position = { x, y, z }; }
    void setPosition(const std::array<T, 3>& p) { m_position = p; }
    void setPositionZ(const T z) { m_position[2] = z; }
    const std::array<T, 3>& position(void) const { return m_position; }
    void addPosition(const std::array<T, 3>& p)
    {
        for (int i = 0; i < 3; ++i)
            m_position[i] += p[i];
    }
    void subtractPosition(const std::array<T, 3>& p)
    {
        for (int i = 0; i < 3; ++i)
            m_position[i] -= p[i];
    }
    void setDirectionCosines(const std::array<T, 6>& c)
    {
        m_directionCosines = c;
        normalizeDirectionCosines();
    }
    const std::array<T, 6>& directionCosines(void) const { return m_directionCosines; }
    const std::array<T, 3>& beamDirection(void) const { return m_beamDirection; }

    void setCollimationAngles(const std::array<T, 4>& a) { m_collimationAngles = a; }
    void setCollimationAngles(const std::array<T, 2>& a) { setCollimationAngles(a[0], a[1]); }
    void setCollimationAngles(const T ax, const T ay) { m_collimationAngles = { -ax / 2, ax / 2, -ay / 2, ay / 2 }; }
    const std::array<T, 4>& collimationAngles(void) const { return m_collimationAngles; }
    T collimationAngleX() const { return m_collimationAngles[1] - m_collimationAngles[0]; }
    T collimationAngleY() const { return m_collimationAngles[3] - m_collimationAngles[2]; }

    void setBeamIntensityWeight(T w) { m_beamIntensityWeight = w; }
    T beamIntensityWeight(void) const { return m_beamIntensityWeight; }
    void setBeamFilter(const BeamFilter<T>* f) { m_beamFilter = f; }
    void setSpecterDistribution(const SpecterDistribution<T>* s) { m_specterDistribution = s; }
    void setHeelFilter(const HeelFilter<T>* h) { m_heelFilter = h; }
    const BeamFilter<T>* beamFilter() const { return m_beamFilter; }
    const SpecterDistribution<T>* specterDistribution() const { return m_specterDistribution; }
    const HeelFilter<T>* heelFilter() const { return m_heelFilter; }
    void setMonoenergeticPhotonEnergy(T e) { m_monoenergeticPhotonEnergy = std::min(std::max(e, T { 0 }), T { 500 }); }
    T monoenergeticPhotonEnergy() const { return m_monoenergeticPhotonEnergy; }
    void setNumberOfHistories(std::size_t n) { m_nHistories = n; }
    std::size_t numberOfHistories(void) const { return m_nHistories; }

    // express position and cosines in the world's basis (b1, b2, b1 x b2)
    void alignToDirectionCosines(const std::array<T, 6>& basis) noexcept
    {
        const T* b1 = basis.data();
        const T* b2 = b1 + 3;
        T b3[3];
        vectormath::cross(b1, b2, b3);
        vectormath::changeBasisInverse(b1, b2, b3, m_position.data());
        vectormath::changeBasisInverse(b1, b2, b3, m_directionCosines.data());
        vectormath::changeBasisInverse(b1, b2, b3, &m_directionCosines[3]);
        vectormath::changeBasisInverse(b1, b2, b3, m_beamDirection.data());
    }

    // theta across the first cosine, phi across the second; the beam direction
    // is rotated about cosine Y by theta, then about cosine X by phi
    Particle<T> sampleParticle(RandomState& state) const noexcept
    {
        const T theta = state.randomUniform(m_collimationAngles[0], m_collimationAngles[1]);
        const T phi = state.randomUniform(m_collimationAngles[2], m_collimationAngles[3]);
        Particle<T> p { m_position, m_beamDirection, 0, m_beamIntensityWeight };
        vectormath::rotate(p.dir.data(), &m_directionCosines[3], theta);
        vectormath::rotate(p.dir.data(), &m_directionCosines[0], phi);
        p.energy = m_specterDistribution ? m_specterDistribution->sampleValue(state) : m_monoenergeticPhotonEnergy;
        if (m_beamFilter)
            p.weight *= m_beamFilter->sampleIntensityWeight(theta);
        if (m_heelFilter)
            p.weight *= m_heelFilter->sampleIntensityWeight(phi, p.energy);
        return p;
    }

protected:
    void normalizeDirectionCosines(void)
    {
        vectormath::normalize(&m_directionCosines[0]);
        vectormath::normalize(&m_directionCosines[3]);
        vectormath::cross(m_directionCosines.data(), m_beamDirection.data());
    }

private:
    std::array<T, 3> m_position { 0, 0, 0 };
    std::array<T, 6> m_directionCosines { 1, 0, 0, 0, 1, 0 };
    std::array<T, 3> m_beamDirection { 0, 0, 1 };
    std::array<T, 4> m_collimationAngles { 0, 0, 0, 0 };
    T m_beamIntensityWeight = 1;
    const BeamFilter<T>* m_beamFilter = nullptr;
    const SpecterDistribution<T>* m_specterDistribution = nullptr;
    const HeelFilter<T>* m_heelFilter = nullptr;
    T m_monoenergeticPhotonEnergy { 0 };
    std::uint64_t m_nHistories = 1000;
};
}
