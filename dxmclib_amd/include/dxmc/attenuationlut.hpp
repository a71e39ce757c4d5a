// AttenuationLut: all per-material tables the transport samples
// (reference attenuationlut.hpp:42-275):
//  * the log-log attenuation LUT and Woodcock majorant (AttenuationLutInterpolator)
//  * RITA<T,56> tables of the Rayleigh form factor squared over q^2
//  * CubicSplineInterpolator<T,16> of the normalised Compton scatter factor S(q)/Z
//  * 12 grouped electron shells per material (impulse approximation, fluorescence)
// The host builds them; Transport flattens them into dxmc_tables_desc for the
// device (csrc/) and the kernels sample exactly these values.
#pragma once
#include <string>
#include <tuple>
#include <mutex>
#include <map>

#include "dxmc/attenuationinterpolator.hpp"
#include "dxmc/constants.hpp"
#include "dxmc/dxmcrandom.hpp"
#include "dxmc/floating.hpp"
#include "dxmc/interpolation.hpp"
#include "dxmc/material.hpp"
#include "dxmc/world.hpp"

#include <array>
#include <vector>

namespace dxmc {

template <Floating T = double>
class AttenuationLut {
public:
    static constexpr int RITA_NODES = 56;
    static constexpr int SPLINE_NODES = 16;
    static constexpr T MAX_PHOTON_ENERGY() { return 2 * ELECTRON_REST_MASS<T>(); }
    static constexpr T MIN_PHOTON_ENERGY() { return T { 0.5 }; }

    AttenuationLut() { }
    AttenuationLut(const World<T>& world, T maxEnergy = 150, T minEnergy = 1) { generate(world, maxEnergy, minEnergy); }

    // tables for a world: the majorant uses each material's maximum voxel density
    void generate(const World<T>& world, T maxEnergy = 150, T minEnergy = 1)
    {
        generate(world.materialMap(), maxEnergy, minEnergy, false);
        m_attenuationData = AttenuationLutInterpolator<T>(world.materialMap(), world.densityArray()->cbegin(), world.densityArray()->cend(),
            world.materialIndexArray()->cbegin(), m_maxEnergy, m_minEnergy, true);
    }

    // same, from precomputed per-material maximum densities (resident worlds)
    void generate(const std::vector<Material>& materials, const std::vector<T>& maxDensity, T maxEnergy, T minEnergy)
    {
        generateWorldTables(materials, maxEnergy, minEnergy);
        setMaxDensity(maxDensity);
    }

    // A world's tables without the majorant; setMaxDensity() adds it once
    // the per-material maximum densities are known (the device computes
    // them when it uploads the world).
    void generateWorldTables(const std::vector<Material>& materials, T maxEnergy, T minEnergy)
    {
        generate(materials, maxEnergy, minEnergy, false);
        m_attenuationData = AttenuationLutInterpolator<T>();
        m_attenuationData.generateMaterials(materials, std::max(m_maxEnergy, T { 50 }), std::max(T { 0.1 }, m_minEnergy));
    }
    void setMaxDensity(const std::vector<T>& maxDensity) { m_attenuationData.generateMajorant(maxDensity); }

    void generate(const std::vector<Material>& materials, T maxEnergy = 150, T minEnergy = 1, bool generatePhotonData = true)
    {
        m_minEnergy = std::max(MIN_PHOTON_ENERGY(), std::min(maxEnergy, minEnergy));
        m_maxEnergy = std::min(MAX_PHOTON_ENERGY(), std::max(maxEnergy, minEnergy));
        generateFormFactors(materials);
        generateScatterFactors(materials);
        m_electronShellConfiguration.clear();
        for (const auto& m : materials)
            m_electronShellConfiguration.push_back(m.getElectronConfiguration<T>());
        if (generatePhotonData)
            m_attenuationData = AttenuationLutInterpolator<T>(materials, m_maxEnergy, m_minEnergy);
    }

    T maxTotalAttenuationInverse(T energy) const { return m_attenuationData.maxAttenuationInverse(energy); }
    std::array<T, 3> photoComptRayAttenuation(std::size_t material, T energy) const { return m_attenuationData(material, energy); }
    T momentumTransferFromFormFactor(std::size_t material, const T qsqrMax, RandomState& state) const { return m_formFactor[material](state, qsqrMax); }
    T comptonScatterFactor(std::size_t material, T q) const { return m_comptonScatterFactor[material](q); }
    const std::array<ElectronShellConfiguration<T>, 12>& electronShellConfiguration(std::size_t m) const { return m_electronShellConfiguration[m]; }

    static T momentumTransfer(T energy, T angle) { return energy * std::sin(angle * T { 0.5 }) * (1 / KEV_TO_ANGSTROM<T>()); }
    static T momentumTransferFromCos(T energy, T cosAngle)
    {
        constexpr T k = 1 / KEV_TO_ANGSTROM<T>();
        return energy * k * std::sqrt(T { 0.5 } - cosAngle * T { 0.5 });
    }
    static T momentumTransferMax(T energy)
    {
        constexpr T k = 1 / KEV_TO_ANGSTROM<T>();
        return energy * k;
    }
    static T cosAngle(const T energy, const T qsqr)
    {
        const T invE = KEV_TO_ANGSTROM<T>() / energy;
        return 1 - 2 * qsqr * invE * invE;
    }

    T minEnergy() const { return m_minEnergy; }
    T maxEnergy() const { return m_maxEnergy; }
    const AttenuationLutInterpolator<T>& interpolator() const { return m_attenuationData; }
    const std::vector<RITA<T, RITA_NODES>>& formFactorTables() const { return m_formFactor; }
    const std::vector<CubicSplineInterpolator<T, SPLINE_NODES>>& scatterFactorTables() const { return m_comptonScatterFactor; }
    const std::vector<std::array<ElectronShellConfiguration<T>, 12>>& shellConfigurations() const { return m_electronShellConfiguration; }

protected:
    // RITA over q^2 in [0, q_app], where q_app is grown (as a momentum
    // transfer: +0.5 while F^2 > 0.5, else +0.1) until F^2(q_app) <= 0.001 or
    // q_app reaches qmax^2 (reference attenuationlut.hpp:212-233)
    // A built table depends only on the material's composition and its range,
    // so the tables are kept per (composition, range) for the life of the
    // process: a server that sees the same materials again skips the build.
    template <typename Table, typename Build>
    static Table cachedTable(int kind, const Material& m, T range, Build build)
    {
        static std::mutex mutex;
        static std::map<std::tuple<int, std::string, T>, Table> cache;
        const auto key = std::make_tuple(kind, m.compositionKey(), range);
        {
            std::lock_guard lock(mutex);
            auto it = cache.find(key);
            if (it != cache.end())
                return it->second;
        }
        Table t = build();
        std::lock_guard lock(mutex);
        return cache.emplace(key, std::move(t)).first->second;
    }

    void generateFormFactors(const std::vector<Material>& materials)
    {
        m_formFactor.clear();
        const T qmax = momentumTransferMax(m_maxEnergy);
        const T qmaxSqr = qmax * qmax;
        for (const auto& m : materials) {
            T qApp = 1;
            T ff = m.getRayleightFormFactorSquared(qApp);
            while (qApp < qmaxSqr && ff > T { 0.001 }) {
                qApp += ff > T { 0.5 } ? T { 0.5 } : T { 0.1 };
                ff = m.getRayleightFormFactorSquared(qApp);
            }
            m_formFactor.push_back(cachedTable<RITA<T, RITA_NODES>>(0, m, qApp, [&] {
                return RITA<T, RITA_NODES>(T { 0 }, qApp, [&](T qsqr) -> T { return m.getRayleightFormFactorSquared(std::sqrt(qsqr)); });
            }));
        }
    }
    // spline of S(q)/Z over [0, q_app]; q_app grows after each evaluation until
    // S/Z >= 0.999 or q_app reaches qmax (reference :240-265)
    void generateScatterFactors(const std::vector<Material>& materials)
    {
        m_comptonScatterFactor.clear();
        for (const auto& m : materials) {
            const T qmaxE = momentumTransferMax(m_maxEnergy);
            T qApp = 0.5;
            T sf = m.getComptonNormalizedScatterFactor(qApp);
            while (sf < T { 0.999 } && qApp < qmaxE) {
                sf = m.getComptonNormalizedScatterFactor(qApp);
                qApp += sf < T { 0.5 } ? T { 0.5 } : T { 0.1 };
            }
            m_comptonScatterFactor.push_back(cachedTable<CubicSplineInterpolator<T, SPLINE_NODES>>(1, m, qApp, [&] {
                return CubicSplineInterpolator<T, SPLINE_NODES>(T { 0 }, qApp, [&](const T q) -> T { return m.getComptonNormalizedScatterFactor(q); });
            }));
        }
    }

private:
    T m_minEnergy = 0;
    T m_maxEnergy = 150;
    std::vector<CubicSplineInterpolator<T, SPLINE_NODES>> m_comptonScatterFactor;
    std::vector<RITA<T, RITA_NODES>> m_formFactor;
    AttenuationLutInterpolator<T> m_attenuationData;
    std::vector<std::array<ElectronShellConfiguration<T>, 12>> m_electronShellConfiguration;
};
}
