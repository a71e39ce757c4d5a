// Attenuation look-up table and Woodcock majorant (reference
// attenuationinterpolator.hpp:30-250), built on the host and sampled by the
// kernels.
//
// Grid: log10 energies, uniform between minEnergy and maxEnergy plus a pair
// of nodes (E_b - 1 eV, E_b) at every binding energy above minEnergy.
// For every material and interaction type (photo, Compton, Rayleigh) each
// interval i stores log-log line coefficients (b, a): mu(E) = 10^(b + a log10 E).
// After fitting, the first node is dropped, so interval i ends at node i of
// the stored grid. Lookups above the last edge compute the interval index
// directly from the uniform step; that index is the one *below* the bracketing
// interval (the previous segment is extrapolated) - a reference property that
// is reproduced, as is the unclamped index of the majorant lookup (here
// clamped only where the reference would read past the table).
#pragma once

#include "dxmc/floating.hpp"
#include "dxmc/material.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace dxmc {

template <Floating T>
class AttenuationLutInterpolator {
public:
    AttenuationLutInterpolator() { }

    // world tables: maxEnergy raised to at least 50 keV
    template <typename DensIt, typename MatIt>
    AttenuationLutInterpolator(const std::vector<Material>& materials, DensIt densBegin, DensIt densEnd, MatIt matBegin, T maxEnergy, T minEnergy, bool worldTables)
    {
        minEnergy = std::max(T { 0.1 }, minEnergy);
        maxEnergy = worldTables ? std::max(maxEnergy, T { 50 }) : std::min(maxEnergy, T { 50 });
        generate(materials, densBegin, densEnd, matBegin, maxEnergy, minEnergy);
    }

    // material-list tables (densities = standard densities): maxEnergy capped at 50 keV
    AttenuationLutInterpolator(const std::vector<Material>& materials, T maxEnergy, T minEnergy)
    {
        std::vector<T> dens(materials.size());
        std::vector<std::uint8_t> idx(materials.size());
        for (std::size_t i = 0; i < materials.size(); ++i) {
            dens[i] = static_cast<T>(materials[i].standardDensity());
            idx[i] = static_cast<std::uint8_t>(i);
        }
        minEnergy = std::max(T { 0.1 }, minEnergy);
        maxEnergy = std::min(maxEnergy, T { 50 });
        generate(materials, dens.cbegin(), dens.cend(), idx.cbegin(), maxEnergy, minEnergy);
    }

    // per-material maximum density over the voxels (majorant input)
    template <typename DensIt, typename MatIt>
    static std::vector<T> maxDensities(std::size_t nMaterials, DensIt densBegin, DensIt densEnd, MatIt matBegin)
    {
        std::vector<T> maxDens(nMaterials, T { 0 });
        auto m = matBegin;
        for (auto d = densBegin; d != densEnd; ++d, ++m) {
            const std::size_t k = static_cast<std::size_t>(*m);
            if (k < nMaterials)
                maxDens[k] = std::max(maxDens[k], static_cast<T>(*d));
        }
        return maxDens;
    }

    template <typename DensIt, typename MatIt>
    void generate(const std::vector<Material>& materials, DensIt densBegin, DensIt densEnd, MatIt matBegin, const T maxEnergy, const T minEnergy)
    {
        generateFromMaxDensity(materials, maxDensities(materials.size(), densBegin, densEnd, matBegin), maxEnergy, minEnergy);
    }

    void generateFromMaxDensity(const std::vector<Material>& materials, const std::vector<T>& maxDens, const T maxEnergy, const T minEnergy)
    {
        generateMaterials(materials, maxEnergy, minEnergy);
        generateMajorant(maxDens);
    }

    // The grid and the per-material coefficients: everything but the
    // majorant, which needs each material's maximum voxel density.
    void generateMaterials(const std::vector<Material>& materials, const T maxEnergy, const T minEnergy)
    {
        m_resolution = std::max(static_cast<std::size_t>(maxEnergy - minEnergy), std::size_t { 10 });
        const T logMin = std::log10(minEnergy);
        const T logMax = std::log10(maxEnergy);

        std::vector<T> edges;
        for (const auto& mat : materials) {
            for (const double e : mat.getBindingEnergies(static_cast<double>(minEnergy))) {
                const T be = static_cast<T>(e);
                edges.push_back(std::log10(be - T { 0.001 }));
                edges.push_back(std::log10(be));
            }
        }
        std::sort(edges.begin(), edges.end());
        edges.erase(std::unique(edges.begin(), edges.end()), edges.end());

        std::vector<T> x;
        x.reserve(m_resolution + 1 + edges.size());
        for (std::size_t i = 0; i <= m_resolution; ++i)
            x.push_back(logMin + (i * (logMax - logMin)) / (m_resolution));
        x.insert(x.end(), edges.cbegin(), edges.cend());
        std::sort(x.begin(), x.end());
        x.erase(std::unique(x.begin(), x.end()), x.end());

        const std::size_t nInt = x.size() - 1;
        m_coefficients.assign(materials.size() * nInt * 6, T { 0 });
        std::vector<T> y(x.size());
        for (std::size_t m = 0; m < materials.size(); ++m) {
            for (int type = 0; type < 3; ++type) {
                for (std::size_t i = 0; i < x.size(); ++i) {
                    const double E = static_cast<double>(std::pow(T { 10 }, x[i]));
                    const double mu = type == 0 ? materials[m].getPhotoelectricAttenuation(E)
                                                : (type == 1 ? materials[m].getComptonAttenuation(E) : materials[m].getRayleightAttenuation(E));
                    y[i] = static_cast<T>(std::log10(mu));
                }
                for (std::size_t i = 0; i < nInt; ++i) {
                    const T a = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
                    const T b = y[i] - x[i] * a;
                    const std::size_t o = m * nInt * 6 + i * 6 + type * 2;
                    m_coefficients[o] = b;
                    m_coefficients[o + 1] = a;
                }
            }
        }

        m_firstNode = x.front();
        m_x.assign(x.begin() + 1, x.end());
        if (!edges.empty()) {
            const auto up = std::upper_bound(m_x.cbegin(), m_x.cend(), edges.back());
            if (up == m_x.cend()) {
                m_linearIndex = m_x.size();
                m_linearStep = m_x[m_x.size() - 1] - m_x[m_x.size() - 2];
                m_linearEnergy = m_x.back();
            } else {
                m_linearIndex = static_cast<std::size_t>(up - m_x.cbegin());
                m_linearStep = m_x[m_linearIndex + 1] - m_x[m_linearIndex];
                m_linearEnergy = m_x[m_linearIndex];
            }
        } else {
            m_linearIndex = 0;
            m_linearStep = m_x[1] - m_x[0];
            m_linearEnergy = m_x.front();
        }
        m_resolution = m_x.size();
        m_maxCoefficients.clear();
    }

    // majorant: 1 / max_m(maxDens_m * sum of the three coefficients) at
    // every node (including the dropped first one), fitted log-log again
    void generateMajorant(const std::vector<T>& maxDens)
    {
        std::vector<T> xm = m_x;
        xm.insert(xm.begin(), m_firstNode);
        std::vector<T> ym(xm.size());
        for (std::size_t i = 0; i < xm.size(); ++i) {
            T worst = 0;
            for (std::size_t m = 0; m < maxDens.size(); ++m) {
                const auto mu = (*this)(m, std::pow(T { 10 }, xm[i]));
                const T total = maxDens[m] * (mu[0] + mu[1] + mu[2]);
                worst = std::max(worst, total);
            }
            ym[i] = std::log10(1 / worst);
        }
        m_maxCoefficients.assign(m_x.size() * 2, T { 0 });
        for (std::size_t i = 0; i + 1 < xm.size(); ++i) {
            const T a = (ym[i + 1] - ym[i]) / (xm[i + 1] - xm[i]);
            const T b = ym[i] - xm[i] * a;
            m_maxCoefficients[i * 2] = b;
            m_maxCoefficients[i * 2 + 1] = a;
        }
    }

    std::size_t intervalIndex(T logE) const
    {
        if (logE > m_linearEnergy) {
            return std::min(static_cast<std::size_t>((logE - m_linearEnergy) / m_linearStep) + m_linearIndex, m_resolution - 1);
        }
        const auto up = std::upper_bound(m_x.cbegin(), m_x.cend(), logE);
        return up != m_x.cend() ? static_cast<std::size_t>(up - m_x.cbegin()) : m_resolution - 1;
    }

    // photo, Compton and Rayleigh mass attenuation [cm2/g]
    std::array<T, 3> operator()(const std::size_t material, const T energy) const
    {
        const T logE = std::log10(energy);
        const std::size_t o = material * m_resolution * 6 + intervalIndex(logE) * 6;
        std::array<T, 3> r;
        for (std::size_t i = 0; i < 3; ++i)
            r[i] = std::pow(T { 10 }, m_coefficients[o + 2 * i] + m_coefficients[o + 2 * i + 1] * logE);
        return r;
    }

    // inverse majorant [cm]
    T maxAttenuationInverse(const T energy) const
    {
        const T logE = std::log10(energy);
        std::size_t k;
        if (logE > m_linearEnergy)
            k = std::min(static_cast<std::size_t>((logE - m_linearEnergy) / m_linearStep) + m_linearIndex, m_resolution - 1);
        else {
            const auto up = std::upper_bound(m_x.cbegin(), m_x.cend(), logE);
            k = up != m_x.cend() ? static_cast<std::size_t>(up - m_x.cbegin()) : m_resolution - 1;
        }
        return std::pow(T { 10 }, m_maxCoefficients[2 * k] + m_maxCoefficients[2 * k + 1] * logE);
    }

    const std::vector<T>& grid() const { return m_x; }
    const std::vector<T>& coefficients() const { return m_coefficients; }
    const std::vector<T>& maxCoefficients() const { return m_maxCoefficients; }
    std::size_t linearIndex() const { return m_linearIndex; }
    T linearStep() const { return m_linearStep; }
    T linearEnergy() const { return m_linearEnergy; }
    std::size_t resolution() const { return m_resolution; }

private:
    std::vector<T> m_x;
    std::vector<T> m_coefficients;
    std::vector<T> m_maxCoefficients;
    std::size_t m_resolution = 40;
    std::size_t m_linearIndex = 0;
    T m_linearStep = 0;
    T m_linearEnergy = 0;
    T m_firstNode = 0;
};
}
