// X-ray tube: energy grid, spectrum, filtration and half value layer
// (reference tube.hpp:35-299). The spectrum is the reference's model:
// thick-target Bethe-Heitler bremsstrahlung of a tungsten anode over
// tabulated electron depth distributions, plus the tungsten K lines, each
// attenuated in the anode along the take-off direction (anodespectrum.hpp),
// then attenuated by the added filtration and normalised to unit sum.
#pragma once

#include "dxmc/anodespectrum.hpp"
#include "dxmc/constants.hpp"
#include "dxmc/floating.hpp"
#include "dxmc/material.hpp"

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

namespace dxmc {

template <Floating T = double>
class Tube {
public:
    Tube(T tubeVoltage = 120.0, T anodeAngleDeg = 12.0, T energyResolution = 1.0)
        : m_voltage(tubeVoltage)
        , m_energyResolution(energyResolution)
    {
        setAnodeAngleDeg(anodeAngleDeg);
    }
    Tube(const Tube<T>& o)
        : m_voltage(o.voltage())
        , m_energyResolution(o.energyResolution())
        , m_anodeAngle(o.anodeAngle())
        , m_filtrationMaterials(o.filtrationMaterials())
    {
    }
    Tube& operator=(const Tube<T>& o)
    {
        m_voltage = o.voltage();
        m_energyResolution = o.energyResolution();
        m_anodeAngle = o.anodeAngle();
        m_filtrationMaterials = o.filtrationMaterials();
        m_hasCachedHVL = false;
        return *this;
    }
    static constexpr T maxVoltage() { return T { 150 }; }
    static constexpr T minVoltage() { return T { 50 }; }

    T voltage() const { return m_voltage; }
    void setVoltage(T v)
    {
        m_voltage = std::min(std::max(v, minVoltage()), maxVoltage());
        m_hasCachedHVL = false;
    }
    T anodeAngle() const { return m_anodeAngle; }
    T anodeAngleDeg() const { return m_anodeAngle * RAD_TO_DEG<T>(); }
    void setAnodeAngle(T a)
    {
        m_anodeAngle = std::min(std::abs(a), PI_VAL<T>() * T { 0.5 });
        m_hasCachedHVL = false;
    }
    void setAnodeAngleDeg(T a) { setAnodeAngle(a * DEG_TO_RAD<T>()); }

    void addFiltrationMaterial(const Material& m, T mm)
    {
        m_filtrationMaterials.emplace_back(m, std::abs(mm));
        m_hasCachedHVL = false;
    }
    std::vector<std::pair<Material, T>>& filtrationMaterials() { return m_filtrationMaterials; }
    const std::vector<std::pair<Material, T>>& filtrationMaterials() const { return m_filtrationMaterials; }
    void setAlFiltration(T mm) { setElementFiltration(13, mm); }
    void setCuFiltration(T mm) { setElementFiltration(29, mm); }
    void setSnFiltration(T mm) { setElementFiltration(50, mm); }
    T AlFiltration() const { return elementFiltration("Al"); }
    T CuFiltration() const { return elementFiltration("Cu"); }
    T SnFiltration() const { return elementFiltration("Sn"); }
    void clearFiltrationMaterials() { m_filtrationMaterials.clear(); }
    void setEnergyResolution(T r) { m_energyResolution = r; }
    T energyResolution() const { return m_energyResolution; }

    // res, 2 res, ... <= voltage (accumulated, as the reference)
    std::vector<T> getEnergy() const
    {
        std::vector<T> e;
        T hv = m_energyResolution;
        while (hv <= m_voltage) {
            e.push_back(hv);
            hv = hv + m_energyResolution;
        }
        return e;
    }
    std::vector<std::pair<T, T>> getSpecter(bool normalize = true) const
    {
        const auto e = getEnergy();
        const auto s = getSpecter(e, normalize);
        std::vector<std::pair<T, T>> r(e.size());
        for (std::size_t i = 0; i < e.size(); ++i)
            r[i] = { e[i], s[i] };
        return r;
    }
    std::vector<T> getSpecter(const std::vector<T>& energies, bool normalize = true) const { return getSpecter(energies, m_anodeAngle, normalize); }
    std::vector<T> getSpecter(const std::vector<T>& energies, const T takeoffAngle, bool normalize = true) const
    {
        auto& anodeModel = anode::SpectrumCache::instance();
        const double T0 = static_cast<double>(m_voltage);
        std::vector<T> s(energies.size());
        for (std::size_t i = 0; i < energies.size(); ++i)
            s[i] = static_cast<T>(anodeModel.bremsstrahlung(T0, static_cast<double>(energies[i]), static_cast<double>(takeoffAngle)));
        addCharacteristic(energies, s, takeoffAngle);
        for (const auto& [mat, mm] : m_filtrationMaterials) {
            const T cm = mm * T { 0.1 };
            for (std::size_t i = 0; i < energies.size(); ++i)
                s[i] = s[i] * static_cast<T>(std::exp(-mat.getTotalAttenuation(energies[i]) * mat.standardDensity() * cm));
        }
        if (normalize) {
            T sum = 0;
            for (const auto v : s)
                sum = sum + v;
            for (auto& v : s)
                v = v / sum;
        }
        return s;
    }

    T mmAlHalfValueLayer()
    {
        if (!m_hasCachedHVL) {
            m_cachedHVL = computeHVL();
            m_hasCachedHVL = true;
        }
        return m_cachedHVL;
    }
    T mmAlHalfValueLayer() const { return m_hasCachedHVL ? m_cachedHVL : computeHVL(); }

protected:
    // each K line goes to the first energy bin at or above it, if that bin
    // is within 2 keV (reference tube.hpp:216-236)
    void addCharacteristic(const std::vector<T>& energies, std::vector<T>& s, T takeoff) const
    {
        for (const auto& [lineEnergy, fraction] : anode::K_LINES) {
            const T e = static_cast<T>(lineEnergy);
            const auto it = std::lower_bound(energies.cbegin(), energies.cend(), e);
            if (it == energies.cend() || std::abs(e - *it) > T { 2 })
                continue;
            const auto i = static_cast<std::size_t>(it - energies.cbegin());
            s[i] = s[i] + static_cast<T>(anode::kLineIntensity(static_cast<double>(m_voltage), lineEnergy, fraction, static_cast<double>(takeoff)));
        }
    }

    // half value layer by the fixed point iteration x <- x + (g(x) - 1/2)
    // on the normalised transmitted fraction g (reference tube.hpp:264-292)
    T computeHVL() const
    {
        const auto e = getEnergy();
        const auto s = getSpecter(e);
        Material al(13);
        std::vector<T> att(e.size());
        for (std::size_t i = 0; i < e.size(); ++i)
            att[i] = static_cast<T>(al.getTotalAttenuation(e[i]) * al.standardDensity());
        T x = 0.5, step = 1;
        int guard = 0;
        do {
            T g = 0;
            for (std::size_t i = 0; i < e.size(); ++i)
                g += s[i] * std::exp(-att[i] * x);
            step = g - T { 0.5 };
            x = x + step;
        } while (std::abs(step) > T { 0.01 } && ++guard < 10000);
        return x * T { 10 };
    }

    void setElementFiltration(int Z, T mm)
    {
        const auto sym = Material::getSymbolFromAtomicNumber(Z);
        for (auto& [m, t] : m_filtrationMaterials)
            if (m.name() == sym) {
                t = std::abs(mm);
                m_hasCachedHVL = false;
                return;
            }
        addFiltrationMaterial(Material(Z), mm);
    }
    T elementFiltration(const char* sym) const
    {
        for (const auto& [m, t] : m_filtrationMaterials)
            if (m.name() == sym)
                return t;
        return T { 0 };
    }

private:
    T m_voltage = 120;
    T m_energyResolution = 1;
    T m_anodeAngle = 0;
    std::vector<std::pair<Material, T>> m_filtrationMaterials;
    T m_cachedHVL = 0;
    bool m_hasCachedHVL = false;
};
}
