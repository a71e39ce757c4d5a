// X-ray tube: energy grid, spectrum, filtration and half value layer.
// Interface of the reference (tube.hpp:35-299). The spectrum itself is this
// project's own thick-target model (the reference's semi-relativistic
// Bethe-Heitler integral over tabulated electron depth distributions is not
// reproduced):
//   N(hv) = 1/hv * integral_{hv}^{T0} exp(-mu_W(hv) x(T) / sin(takeoff)) dT
// i.e. Kramers' constant energy-loss emission with the electron energy
// falling with depth by the Thomson-Whiddington law
//   x(T) = R(T0) (1 - T^2/T0^2),  R(T0) = 0.0119 T0^1.513 mg/cm2 (tungsten)
// plus the four tungsten K lines (59.3, 58.0, 67.2, 69.1 keV; relative
// intensities 0.505, 0.291, 0.162, 0.042) with a (T0/E_K - 1)^1.63 yield,
// attenuated by the anode at the mean production depth, then filtered by
// the added filtration. Spectrum and heel tables are inputs to the
// transport; GPU and oracle sample identical tables.
#pragma once

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
        std::vector<T> s(energies.size());
        for (std::size_t i = 0; i < energies.size(); ++i)
            s[i] = static_cast<T>(bremsstrahlung(m_voltage, energies[i], takeoffAngle));
        addCharacteristic(energies, s, takeoffAngle);
        for (const auto& [mat, mm] : m_filtrationMaterials) {
            const double cm = mm * 0.1;
            for (std::size_t i = 0; i < energies.size(); ++i)
                s[i] *= static_cast<T>(std::exp(-mat.getTotalAttenuation(energies[i]) * mat.standardDensity() * cm));
        }
        if (normalize) {
            T sum = 0;
            for (const auto v : s)
                sum += v;
            for (auto& v : s)
                v /= sum;
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
    static double anodeRange(double T0) { return 0.0119 * std::pow(T0, 1.513); } // mg/cm2

    static double bremsstrahlung(double T0, double hv, double takeoff)
    {
        if (hv <= 0 || hv >= T0)
            return 0;
        const double mu = Material::getTotalAttenuation(74, hv) * 1e-3 / std::sin(takeoff); // per mg/cm2
        const double R = anodeRange(T0);
        // integral over T in [hv, T0] by 64-point midpoint rule
        constexpr int n = 64;
        const double dT = (T0 - hv) / n;
        double s = 0;
        for (int i = 0; i < n; ++i) {
            const double t = hv + (i + 0.5) * dT;
            const double x = R * (1 - (t * t) / (T0 * T0));
            s += std::exp(-mu * x);
        }
        return s * dT / hv;
    }

    void addCharacteristic(const std::vector<T>& energies, std::vector<T>& s, T takeoff) const
    {
        constexpr double kEdge = 69.525;
        if (m_voltage <= kEdge || energies.empty())
            return;
        constexpr double lineE[4] = { 59.3, 58.0, 67.2, 69.1 };
        constexpr double lineF[4] = { 0.505, 0.291, 0.162, 0.042 };
        double brems = 0;
        for (const auto e : energies)
            brems += bremsstrahlung(m_voltage, e, takeoff);
        const double yield = 0.1 * std::pow(m_voltage / kEdge - 1, 1.63) * brems;
        const double depth = anodeRange(m_voltage) / 3;
        for (int i = 0; i < 4; ++i) {
            const auto it = std::lower_bound(energies.cbegin(), energies.cend(), static_cast<T>(lineE[i]));
            if (it == energies.cend() || std::abs(lineE[i] - *it) > 2.0)
                continue;
            const double mu = Material::getTotalAttenuation(74, lineE[i]) * 1e-3 / std::sin(takeoff);
            s[it - energies.cbegin()] += static_cast<T>(lineF[i] * yield * std::exp(-mu * depth));
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
