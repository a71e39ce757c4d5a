// Beam filters (reference beamfilters.hpp:42-766).
//  BowTieFilter  per-photon weight from the fan angle (sampled on the device)
//  XCareFilter   per-exposure weight from the gantry angle (host)
//  HeelFilter    per-photon weight from the cone angle and energy (device)
//  AECFilter     per-exposure weight from the tube position along z (host)
#pragma once

#include "dxmc/constants.hpp"
#include "dxmc/floating.hpp"
#include "dxmc/interpolation.hpp"
#include "dxmc/tube.hpp"
#include "dxmc/world.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

namespace dxmc {

template <Floating T = double>
class BeamFilter {
public:
    virtual ~BeamFilter() = default;
    virtual T sampleIntensityWeight(const T angle) const = 0;
    const std::string& filterName(void) const { return m_filterName; }
    void setFilterName(const std::string& name) { m_filterName = name; }

private:
    std::string m_filterName;
};

// Fluence profile over |fan angle|; weights normalised to mean 1 over the
// supplied samples; linear interpolation between samples, constant outside.
template <Floating T = double>
class BowTieFilter : public BeamFilter<T> {
public:
    BowTieFilter(const std::vector<T>& angles, const std::vector<T>& weights)
    {
        if (angles.size() == weights.size()) {
            m_data.resize(angles.size());
            for (std::size_t i = 0; i < angles.size(); ++i)
                m_data[i] = { std::abs(angles[i]), weights[i] };
        }
        std::sort(m_data.begin(), m_data.end());
        normalize();
    }
    BowTieFilter(const std::vector<std::pair<T, T>>& pairs)
        : m_data(pairs)
    {
        for (auto& p : m_data)
            p.first = std::abs(p.first);
        std::sort(m_data.begin(), m_data.end());
        normalize();
    }
    BowTieFilter(const BowTieFilter& other)
        : BeamFilter<T>()
        , m_data(other.data())
    {
    }

    // bisection over the sorted angles, as the reference (beamfilters.hpp:150-190)
    T sampleIntensityWeight(T signedAngle) const override
    {
        const T angle = std::abs(signedAngle);
        std::size_t lo = 0, hi = m_data.size() - 1;
        std::size_t it = (hi - lo) / 2;
        while (it != lo) {
            if (angle < m_data[it].first) {
                hi = it;
                it = lo;
            } else {
                lo = it;
            }
            it += (hi - lo) / 2;
        }
        if (angle < m_data[lo].first)
            return m_data[0].second;
        if (angle > m_data[hi].first)
            return m_data.back().second;
        const T x0 = m_data[lo].first, x1 = m_data[hi].first;
        const T y0 = m_data[lo].second, y1 = m_data[hi].second;
        return y0 + (angle - x0) * (y1 - y0) / (x1 - x0);
    }
    const std::vector<std::pair<T, T>>& data() const { return m_data; }

protected:
    void normalize()
    {
        double acc = 0.0; // the reference accumulates the mean in double
        for (const auto& p : m_data)
            acc = acc + p.second;
        const auto mean = acc / m_data.size();
        for (auto& p : m_data)
            p.second = static_cast<T>(p.second / mean);
    }

private:
    std::vector<std::pair<T, T>> m_data;
};

// Organ-based tube current modulation: low weight over a span of gantry
// angles (with linear ramps), high weight elsewhere so the mean over a
// rotation is 1.
template <Floating T = double>
class XCareFilter : public BeamFilter<T> {
public:
    XCareFilter() = default;
    T filterAngle() const { return m_filterAngle; }
    T filterAngleDeg() const { return m_filterAngle * RAD_TO_DEG<T>(); }
    void setFilterAngle(T angle)
    {
        constexpr T twoPi = 2 * PI_VAL<T>();
        m_filterAngle = std::fmod(angle, twoPi);
        if (m_filterAngle < 0)
            m_filterAngle += twoPi;
    }
    void setFilterAngleDeg(T a) { setFilterAngle(a * DEG_TO_RAD<T>()); }
    T spanAngle() const { return m_spanAngle; }
    T spanAngleDeg() const { return m_spanAngle * RAD_TO_DEG<T>(); }
    void setSpanAngle(T a)
    {
        if (a > T { 5 } * DEG_TO_RAD<T>() && a < PI_VAL<T>())
            m_spanAngle = a;
    }
    void setSpanAngleDeg(T a) { setSpanAngle(a * DEG_TO_RAD<T>()); }
    T rampAngle() const { return m_rampAngle; }
    T rampAngleDeg() const { return m_rampAngle * RAD_TO_DEG<T>(); }
    void setRampAngle(T a)
    {
        if (a >= 0 && a <= T { 0.5 } * m_spanAngle)
            m_rampAngle = a;
    }
    void setRampAngleDeg(T a) { setRampAngle(a * DEG_TO_RAD<T>()); }
    T lowWeight() const { return m_lowWeight; }
    void setLowWeight(T w)
    {
        if (w > 0 && w <= 1)
            m_lowWeight = w;
    }
    T highWeight() const
    {
        constexpr T twoPi = 2 * PI_VAL<T>();
        return (twoPi - m_spanAngle * m_lowWeight + m_lowWeight * m_rampAngle) / (twoPi - m_spanAngle + m_rampAngle);
    }
    T sampleIntensityWeight(const T angle) const override
    {
        constexpr T twoPi = 2 * PI_VAL<T>();
        T a = std::fmod(angle - m_filterAngle + PI_VAL<T>(), twoPi);
        if (a < 0)
            a += twoPi;
        const T high = highWeight();
        const T start = PI_VAL<T>() - m_spanAngle * T { 0.5 };
        if (a < start)
            return high;
        const T ramp1 = start + m_rampAngle;
        if (a < ramp1)
            return interp<T>(start, ramp1, high, m_lowWeight, a);
        const T ramp2 = ramp1 + m_spanAngle - m_rampAngle;
        if (a < ramp2)
            return m_lowWeight;
        const T end = start + m_spanAngle;
        if (a < end)
            return interp<T>(ramp2, end, m_lowWeight, high, a);
        return high;
    }

private:
    T m_filterAngle = 0;
    T m_spanAngle = 120 * DEG_TO_RAD<T>();
    T m_rampAngle = 20 * DEG_TO_RAD<T>();
    T m_lowWeight = T { 0.6 };
};

// Heel effect: relative fluence on a 5-angle x energy grid, each energy row
// normalised to mean 1; linear in angle between grid points.
template <Floating T = double>
class HeelFilter {
public:
    HeelFilter(const Tube<T>& tube, const T spanAngle = 0) { update(tube, spanAngle); }

    void update(const Tube<T>& tube, const T spanAngle = 0)
    {
        m_energySize = static_cast<std::size_t>((tube.voltage() - m_energyStart) / m_energyStep);
        if (m_energySize < 2)
            m_energySize = 2;
        m_energies.clear();
        for (std::size_t i = 0; i < m_energySize; ++i)
            m_energies.push_back(m_energyStart + i * m_energyStep);
        const T span = std::abs(spanAngle);
        m_angleStart = tube.anodeAngle() < span / 2 ? -tube.anodeAngle() : -span / 2;
        m_angleStep = (span / 2 - m_angleStart) / m_angleSize;
        m_weights.assign(m_energySize * m_angleSize, T { 0 });
        for (std::size_t i = 0; i < m_angleSize; ++i) {
            const T angle = m_angleStart + i * m_angleStep + tube.anodeAngle();
            const auto s = tube.getSpecter(m_energies, angle, false);
            for (std::size_t j = 0; j < m_energySize; ++j)
                m_weights[j * m_angleSize + i] = s[j];
        }
        for (std::size_t j = 0; j < m_energySize; ++j) {
            T* row = &m_weights[j * m_angleSize];
            // summation order of libstdc++'s std::reduce (pairs of pairs, then the tail)
            const T mean = static_cast<T>(stlReduce(row, row + m_angleSize, 0.0) / m_angleSize);
            for (std::size_t i = 0; i < m_angleSize; ++i)
                row[i] = mean > T { 0 } ? row[i] / mean : T { 1 };
        }
    }

    T sampleIntensityWeight(const T angle, const T energy) const
    {
        std::size_t e = energy < m_energyStart ? 0 : static_cast<std::size_t>((energy - m_energyStart + T { 0.5 } * m_energyStep) / m_energyStep);
        if (e >= m_energySize)
            e = m_energySize - 1;
        std::size_t a = angle < m_angleStart ? 0 : static_cast<std::size_t>((angle - m_angleStart) / m_angleStep);
        if (a >= m_angleSize)
            a = m_angleSize - 1;
        const std::size_t w = e * m_angleSize + a;
        if (a < m_angleSize - 1) {
            const T a0 = m_angleStart + m_angleStep * a;
            const T a1 = m_angleStart + m_angleStep * (a + 1);
            return interp(a0, a1, m_weights[w], m_weights[w + 1], angle);
        }
        return m_weights[w];
    }
    std::size_t energySize() const { return m_energySize; }
    std::size_t angleSize() const { return m_angleSize; }
    T energyStart() const { return m_energyStart; }
    T energyStep() const { return m_energyStep; }
    T angleStart() const { return m_angleStart; }
    T angleStep() const { return m_angleStep; }
    const std::vector<T>& weights() const { return m_weights; }
    const std::string& filterName(void) const { return m_filterName; }
    void setFilterName(const std::string& n) { m_filterName = n; }

private:
    T m_energyStep = 2.0;
    T m_energyStart = 10.0;
    std::size_t m_energySize = 65;
    T m_angleStep = 0.07;
    T m_angleStart = 0.07;
    std::size_t m_angleSize = 5;
    std::vector<T> m_energies;
    std::vector<T> m_weights;
    std::string m_filterName;
};

// Automatic exposure control along z: exposure weight interpolated from the
// slice mass of the world via a mass -> relative tube current table.
template <Floating T = double>
class AECFilter {
public:
    AECFilter(const std::vector<T>& density, const std::array<T, 3> spacing, const std::array<std::size_t, 3> dim, const std::vector<T>& exposure)
    {
        generateMassWeightMap(density, spacing, dim, exposure);
    }
    AECFilter(const std::vector<T>& mass, const std::vector<T>& intensity)
        : m_mass(mass)
        , m_massIntensity(intensity)
        , m_positionIntensity(1, T { 1 })
    {
        m_positionStep = 1;
        m_positionMax = 1;
    }
    T sampleIntensityWeight(const std::array<T, 3>& pos) const
    {
        const T p = pos[2];
        if (p < m_positionMin + m_positionStep)
            return m_positionIntensity.front();
        if (p >= m_positionMax - m_positionStep)
            return m_positionIntensity.back();
        const std::size_t i = static_cast<std::size_t>((p - m_positionMin) / m_positionStep);
        const T x0 = i * m_positionStep + m_positionMin;
        return interp(x0, x0 + m_positionStep, m_positionIntensity[i], m_positionIntensity[i + 1], p);
    }
    void updateFromWorld(const World<T>& w) { generatePositionWeightMap(*w.densityArray(), w.spacing(), w.dimensions(), w.origin()); }
    bool isValid() const { return m_valid; }
    const std::vector<T>& mass() const { return m_mass; }
    const std::vector<T>& massIntensity() const { return m_massIntensity; }
    const std::string& filterName(void) const { return m_filterName; }
    void setFilterName(const std::string& n) { m_filterName = n; }

protected:
    static T sliceMass(const std::vector<T>& d, std::size_t slice, std::size_t sliceSize, T area)
    {
        T s = 0;
        for (std::size_t i = 0; i < sliceSize; ++i)
            s += d[slice * sliceSize + i];
        return s * area;
    }
    void generateMassWeightMap(const std::vector<T>& d, const std::array<T, 3>& sp, const std::array<std::size_t, 3>& dim, const std::vector<T>& exposure)
    {
        m_valid = false;
        if (d.size() != dim[0] * dim[1] * dim[2] || exposure.size() != dim[2]) {
            m_mass.assign(1, 0);
            m_massIntensity.assign(1, 1);
            return;
        }
        const std::size_t ss = dim[0] * dim[1];
        std::vector<std::pair<T, T>> me(dim[2]);
        for (std::size_t k = 0; k < dim[2]; ++k)
            me[k] = { sliceMass(d, k, ss, sp[0] * sp[1]), exposure[k] };
        std::sort(me.begin(), me.end());
        T mean = 0;
        for (const auto e : exposure)
            mean += e;
        mean /= dim[2];
        m_mass.resize(dim[2]);
        m_massIntensity.resize(dim[2]);
        for (std::size_t k = 0; k < dim[2]; ++k) {
            m_mass[k] = me[k].first;
            m_massIntensity[k] = me[k].second / mean;
        }
        generatePositionWeightMap(d, sp, dim, { 0, 0, 0 });
    }
    void generatePositionWeightMap(const std::vector<T>& d, const std::array<T, 3>& sp, const std::array<std::size_t, 3>& dim, const std::array<T, 3>& origin)
    {
        m_positionMin = origin[2] - sp[2] * dim[2] * T { 0.5 };
        m_positionMax = m_positionMin + sp[2] * dim[2];
        m_positionStep = (m_positionMax - m_positionMin) / dim[2];
        m_positionIntensity.resize(dim[2]);
        const std::size_t ss = dim[0] * dim[1];
        for (std::size_t k = 0; k < dim[2]; ++k)
            m_positionIntensity[k] = massIntensity(sliceMass(d, k, ss, sp[0] * sp[1]));
        m_valid = true;
    }
    T massIntensity(T mass) const
    {
        if (mass <= m_mass.front())
            return m_massIntensity.front();
        if (mass >= m_mass.back())
            return m_massIntensity.back();
        return interpolate(m_mass, m_massIntensity, mass);
    }

private:
    std::vector<T> m_mass;
    std::vector<T> m_massIntensity;
    T m_positionStep = 0, m_positionMin = 0, m_positionMax = 0;
    std::vector<T> m_positionIntensity;
    std::string m_filterName;
    bool m_valid = false;
};
}
