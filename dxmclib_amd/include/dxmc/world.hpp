// Voxel world (reference world.hpp:37-227) and the CTDI phantom (:229-390).
// Voxel layout is x fastest: index = z*nx*ny + y*nx + x. Extents are centred
// on the origin; the "safe" extent is nudged one ulp inwards so that a point
// inside it always maps to a valid voxel.
#pragma once

#include "dxmc/floating.hpp"
#include "dxmc/material.hpp"
#include "dxmc/vectormath.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

namespace dxmc {

template <Floating T = double>
class World {
public:
    World() { updateMatrixExtent(); }
    virtual ~World() = default;

    void setDimensions(const std::array<std::size_t, 3>& d)
    {
        m_valid = false;
        m_dimensions = d;
        updateMatrixExtent();
    }
    void setSpacing(const std::array<T, 3>& s)
    {
        m_valid = false;
        m_spacing = s;
        updateMatrixExtent();
    }
    void setOrigin(const std::array<T, 3>& o)
    {
        m_valid = false;
        m_origin = o;
        updateMatrixExtent();
    }
    void setDirectionCosines(const std::array<T, 6>& c)
    {
        m_valid = false;
        m_directionCosines = c;
        vectormath::normalize(m_directionCosines.data());
        vectormath::normalize(&m_directionCosines[3]);
    }

    std::size_t size(void) const { return m_dimensions[0] * m_dimensions[1] * m_dimensions[2]; }
    const std::array<std::size_t, 3>& dimensions(void) const { return m_dimensions; }
    const std::array<T, 3>& spacing(void) const { return m_spacing; }
    const std::array<T, 3>& origin(void) const { return m_origin; }
    const std::array<T, 6>& directionCosines(void) const { return m_directionCosines; }
    std::array<T, 3> depthDirection(void) const
    {
        std::array<T, 3> d;
        vectormath::cross(m_directionCosines.data(), d.data());
        return d;
    }

    void setDensityArray(std::shared_ptr<std::vector<T>> a)
    {
        m_valid = false;
        m_density = a;
    }
    std::shared_ptr<std::vector<T>> densityArray(void) { return m_density; }
    const std::shared_ptr<std::vector<T>> densityArray(void) const { return m_density; }
    void setMaterialIndexArray(std::shared_ptr<std::vector<std::uint8_t>> a)
    {
        m_valid = false;
        m_materialIndex = a;
    }
    std::shared_ptr<std::vector<std::uint8_t>> materialIndexArray(void) { return m_materialIndex; }
    const std::shared_ptr<std::vector<std::uint8_t>> materialIndexArray(void) const { return m_materialIndex; }
    void setMeasurementMapArray(std::shared_ptr<std::vector<std::uint8_t>> a)
    {
        m_valid = false;
        m_measurementMap = a;
    }
    std::shared_ptr<std::vector<std::uint8_t>> measurementMapArray(void) { return m_measurementMap; }
    const std::shared_ptr<std::vector<std::uint8_t>> measurementMapArray(void) const { return m_measurementMap; }

    const std::vector<Material>& materialMap(void) const { return m_materialMap; }
    bool addMaterialToMap(const Material& m)
    {
        m_valid = false;
        if (!m.isValid())
            return false;
        m_materialMap.push_back(m);
        return true;
    }
    void clearMaterialMap(void)
    {
        m_valid = false;
        m_materialMap.clear();
    }
    const std::array<T, 6>& matrixExtent(void) const { return m_matrixExtent; }
    const std::array<T, 6>& matrixExtentSafe(void) const { return m_matrixExtentSafe; }

    void makeValid(void)
    {
        if (!m_valid)
            m_valid = validate();
    }
    bool isValid(void) const { return m_valid; }
    [[nodiscard]] bool isValid(void)
    {
        makeValid();
        return m_valid;
    }
    // true when a measurement map was supplied and it flags at least one voxel
    bool hasMeasurements() const
    {
        return m_measurementMap && std::any_of(m_measurementMap->cbegin(), m_measurementMap->cend(), [](auto v) { return v != 0; });
    }

protected:
    void updateMatrixExtent(void)
    {
        for (std::size_t i = 0; i < 3; i++) {
            const T half = (m_dimensions[i] * m_spacing[i]) * T { 0.5 };
            const T lo = m_origin[i] - half;
            const T hi = m_origin[i] + half;
            m_matrixExtentSafe[i * 2] = std::nextafter(lo, hi);
            m_matrixExtentSafe[i * 2 + 1] = std::nextafter(hi, lo);
            m_matrixExtent[i * 2] = lo;
            m_matrixExtent[i * 2 + 1] = hi;
        }
    }

    // reference world.hpp:182-227: sizes, spacing, arrays, valid materials,
    // orthogonal cosines and material indices inside the material map
    bool validate()
    {
        if (size() == 0)
            return false;
        if ((m_spacing[0] * m_spacing[1] * m_spacing[2]) <= T { 0 })
            return false;
        if (!m_density || !m_materialIndex)
            return false;
        if (!m_measurementMap)
            m_measurementMap = std::make_shared<std::vector<std::uint8_t>>(size(), 0);
        const auto n = size();
        if (m_density->size() != n || m_materialIndex->size() != n || m_measurementMap->size() != n)
            return false;
        for (const auto& m : m_materialMap)
            if (!m.isValid())
                return false;
        const auto depth = depthDirection();
        const T ortho = vectormath::dot(depth.data(), m_directionCosines.data()) + vectormath::dot(depth.data(), &m_directionCosines[3])
            + vectormath::dot(&m_directionCosines[0], &m_directionCosines[3]);
        if (std::abs(ortho) > T { 0.001 })
            return false;
        // every index below the number of materials (unsigned: the largest
        // decides; a branch-free maximum, about 1 ms per 1e7 voxels)
        std::uint8_t mx = 0;
        for (const std::uint8_t v : *m_materialIndex)
            mx = v > mx ? v : mx;
        return mx < m_materialMap.size();
    }

    std::array<T, 3> m_spacing = { 1, 1, 1 };
    std::array<T, 3> m_origin = { 0, 0, 0 };
    std::array<T, 6> m_directionCosines = { 1, 0, 0, 0, 1, 0 };
    std::array<std::size_t, 3> m_dimensions = { 0, 0, 0 };
    std::array<T, 6> m_matrixExtent { 0, 0, 0, 0, 0, 0 };
    std::array<T, 6> m_matrixExtentSafe { 0, 0, 0, 0, 0, 0 };
    std::shared_ptr<std::vector<T>> m_density = nullptr;
    std::shared_ptr<std::vector<std::uint8_t>> m_materialIndex = nullptr;
    std::shared_ptr<std::vector<std::uint8_t>> m_measurementMap = nullptr;
    std::vector<Material> m_materialMap;
    bool m_valid = false;
};

// CTDI PMMA phantom used for CT dose calibration: a PMMA cylinder of the
// given diameter (mm) in air, 1x1x2.5 mm voxels, 60 slices, five air holes
// (centre + four at 13 mm depth) whose voxels within +-50 mm of the centre are
// flagged in the measurement map (forced photoelectric scoring).
template <Floating T = double>
class CTDIPhantom final : public World<T> {
public:
    enum class HolePosition { Center,
        West,
        East,
        South,
        North };

    CTDIPhantom(std::size_t diameter = 320)
    {
        this->setDirectionCosines({ 1, 0, 0, 0, 1, 0 });
        const std::array<T, 3> sp { 1, 1, 2.5 };
        this->setSpacing(sp);
        this->setOrigin({ 0, 0, 0 });
        const std::size_t nxy = diameter % 2 == 0 ? diameter + 3 : diameter + 2;
        this->setDimensions({ nxy, nxy, 60 });
        Material air("Air, Dry (near sea level)");
        Material pmma("Polymethyl Methacralate (Lucite, Perspex)");
        this->addMaterialToMap(air);
        this->addMaterialToMap(air); // index 1: air inside the holes
        this->addMaterialToMap(pmma);

        const T holeRadius = T { 13.1 } / T { 2 };
        const T holeDepth = T { 13 };
        const T radius = static_cast<T>(diameter) / T { 2 };
        const auto& dim = this->dimensions();

        auto dens = std::make_shared<std::vector<T>>(this->size(), airDensity());
        auto mat = std::make_shared<std::vector<std::uint8_t>>(this->size(), 0);
        auto meas = std::make_shared<std::vector<std::uint8_t>>(this->size(), 0);
        this->setDensityArray(dens);
        this->setMaterialIndexArray(mat);
        this->setMeasurementMapArray(meas);

        const std::array<std::size_t, 2> dim2 = { dim[0], dim[1] };
        const std::array<T, 2> sp2 = { sp[0], sp[1] };
        const std::array<T, 2> centre = { dim[0] * sp[0] * T { 0.5 }, dim[1] * sp[1] * T { 0.5 } };
        std::array<std::array<T, 2>, 5> holes = { { { 0, 0 }, { 0, radius - holeDepth }, { 0, -radius + holeDepth }, { radius - holeDepth, 0 }, { -radius + holeDepth, 0 } } };
        for (auto& h : holes)
            for (std::size_t i = 0; i < 2; ++i)
                h[i] += centre[i];

        for (std::size_t k = 0; k < dim[2]; ++k) {
            const T z = sp[2] * (k - dim[2] * 0.5 + 0.5);
            if (std::abs(z) <= 50.0) {
                const std::size_t off = k * dim[0] * dim[1];
                for (std::size_t i = 0; i < 5; ++i)
                    for (auto idx : circleIndices2D(dim2, sp2, holes[i], holeRadius))
                        m_holes[i].push_back(idx + off);
            }
        }
        const auto disc = circleIndices2D(dim2, sp2, centre, radius);
        std::array<std::vector<std::size_t>, 5> holeDiscs;
        for (std::size_t i = 0; i < 5; ++i)
            holeDiscs[i] = circleIndices2D(dim2, sp2, holes[i], holeRadius);
        for (std::size_t k = 0; k < dim[2]; ++k) {
            const std::size_t off = k * dim[0] * dim[1];
            for (auto idx : disc) {
                (*dens)[idx + off] = static_cast<T>(pmma.standardDensity());
                (*mat)[idx + off] = 2;
            }
            for (const auto& hd : holeDiscs)
                for (auto idx : hd) {
                    (*dens)[idx + off] = static_cast<T>(air.standardDensity());
                    (*mat)[idx + off] = 1;
                }
        }
        for (const auto& h : m_holes)
            for (auto idx : h)
                (*meas)[idx] = 1;
        this->makeValid();
    }

    const std::vector<std::size_t>& holeIndices(HolePosition p) const
    {
        switch (p) {
        case HolePosition::West:
            return m_holes[4];
        case HolePosition::East:
            return m_holes[3];
        case HolePosition::North:
            return m_holes[2];
        case HolePosition::South:
            return m_holes[1];
        default:
            return m_holes[0];
        }
    }
    static constexpr T airDensity() { return T { 0.001205 }; }
    static constexpr std::uint64_t ctdiMinHistories() { return 100000000; }

protected:
    // voxel (i,j) belongs to the disc when the offset of its centre from
    // 'centre', measured as centre - (i - 1/2) * spacing, lies within radius
    static std::vector<std::size_t> circleIndices2D(const std::array<std::size_t, 2>& dim, const std::array<T, 2>& spacing, const std::array<T, 2>& centre, T radius)
    {
        std::vector<std::size_t> idx;
        const int lim[4] = {
            std::max(static_cast<int>((centre[0] - radius) / spacing[0]), 0),
            std::min(static_cast<int>((centre[0] + radius) / spacing[0]) + 1, static_cast<int>(dim[0])),
            std::max(static_cast<int>((centre[1] - radius) / spacing[1]), 0),
            std::min(static_cast<int>((centre[1] + radius) / spacing[1]) + 1, static_cast<int>(dim[1])),
        };
        const T r2 = radius * radius;
        for (int i = lim[0]; i < lim[1]; ++i) {
            const T px = centre[0] - i * spacing[0] + spacing[0] / 2;
            for (int j = lim[2]; j < lim[3]; ++j) {
                const T py = centre[1] - j * spacing[1] + spacing[1] / 2;
                if (px * px + py * py <= r2)
                    idx.push_back(static_cast<std::size_t>(i) + static_cast<std::size_t>(j) * dim[0]);
            }
        }
        return idx;
    }

private:
    std::array<std::vector<std::size_t>, 5> m_holes;
};
}
