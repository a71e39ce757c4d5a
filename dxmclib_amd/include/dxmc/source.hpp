// Radiation sources (reference source.hpp:46-1367): each produces the
// exposure table the transport runs over and the factor that converts energy
// imparted to absolute dose.
//
//  PencilSource       mono-energetic pencil beam, air kerma calibration
//  IsotropicSource    fixed cone with a user spectrum
//  IsotropicCTSource  the same rotated about z
//  DXSource           projection radiography (DAP calibration)
//  CBCTSource         cone beam CT arc (DAP calibration)
//  CTAxialSource      axial CT, one rotation per step (CTDIvol calibration)
//  CTSpiralSource     helical CT (CTDIvol calibration)
//  CTAxialDualSource  axial dual-source CT: two tubes, alternating exposures
//  CTSpiralDualSource helical dual-source CT
//  CTTopogramSource   CT scout: the tube fixed at the start angle, moving in z
// CT calibration runs a nested transport in a CTDI phantom on the GPU.
#pragma once

#include "dxmc/beamfilters.hpp"
#include "dxmc/constants.hpp"
#include "dxmc/dxmcrandom.hpp"
#include "dxmc/exposure.hpp"
#include "dxmc/floating.hpp"
#include "dxmc/lowenergycorrectionmodel.hpp"
#include "dxmc/progressbar.hpp"
#include "dxmc/transport.hpp"
#include "dxmc/tube.hpp"
#include "dxmc/world.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace dxmc {

template <Floating T = double>
class Source {
public:
    enum class Type {
        None,
        CTSpiral,
        CTAxial,
        DX,
        CTDual,
        Pencil,
        Isotropic,
        IsotropicCT,
        CTTopogram,
        CBCT,
        Other
    };
    Source() = default;
    virtual ~Source() = default;

    virtual Exposure<T> getExposure(std::uint64_t i) const = 0;
    virtual T maxPhotonEnergyProduced() const { return Tube<T>::maxVoltage(); }
    void setPosition(const std::array<T, 3>& p) { m_position = p; }
    void setPosition(T x, T y, T z) { m_position = { x, y, z }; }
    std::array<T, 3>& position(void) { return m_position; }
    const std::array<T, 3>& position(void) const { return m_position; }
    void setDirectionCosines(const std::array<T, 6>& c)
    {
        m_directionCosines = c;
        vectormath::normalize(&m_directionCosines[0]);
        vectormath::normalize(&m_directionCosines[3]);
    }
    const std::array<T, 6>& directionCosines(void) const { return m_directionCosines; }
    std::array<T, 6>& directionCosines(void) { return m_directionCosines; }
    void setHistoriesPerExposure(std::uint64_t h) { m_historiesPerExposure = h; }
    std::uint64_t historiesPerExposure(void) const { return m_historiesPerExposure; }
    virtual std::uint64_t totalExposures(void) const = 0;
    Type type() const { return m_type; }
    virtual T getCalibrationValue(LOWENERGYCORRECTION model = LOWENERGYCORRECTION::NONE, ProgressBar<T>* progress = nullptr) const = 0;
    virtual bool isValid(void) const = 0;
    virtual bool validate(void) = 0;
    virtual void updateFromWorld(const World<T>&) { }

protected:
    std::array<T, 3> m_position = { 0, 0, 0 };
    std::array<T, 6> m_directionCosines = { 1, 0, 0, 0, 1, 0 };
    std::uint64_t m_historiesPerExposure = 1000000;
    Type m_type = Type::None;
};

template <Floating T = double>
class PencilSource final : public Source<T> {
public:
    PencilSource() { this->m_type = Source<T>::Type::Pencil; }
    Exposure<T> getExposure(std::uint64_t) const override
    {
        Exposure<T> e(this->m_position, this->m_directionCosines, std::array<T, 2> { 0, 0 }, this->m_historiesPerExposure);
        e.setMonoenergeticPhotonEnergy(m_photonEnergy);
        return e;
    }
    void setPhotonEnergy(T e) { m_photonEnergy = std::min(std::max(e, T { 1 }), ELECTRON_REST_MASS<T>()); }
    T photonEnergy() const { return m_photonEnergy; }
    T maxPhotonEnergyProduced() const override { return m_photonEnergy; }
    std::uint64_t totalExposures(void) const override { return m_totalExposures; }
    void setTotalExposures(std::uint64_t n)
    {
        if (n > 0)
            m_totalExposures = n;
    }
    void setAirDose(T gycm2)
    {
        if (gycm2 > 0)
            m_airDose = gycm2;
    }
    T airDose(void) const { return m_airDose; }
    // air dose / (N * E * (mu_en/rho)_air * keV->mJ)
    T getCalibrationValue(LOWENERGYCORRECTION, ProgressBar<T>* = nullptr) const override
    {
        Material air("Air, Dry (near sea level)");
        const T n = static_cast<T>(totalExposures() * this->historiesPerExposure());
        const T mea = static_cast<T>(air.getMassEnergyAbsorbtion(m_photonEnergy));
        return m_airDose / (n * m_photonEnergy * mea * KEV_TO_MJ<T>());
    }
    bool isValid(void) const override { return true; }
    bool validate(void) override { return true; }

private:
    T m_photonEnergy = 100;
    T m_airDose = 1;
    std::uint64_t m_totalExposures = 10;
};

template <Floating T = double>
class IsotropicSource : public Source<T> {
public:
    IsotropicSource()
        : m_specterDistribution(std::vector<T> { 1 }, std::vector<T> { 60 })
    {
        this->m_type = Source<T>::Type::Isotropic;
    }
    Exposure<T> getExposure(std::uint64_t) const override
    {
        return Exposure<T>(this->m_position, this->m_directionCosines, m_collimationAngles, this->m_historiesPerExposure, T { 1 }, &m_specterDistribution);
    }
    T maxPhotonEnergyProduced() const override { return m_maxPhotonEnergy; }
    void setTotalExposures(std::uint64_t n) { m_totalExposures = n; }
    std::uint64_t totalExposures() const override { return m_totalExposures; }
    T getCalibrationValue(LOWENERGYCORRECTION, ProgressBar<T>* = nullptr) const override { return T { 1 }; }
    bool isValid() const override { return true; }
    bool validate() override { return true; }
    void setSpecter(const std::vector<T>& weights, const std::vector<T>& energies)
    {
        m_maxPhotonEnergy = *std::max_element(energies.cbegin(), energies.cend());
        m_specterDistribution = SpecterDistribution<T>(weights, energies);
    }
    void setCollimationAngles(T x0, T x1, T y0, T y1)
    {
        m_collimationAngles[0] = std::clamp(x0, -PI_VAL<T>() / 2, PI_VAL<T>() / 2);
        m_collimationAngles[1] = std::clamp(x1, -PI_VAL<T>() / 2, PI_VAL<T>() / 2);
        m_collimationAngles[2] = std::clamp(y0, -PI_VAL<T>(), PI_VAL<T>());
        m_collimationAngles[3] = std::clamp(y1, -PI_VAL<T>(), PI_VAL<T>());
    }
    void setCollimationAngles(T x, T y) { setCollimationAngles(-x / 2, x / 2, -y / 2, y / 2); }
    const std::array<T, 4>& collimationAngles() const { return m_collimationAngles; }
    const SpecterDistribution<T>& specterDistribution() const { return m_specterDistribution; }

protected:
    std::uint64_t m_totalExposures = 1;
    std::array<T, 4> m_collimationAngles = { 0, 0, 0, 0 };
    SpecterDistribution<T> m_specterDistribution;
    T m_maxPhotonEnergy = 60;
};

template <Floating T = double>
class IsotropicCTSource final : public IsotropicSource<T> {
public:
    IsotropicCTSource() { this->m_type = Source<T>::Type::IsotropicCT; }
    Exposure<T> getExposure(std::uint64_t i) const override
    {
        const std::array<T, 3> axis = { 0, 0, 1 };
        const T angle = (i * 2 * PI_VAL<T>()) / this->m_totalExposures;
        auto c = this->m_directionCosines;
        auto p = this->m_position;
        vectormath::rotate(p.data(), axis.data(), angle);
        vectormath::rotate(&c[0], axis.data(), angle);
        vectormath::rotate(&c[3], axis.data(), angle);
        return Exposure<T>(p, c, this->m_collimationAngles, this->m_historiesPerExposure, T { 1 }, &this->m_specterDistribution);
    }
};

// Sources with a tube and a field defined at the detector distance.
template <Floating T = double>
class DAPSource : public Source<T> {
public:
    DAPSource()
    {
        m_tube.setAlFiltration(2.0);
        updateFieldSize(m_fieldSize);
        this->setDirectionCosines(zeroDirectionCosines());
    }
    Tube<T>& tube(void)
    {
        m_specterValid = false;
        return m_tube;
    }
    const Tube<T>& tube(void) const { return m_tube; }
    T maxPhotonEnergyProduced() const override { return m_tube.voltage(); }
    void setCollimationAngles(const std::array<T, 2>& a) { updateCollimationAngles({ std::abs(a[0]), std::abs(a[1]) }); }
    const std::array<T, 2>& collimationAngles(void) const { return m_collimationAngles; }
    void setCollimationAnglesDeg(const std::array<T, 2>& a) { updateCollimationAngles({ std::abs(a[0]) * DEG_TO_RAD<T>(), std::abs(a[1]) * DEG_TO_RAD<T>() }); }
    void setFieldSize(const std::array<T, 2>& mm) { updateFieldSize({ std::abs(mm[0]), std::abs(mm[1]) }); }
    const std::array<T, 2>& fieldSize(void) const { return m_fieldSize; }
    void setSourceDetectorDistance(T mm)
    {
        m_sdd = std::abs(mm);
        updateFieldSize(m_fieldSize);
    }
    T sourceDetectorDistance() const { return m_sdd; }
    // primary angle about z, secondary about x, then the tube rotation about the beam
    void setSourceAngles(T primary, T secondary)
    {
        constexpr T ERR = T { 1E-6 };
        secondary = std::clamp(secondary, -PI_VAL<T>() / 2 + ERR, PI_VAL<T>() / 2 - ERR);
        while (primary > PI_VAL<T>())
            primary -= PI_VAL<T>();
        while (primary < -PI_VAL<T>())
            primary += PI_VAL<T>();
        auto c = zeroDirectionCosines();
        const std::array<T, 3> z = { 0, 0, 1 }, x = { 1, 0, 0 };
        vectormath::rotate(c.data(), z.data(), primary);
        vectormath::rotate(&c[3], z.data(), primary);
        vectormath::rotate(c.data(), x.data(), -secondary);
        vectormath::rotate(&c[3], x.data(), -secondary);
        std::array<T, 3> beam;
        vectormath::cross(c.data(), beam.data());
        vectormath::rotate(c.data(), beam.data(), m_tubeRotationAngle);
        vectormath::rotate(&c[3], beam.data(), m_tubeRotationAngle);
        this->setDirectionCosines(c);
    }
    void setSourceAnglesDeg(T p, T s) { setSourceAngles(p * DEG_TO_RAD<T>(), s * DEG_TO_RAD<T>()); }
    void setTubeRotation(T angle)
    {
        const T diff = angle - m_tubeRotationAngle;
        auto c = this->directionCosines();
        std::array<T, 3> beam;
        vectormath::cross(c.data(), beam.data());
        vectormath::rotate(c.data(), beam.data(), diff);
        vectormath::rotate(&c[3], beam.data(), diff);
        this->setDirectionCosines(c);
        m_tubeRotationAngle = angle;
    }
    T tubeRotation() const { return m_tubeRotationAngle; }
    void setTubeRotationDeg(T a) { setTubeRotation(a * DEG_TO_RAD<T>()); }
    void setDap(T gycm2)
    {
        if (gycm2 > 0)
            m_dap = gycm2;
    }
    T dap(void) const { return m_dap; }
    // DAP / (N * sum_E E w(E) (mu_en/rho)_air(E))
    T getCalibrationValue(LOWENERGYCORRECTION, ProgressBar<T>* = nullptr) const override
    {
        const auto s = m_tube.getSpecter(true);
        Material air("Air, Dry (near sea level)");
        T out = 0;
        for (const auto& [keV, w] : s)
            out += keV * w * static_cast<T>(air.getMassEnergyAbsorbtion(keV));
        out *= this->totalExposures() * this->historiesPerExposure();
        return m_dap / out;
    }
    bool isValid(void) const override { return m_specterValid; }
    bool validate(void) override
    {
        updateSpecterDistribution();
        return m_specterValid;
    }
    void setModelHeelEffect(bool on)
    {
        m_modelHeelEffect = on;
        m_specterValid = false;
    }
    bool modelHeelEffect() const { return m_modelHeelEffect; }
    const SpecterDistribution<T>* specterDistribution() const { return m_specterDistribution.get(); }
    const HeelFilter<T>* heelFilter() const { return m_heelFilter.get(); }

protected:
    void updateFieldSize(const std::array<T, 2>& f)
    {
        for (std::size_t i = 0; i < 2; ++i) {
            m_fieldSize[i] = f[i];
            m_collimationAngles[i] = std::atan(m_fieldSize[i] * T { 0.5 } / m_sdd) * T { 2 };
        }
        m_specterValid = false;
    }
    void updateCollimationAngles(const std::array<T, 2>& a)
    {
        for (std::size_t i = 0; i < 2; ++i) {
            m_collimationAngles[i] = a[i];
            m_fieldSize[i] = std::tan(m_collimationAngles[i] / 2) * m_sdd * 2;
        }
        m_specterValid = false;
    }
    void updateSpecterDistribution()
    {
        if (m_specterValid)
            return;
        const auto e = m_tube.getEnergy();
        m_specterDistribution = std::make_shared<SpecterDistribution<T>>(m_tube.getSpecter(e), e);
        m_heelFilter = m_modelHeelEffect ? std::make_shared<HeelFilter<T>>(m_tube, m_collimationAngles[1]) : nullptr;
        m_specterValid = true;
    }
    static constexpr std::array<T, 6> zeroDirectionCosines() { return { -1, 0, 0, 0, 0, 1 }; }

    T m_sdd = 1000;
    T m_dap = 1;
    std::array<T, 2> m_fieldSize = { 100, 100 };
    std::array<T, 2> m_collimationAngles = { 0, 0 };
    Tube<T> m_tube;
    T m_tubeRotationAngle = 0;
    std::shared_ptr<SpecterDistribution<T>> m_specterDistribution;
    std::shared_ptr<HeelFilter<T>> m_heelFilter;
    bool m_modelHeelEffect = true;
    bool m_specterValid = false;
};

template <Floating T = double>
class DXSource final : public DAPSource<T> {
public:
    DXSource() { this->m_type = Source<T>::Type::DX; }
    Exposure<T> getExposure(std::uint64_t) const override
    {
        return Exposure<T>(tubePosition(), this->m_directionCosines, this->m_collimationAngles, this->m_historiesPerExposure, T { 1 },
            this->m_specterDistribution.get(), this->m_heelFilter.get());
    }
    std::uint64_t totalExposures(void) const override { return m_totalExposures; }
    void setTotalExposures(std::uint64_t n) { m_totalExposures = std::max<std::uint64_t>(n, 1); }
    // the tube sits one SDD behind the position along the beam
    std::array<T, 3> tubePosition(void) const
    {
        std::array<T, 3> beam, pos;
        vectormath::cross(this->m_directionCosines.data(), beam.data());
        for (std::size_t i = 0; i < 3; ++i)
            pos[i] = this->m_position[i] - beam[i] * this->m_sdd;
        return pos;
    }

private:
    std::uint64_t m_totalExposures = 1000;
};

template <Floating T = double>
class CBCTSource final : public DAPSource<T> {
public:
    CBCTSource()
    {
        this->m_type = Source<T>::Type::CBCT;
        this->setSourceDetectorDistance(500);
    }
    std::array<T, 3> rotationAxis() const { return { this->m_directionCosines[3], this->m_directionCosines[4], this->m_directionCosines[5] }; }
    void setSpanAngle(T a)
    {
        m_angleSpan = std::max(a, m_angleStep);
        m_totalExposures = std::max<std::size_t>(static_cast<std::size_t>(m_angleSpan / m_angleStep), 2);
    }
    void setSpanAngleDeg(T a) { setSpanAngle(a * DEG_TO_RAD<T>()); }
    T spanAngle() const { return m_angleSpan; }
    void setStepAngle(T a)
    {
        m_angleStep = std::max(a, PI_VAL<T>() / T { 360 });
        m_totalExposures = std::max<std::size_t>(static_cast<std::size_t>(m_angleSpan / m_angleStep), 2);
    }
    void setStepAngleDeg(T a) { setStepAngle(a * DEG_TO_RAD<T>()); }
    T stepAngle() const { return m_angleStep; }
    Exposure<T> getExposure(std::uint64_t i) const override
    {
        const T angle = i * m_angleStep;
        const auto tube = tubePosition();
        const auto& iso = this->position();
        std::array<T, 3> p;
        for (std::size_t k = 0; k < 3; ++k)
            p[k] = tube[k] - iso[k];
        const auto axis = rotationAxis();
        vectormath::rotate(p.data(), axis.data(), angle);
        for (std::size_t k = 0; k < 3; ++k)
            p[k] += iso[k];
        auto c = this->m_directionCosines;
        vectormath::rotate(c.data(), axis.data(), angle);
        vectormath::rotate(&c[3], axis.data(), angle);
        return Exposure<T>(p, c, this->m_collimationAngles, this->m_historiesPerExposure, T { 1 }, this->m_specterDistribution.get(),
            this->m_heelFilter.get());
    }
    std::uint64_t totalExposures(void) const override { return m_totalExposures; }
    std::array<T, 3> tubePosition(void) const
    {
        std::array<T, 3> beam, pos;
        vectormath::cross(this->m_directionCosines.data(), beam.data());
        for (std::size_t i = 0; i < 3; ++i)
            pos[i] = this->m_position[i] - beam[i] * this->m_sdd * T { 0.5 };
        return pos;
    }

private:
    std::size_t m_totalExposures = 180;
    T m_angleSpan = PI_VAL<T>();
    T m_angleStep = PI_VAL<T>() / T { 180 };
};

// CTDIw of CT calibration runs, kept per the run's complete input: the
// exposure records and all tables of the prepared nested transport (tube
// spectrum, heel and bowtie tables, geometry and weights of every exposure,
// histories, attenuation and sampling tables), the low-energy model and
// the phantom diameter. A run is deterministic (fixed seed, one random
// stream per history), so an equal input gives the same CTDIw; a repeated
// DOSE-mode CT request then skips its nested transport of 1e8 or more
// histories. At most 256 entries; an unfinished (cancelled) run is not kept.
template <Floating T>
class CalibrationCache {
public:
    template <typename Prepared>
    static std::string key(const Prepared& p, int model, std::uint64_t diameter)
    {
        std::string k;
        auto put = [&k](const void* data, std::size_t bytes) { k.append(static_cast<const char*>(data), bytes); };
        auto vec = [&put](const auto& v) {
            const std::uint64_t n = v.size();
            put(&n, sizeof n);
            put(v.data(), v.size() * sizeof(v[0]));
        };
        put(&model, sizeof model);
        put(&diameter, sizeof diameter);
        const TableSet& t = p.tables;
        vec(t.grid);
        vec(t.coefficients);
        vec(t.rita);
        vec(t.spline);
        vec(t.spectrumOffset);
        vec(t.spectrumAlias);
        vec(t.spectrumProbability);
        vec(t.spectrumEnergy);
        vec(t.heelWeight);
        vec(t.bowtieOffset);
        vec(t.bowtieAngle);
        vec(t.bowtieWeight);
        for (const auto& h : t.heel) {
            put(&h.energy_start, 4 * sizeof(double));
            put(&h.energy_size, 3 * sizeof(std::uint32_t));
        }
        for (const auto& e : p.exposures) {
            put(e.position, 18 * sizeof(double)); // position .. mono_energy
            put(&e.n_histories, sizeof e.n_histories);
            put(&e.spectrum, 3 * sizeof(std::int32_t));
        }
        return k;
    }
    static bool find(const std::string& k, T& ctdiw)
    {
        std::lock_guard lock(mutex());
        const auto it = entries().find(k);
        if (it == entries().end())
            return false;
        ctdiw = it->second;
        return true;
    }
    static void store(const std::string& k, T ctdiw)
    {
        std::lock_guard lock(mutex());
        if (entries().size() >= 256)
            entries().clear();
        entries()[k] = ctdiw;
    }
    static void clear()
    {
        std::lock_guard lock(mutex());
        entries().clear();
    }

private:
    static std::mutex& mutex()
    {
        static std::mutex m;
        return m;
    }
    static std::map<std::string, T>& entries()
    {
        static std::map<std::string, T> e;
        return e;
    }
};

template <Floating T>
class CTAxialSource;

template <Floating T>
class CTBaseSource : public Source<T> {
public:
    CTBaseSource()
    {
        m_tube.setAlFiltration(7.0);
        this->setDirectionCosines({ -1, 0, 0, 0, 0, 1 });
    }
    Tube<T>& tube(void)
    {
        m_specterValid = false;
        return m_tube;
    }
    const Tube<T>& tube(void) const { return m_tube; }
    T maxPhotonEnergyProduced() const override { return m_tube.voltage(); }
    void setBowTieFilter(std::shared_ptr<BowTieFilter<T>> f) { m_bowTieFilter = f; }
    std::shared_ptr<BowTieFilter<T>> bowTieFilter(void) { return m_bowTieFilter; }
    const std::shared_ptr<BowTieFilter<T>> bowTieFilter() const { return m_bowTieFilter; }
    void setSourceDetectorDistance(T sdd)
    {
        m_sdd = std::abs(sdd);
        m_specterValid = false;
    }
    T sourceDetectorDistance(void) const { return m_sdd; }
    void setCollimation(T c)
    {
        m_collimation = std::abs(c);
        m_specterValid = false;
    }
    T collimation(void) const { return m_collimation; }
    void setFieldOfView(T fov) { m_fov = std::abs(fov); }
    T fieldOfView(void) const { return m_fov; }
    void setGantryTiltAngle(T a) { m_gantryTiltAngle = std::clamp(a, -PI_VAL<T>(), PI_VAL<T>()); }
    T gantryTiltAngle() const { return m_gantryTiltAngle; }
    void setGantryTiltAngleDeg(T a) { setGantryTiltAngle(a * DEG_TO_RAD<T>()); }
    void setStartAngle(T a) { m_startAngle = a; }
    T startAngle(void) const { return m_startAngle; }
    void setStartAngleDeg(T a) { m_startAngle = DEG_TO_RAD<T>() * a; }
    virtual void setScanLenght(T l) { m_scanLenght = std::abs(l); }
    T scanLenght(void) const { return m_scanLenght; }
    void setCtdiVol(T c)
    {
        if (c > 0)
            m_ctdivol = c;
    }
    T ctdiVol(void) const { return m_ctdivol; }
    void setCtdiPhantomDiameter(std::uint64_t mm) { m_ctdiPhantomDiameter = std::max<std::uint64_t>(mm, 160); }
    std::uint64_t ctdiPhantomDiameter(void) const { return m_ctdiPhantomDiameter; }
    void setModelHeelEffect(bool on)
    {
        m_modelHeelEffect = on;
        m_specterValid = false;
    }
    bool modelHeelEffect() const { return m_modelHeelEffect; }
    bool isValid(void) const override { return m_specterValid; }
    bool validate(void) override
    {
        updateSpecterDistribution();
        return m_specterValid;
    }
    const SpecterDistribution<T>* specterDistribution() const { return m_specterDistribution.get(); }
    const HeelFilter<T>* heelFilter() const { return m_heelFilter.get(); }

protected:
    virtual void updateSpecterDistribution()
    {
        if (m_specterValid)
            return;
        const auto e = m_tube.getEnergy();
        m_specterDistribution = std::make_shared<SpecterDistribution<T>>(m_tube.getSpecter(e), e);
        const T span = std::atan(m_collimation * T { 0.5 } / m_sdd) * T { 2 };
        m_heelFilter = m_modelHeelEffect ? std::make_shared<HeelFilter<T>>(m_tube, span) : nullptr;
        m_specterValid = true;
    }

    // Gantry geometry of every CT exposure: the tube starts at (0, -SDD/2, 0)
    // (the SDD of the source, also for a second tube), the gantry is tilted
    // about x, the tube and the first cosine rotate about the (tilted) second
    // cosine by 'angle', and the tube moves by zOffset along z.
    void gantryPose(T angle, T zOffset, std::array<T, 3>& pos, std::array<T, 6>& c) const
    {
        pos = { 0, -m_sdd / T { 2 }, 0 };
        c = this->m_directionCosines;
        T* rotationAxis = &c[3];
        T* otherAxis = &c[0];
        const std::array<T, 3> tilt = { 1, 0, 0 };
        auto tiltCorrection = pos;
        vectormath::rotate(tiltCorrection.data(), tilt.data(), m_gantryTiltAngle);
        vectormath::rotate(rotationAxis, tilt.data(), m_gantryTiltAngle);
        vectormath::rotate(otherAxis, tilt.data(), m_gantryTiltAngle);
        vectormath::rotate(pos.data(), rotationAxis, angle);
        pos[2] += zOffset + tiltCorrection[2];
        vectormath::rotate(otherAxis, rotationAxis, angle);
        for (std::size_t i = 0; i < 3; ++i)
            pos[i] += this->m_position[i];
    }
    std::array<T, 2> fanAngles(T fov, T sdd) const { return { std::atan(fov / sdd) * T { 2 }, std::atan(m_collimation / sdd) * T { 2 } }; }

    // CTDIw from a nested transport in a CTDI phantom (forced scoring in
    // the holes), reference source.hpp:925-988; U is CTAxialSource or
    // CTAxialDualSource
    template <typename U>
    static T ctCalibration(U& copy, LOWENERGYCORRECTION model, ProgressBar<T>* progress);

    T m_sdd = 1190;
    T m_collimation = T { 38.4 };
    T m_fov = 500;
    T m_startAngle = 0;
    T m_scanLenght = 100;
    T m_ctdivol = 1;
    T m_gantryTiltAngle = 0;
    std::uint64_t m_ctdiPhantomDiameter = 320;
    std::shared_ptr<BowTieFilter<T>> m_bowTieFilter;
    Tube<T> m_tube;
    std::shared_ptr<SpecterDistribution<T>> m_specterDistribution;
    std::shared_ptr<HeelFilter<T>> m_heelFilter;
    bool m_modelHeelEffect = true;
    bool m_specterValid = false;
};

template <Floating T>
class CTSource : public CTBaseSource<T> {
public:
    CTSource() = default;
    void setExposureAngleStep(T a) { m_exposureAngleStep = std::clamp(std::abs(a), DEG_TO_RAD<T>() / 10, PI_VAL<T>() / 2); }
    T exposureAngleStep(void) const { return m_exposureAngleStep; }
    void setExposureAngleStepDeg(T a) { setExposureAngleStep(a * DEG_TO_RAD<T>()); }
    T exposureAngleStepDeg(void) const { return m_exposureAngleStep * RAD_TO_DEG<T>(); }
    void setAecFilter(std::shared_ptr<AECFilter<T>> f) { m_aecFilter = f; }
    std::shared_ptr<AECFilter<T>> aecFilter(void) { return m_aecFilter; }
    bool useXCareFilter() const { return m_useXCareFilter; }
    void setUseXCareFilter(bool u) { m_useXCareFilter = u; }
    XCareFilter<T>& xcareFilter() { return m_xcareFilter; }
    const XCareFilter<T>& xcareFilter() const { return m_xcareFilter; }
    void updateFromWorld(const World<T>& w) override
    {
        if (m_aecFilter)
            m_aecFilter->updateFromWorld(w);
    }
    virtual std::uint64_t exposuresPerRotatition() const { return static_cast<std::size_t>(2 * PI_VAL<T>() / m_exposureAngleStep); }

protected:
    // what an exposure takes from the tube that makes it
    struct TubeSide {
        T sdd, fov, weight;
        const SpecterDistribution<T>* spectrum;
        const HeelFilter<T>* heel;
        const BeamFilter<T>* bowtie;
    };
    TubeSide tubeA() const
    {
        return { this->m_sdd, this->m_fov, T { 1 }, this->m_specterDistribution.get(), this->m_heelFilter.get(), this->m_bowTieFilter.get() };
    }
    // an axial or spiral exposure at gantry angle 'angle' and z offset, with
    // the AEC weight at the tube position and the XCare weight at the angle
    Exposure<T> gantryExposure(T angle, T zOffset, std::uint64_t histories, const TubeSide& tube) const
    {
        std::array<T, 3> pos;
        std::array<T, 6> c;
        this->gantryPose(angle, zOffset, pos, c);
        T weight = tube.weight;
        if (m_aecFilter)
            weight *= m_aecFilter->sampleIntensityWeight(pos);
        if (m_useXCareFilter)
            weight *= m_xcareFilter.sampleIntensityWeight(angle);
        return Exposure<T>(pos, c, this->fanAngles(tube.fov, tube.sdd), histories, weight, tube.spectrum, tube.heel, tube.bowtie);
    }
    Exposure<T> gantryExposure(T angle, T zOffset, std::uint64_t histories) const { return gantryExposure(angle, zOffset, histories, tubeA()); }


    T m_exposureAngleStep = DEG_TO_RAD<T>();
    std::shared_ptr<AECFilter<T>> m_aecFilter;
    XCareFilter<T> m_xcareFilter;
    bool m_useXCareFilter = false;
};

template <Floating T>
class CTSpiralSource;

template <Floating T = double>
class CTAxialSource final : public CTSource<T> {
public:
    CTAxialSource()
    {
        this->m_type = Source<T>::Type::CTAxial;
        m_step = this->m_collimation;
        this->m_scanLenght = m_step;
    }
    CTAxialSource(const CTSpiralSource<T>& other);

    Exposure<T> getExposure(std::uint64_t i) const override
    {
        const std::uint64_t perRotation = static_cast<std::uint64_t>(2 * PI_VAL<T>() / this->m_exposureAngleStep);
        const std::uint64_t rotation = i / perRotation;
        const T angle = this->m_startAngle + this->m_exposureAngleStep * (i - (rotation * perRotation));
        return this->gantryExposure(angle, m_step * rotation, this->m_historiesPerExposure);
    }
    void setStep(T step)
    {
        const T n = this->m_scanLenght / m_step;
        m_step = std::max(std::abs(step), T { 0.01 });
        setScanLenght(m_step * n);
    }
    T step(void) const { return m_step; }
    void setScanLenght(T l) override { this->m_scanLenght = std::max(m_step * std::ceil(std::abs(l) / m_step), m_step); }
    std::uint64_t totalExposures(void) const override
    {
        const std::uint64_t perRotation = static_cast<std::uint64_t>(PI_VAL<T>() * T { 2 } / this->m_exposureAngleStep);
        return perRotation * static_cast<std::uint64_t>(std::round(this->m_scanLenght / m_step));
    }
    T getCalibrationValue(LOWENERGYCORRECTION model, ProgressBar<T>* progress = nullptr) const override
    {
        auto copy = *this;
        return CTBaseSource<T>::ctCalibration(copy, model, progress);
    }

private:
    T m_step = 10;
};

template <Floating T = double>
class CTSpiralSource final : public CTSource<T> {
public:
    CTSpiralSource() { this->m_type = Source<T>::Type::CTSpiral; }
    Exposure<T> getExposure(std::uint64_t i) const override
    {
        const T angle = this->m_startAngle + this->m_exposureAngleStep * i;
        constexpr T twoPi = T { 2 } * PI_VAL<T>();
        const T z = (i * this->m_exposureAngleStep) * this->m_collimation * m_pitch / twoPi;
        return this->gantryExposure(angle, z, this->m_historiesPerExposure);
    }
    void setPitch(T p) { m_pitch = std::max(T { 0.01 }, p); }
    T pitch(void) const { return m_pitch; }
    void setScanLenght(T l) override { this->m_scanLenght = std::max(std::abs(l), this->m_collimation * m_pitch * T { 0.5 }); }
    std::uint64_t totalExposures(void) const override
    {
        constexpr T twoPi = 2 * PI_VAL<T>();
        return static_cast<std::uint64_t>(this->m_scanLenght * twoPi / (this->m_collimation * m_pitch * this->m_exposureAngleStep));
    }
    T getCalibrationValue(LOWENERGYCORRECTION model, ProgressBar<T>* progress = nullptr) const override
    {
        CTAxialSource<T> copy(*this);
        return CTBaseSource<T>::ctCalibration(copy, model, progress) * m_pitch;
    }

private:
    T m_pitch = 1;
};

// Dual-source CT (reference source.hpp:1079-1190): a second tube B on the
// same gantry, a quarter turn ahead by default. Exposures alternate between
// the tubes (even index A, odd index B); each tube has its own spectrum,
// heel and bowtie filter, field of view, SDD and start angle, and its
// exposures carry the weight 2 w_X / (w_A + w_B) with w_X = mAs_X times the
// sum of tube X's unnormalised spectrum. The tube position follows the
// source's SDD for both tubes, as in the reference.
template <Floating T = double>
class CTDualSource : public CTSource<T> {
public:
    CTDualSource()
    {
        this->m_type = Source<T>::Type::None;
        m_sddB = this->m_sdd;
        m_fovB = this->m_fov;
        m_startAngleB = this->m_startAngle + PI_VAL<T>() * T { 0.5 };
        m_tubeB.setAlFiltration(this->m_tube.AlFiltration());
    }
    T tubeAmas() const { return m_tubeAmas; }
    T tubeBmas() const { return m_tubeBmas; }
    void setTubeAmas(T mas)
    {
        this->m_specterValid = false;
        m_tubeAmas = std::max(T { 0 }, mas);
    }
    void setTubeBmas(T mas)
    {
        this->m_specterValid = false;
        m_tubeBmas = std::max(T { 0 }, mas);
    }
    Tube<T>& tubeB(void)
    {
        this->m_specterValid = false;
        return m_tubeB;
    }
    const Tube<T>& tubeB(void) const { return m_tubeB; }
    T maxPhotonEnergyProduced() const override { return std::max(this->m_tube.voltage(), m_tubeB.voltage()); }
    std::uint64_t exposuresPerRotatition() const override { return 2 * static_cast<std::size_t>((2 * PI_VAL<T>()) / this->m_exposureAngleStep); }
    void setBowTieFilterB(std::shared_ptr<BowTieFilter<T>> f) { m_bowTieFilterB = f; }
    std::shared_ptr<BowTieFilter<T>> bowTieFilterB(void) { return m_bowTieFilterB; }
    const std::shared_ptr<BowTieFilter<T>> bowTieFilterB() const { return m_bowTieFilterB; }
    void setSourceDetectorDistanceB(T sdd)
    {
        this->m_specterValid = false;
        m_sddB = std::abs(sdd);
    }
    T sourceDetectorDistanceB(void) const { return m_sddB; }
    void setFieldOfViewB(T fov) { m_fovB = std::abs(fov); }
    T fieldOfViewB(void) const { return m_fovB; }
    void setStartAngleB(T a) { m_startAngleB = a; }
    T startAngleB(void) const { return m_startAngleB; }
    void setStartAngleDegB(T a) { m_startAngleB = DEG_TO_RAD<T>() * a; }
    T startAngleDegB(void) const { return RAD_TO_DEG<T>() * m_startAngleB; }
    // the exposure weights of tube A and B (2 w_X / (w_A + w_B), after validate)
    T tubeAweight() const { return m_tubeAweight; }
    T tubeBweight() const { return m_tubeBweight; }
    const SpecterDistribution<T>* specterDistributionB() const { return m_specterDistributionB.get(); }
    const HeelFilter<T>* heelFilterB() const { return m_heelFilterB.get(); }

protected:
    void updateSpecterDistribution() override
    {
        if (this->m_specterValid)
            return;
        const auto energyA = this->m_tube.getEnergy();
        const auto energyB = m_tubeB.getEnergy();
        auto specterA = this->m_tube.getSpecter(energyA, false);
        auto specterB = m_tubeB.getSpecter(energyB, false);
        const T sumA = std::accumulate(specterA.cbegin(), specterA.cend(), T { 0 });
        const T sumB = std::accumulate(specterB.cbegin(), specterB.cend(), T { 0 });
        const T weightA = m_tubeAmas * sumA;
        const T weightB = m_tubeBmas * sumB;
        for (auto& v : specterA)
            v = v / sumA;
        for (auto& v : specterB)
            v = v / sumB;
        m_tubeAweight = weightA * T { 2 } / (weightA + weightB);
        m_tubeBweight = weightB * T { 2 } / (weightA + weightB);
        this->m_specterDistribution = std::make_shared<SpecterDistribution<T>>(specterA, energyA);
        m_specterDistributionB = std::make_shared<SpecterDistribution<T>>(specterB, energyB);
        // both heel filters span tube A's collimation angle, as in the
        // reference, and are kept regardless of the heel switch there
        const T span = std::atan(this->m_collimation * T { 0.5 } / this->m_sdd) * T { 2 };
        this->m_heelFilter = std::make_shared<HeelFilter<T>>(this->m_tube, span);
        m_heelFilterB = std::make_shared<HeelFilter<T>>(m_tubeB, span);
        this->m_specterValid = true;
    }
    // tube A for an even exposure index, tube B for an odd one, with its
    // start angle
    typename CTSource<T>::TubeSide tubeFor(std::uint64_t totalIndex, T& startAngle) const
    {
        if (totalIndex % 2 == 0) {
            startAngle = this->m_startAngle;
            return { this->m_sdd, this->m_fov, m_tubeAweight, this->m_specterDistribution.get(), this->m_heelFilter.get(),
                this->m_bowTieFilter.get() };
        }
        startAngle = m_startAngleB;
        return { m_sddB, m_fovB, m_tubeBweight, m_specterDistributionB.get(), m_heelFilterB.get(), m_bowTieFilterB.get() };
    }

    T m_sddB;
    T m_fovB;
    T m_startAngleB;
    T m_tubeAmas = 100;
    T m_tubeBmas = 100;
    T m_tubeAweight = -1;
    T m_tubeBweight = -1;
    std::shared_ptr<BowTieFilter<T>> m_bowTieFilterB;
    Tube<T> m_tubeB;
    std::shared_ptr<SpecterDistribution<T>> m_specterDistributionB;
    std::shared_ptr<HeelFilter<T>> m_heelFilterB;
};

template <Floating T>
class CTSpiralDualSource;

// axial dual-source CT (reference source.hpp:1369-1472): exposure pairs
// (A, B) at each gantry step, one rotation per table step
template <Floating T = double>
class CTAxialDualSource final : public CTDualSource<T> {
public:
    CTAxialDualSource()
    {
        this->m_type = Source<T>::Type::CTDual;
        m_step = this->m_collimation;
        this->m_scanLenght = m_step;
    }
    CTAxialDualSource(const CTSpiralDualSource<T>& other);
    Exposure<T> getExposure(std::uint64_t totalIndex) const override
    {
        T startAngle;
        const auto tube = this->tubeFor(totalIndex, startAngle);
        const std::uint64_t i = totalIndex / 2;
        const std::uint64_t perRotation = static_cast<std::uint64_t>(2 * PI_VAL<T>() / this->m_exposureAngleStep);
        const std::uint64_t rotation = i / perRotation;
        const T angle = startAngle + this->m_exposureAngleStep * (i - (rotation * perRotation));
        return this->gantryExposure(angle, m_step * rotation, this->m_historiesPerExposure, tube);
    }
    void setStep(T step)
    {
        const T n = this->m_scanLenght / m_step;
        m_step = std::max(std::abs(step), T { 0.01 });
        setScanLenght(m_step * n);
    }
    T step(void) const { return m_step; }
    void setScanLenght(T l) override { this->m_scanLenght = std::max(m_step * std::ceil(std::abs(l) / m_step), m_step); }
    std::uint64_t totalExposures(void) const override
    {
        const std::uint64_t perRotation = static_cast<std::uint64_t>(PI_VAL<T>() * T { 2 } / this->m_exposureAngleStep);
        return perRotation * static_cast<std::uint64_t>(std::round(this->m_scanLenght / m_step)) * 2;
    }
    T getCalibrationValue(LOWENERGYCORRECTION model, ProgressBar<T>* progress = nullptr) const override
    {
        auto copy = *this;
        return CTBaseSource<T>::ctCalibration(copy, model, progress);
    }

private:
    T m_step = 10;
};

// helical dual-source CT (reference source.hpp:1474-1592)
template <Floating T = double>
class CTSpiralDualSource final : public CTDualSource<T> {
public:
    CTSpiralDualSource() { this->m_type = Source<T>::Type::CTDual; }
    Exposure<T> getExposure(std::uint64_t totalIndex) const override
    {
        T startAngle;
        const auto tube = this->tubeFor(totalIndex, startAngle);
        const std::uint64_t i = totalIndex / 2;
        const T angle = startAngle + this->m_exposureAngleStep * i;
        constexpr T twoPi = T { 2 } * PI_VAL<T>();
        const T z = (i * this->m_exposureAngleStep) * this->m_collimation * m_pitch / twoPi;
        return this->gantryExposure(angle, z, this->m_historiesPerExposure, tube);
    }
    std::uint64_t totalExposures(void) const override
    {
        const auto single = static_cast<std::uint64_t>(this->m_scanLenght * 2 * PI_VAL<T>() / (this->m_collimation * m_pitch * this->m_exposureAngleStep));
        return single * 2;
    }
    T getCalibrationValue(LOWENERGYCORRECTION model, ProgressBar<T>* progress = nullptr) const override
    {
        CTAxialDualSource<T> copy(*this);
        return CTBaseSource<T>::ctCalibration(copy, model, progress) * m_pitch;
    }
    T pitch(void) const { return m_pitch; }
    void setPitch(T p) { m_pitch = std::max(T { 0.01 }, p); }
    void setScanLenght(T l) override { CTBaseSource<T>::setScanLenght(std::max(std::abs(l), this->m_collimation * m_pitch * T { 0.5 })); }

private:
    T m_pitch = 1;
};

template <Floating T>
CTAxialDualSource<T>::CTAxialDualSource(const CTSpiralDualSource<T>& o)
    : CTDualSource<T>(o)
{
    m_step = this->m_collimation;
    setScanLenght(o.scanLenght());
}

// CT topogram (scout view; reference source.hpp:1609-1700): the tube stays
// at the start angle and steps along z over the scan length, one exposure
// per millimetre (at least one). Its dose calibration is that of an axial
// scan of one collimation width holding the topogram's CTDIvol times scan
// length over collimation, scaled by the ratio of exposure counts.
template <Floating T = double>
class CTTopogramSource final : public CTBaseSource<T> {
public:
    CTTopogramSource() { this->m_type = Source<T>::Type::CTTopogram; }
    Exposure<T> getExposure(std::uint64_t i) const override
    {
        const std::uint64_t n = totalExposures();
        // one exposure: z 0 (the reference divides by n - 1 = 0 there)
        const T step = n > 1 ? this->scanLenght() / (n - 1) : T { 0 };
        std::array<T, 3> pos;
        std::array<T, 6> c;
        this->gantryPose(this->m_startAngle, step * i, pos, c);
        return Exposure<T>(pos, c, this->fanAngles(this->m_fov, this->m_sdd), this->m_historiesPerExposure, T { 1 }, this->m_specterDistribution.get(),
            this->m_heelFilter.get(), this->m_bowTieFilter.get());
    }
    std::uint64_t totalExposures() const override { return std::max<std::uint64_t>(static_cast<std::uint64_t>(std::ceil(this->scanLenght())), 1); }
    T getCalibrationValue(LOWENERGYCORRECTION model, ProgressBar<T>* progress = nullptr) const override
    {
        CTAxialSource<T> copy;
        CTBaseSource<T>& base = copy;
        base = *this;
        copy.setCtdiVol(this->ctdiVol() * this->scanLenght() / this->collimation());
        copy.setScanLenght(0);
        copy.setStep(this->m_collimation);
        // angle step matched to the number of exposures, at most 5 degrees
        constexpr T maxStep = (2 * PI_VAL<T>()) / 72;
        copy.setExposureAngleStep(std::min(2 * PI_VAL<T>() / totalExposures(), maxStep));
        const T factor = CTBaseSource<T>::ctCalibration(copy, model, progress);
        return (factor * totalExposures()) / copy.totalExposures();
    }
};

template <Floating T>
CTAxialSource<T>::CTAxialSource(const CTSpiralSource<T>& o)
    : CTSource<T>(o)
{
    this->m_type = Source<T>::Type::CTAxial;
    m_step = this->m_collimation;
    setScanLenght(o.scanLenght());
}

template <Floating T>
template <typename U>
T CTBaseSource<T>::ctCalibration(U& src, LOWENERGYCORRECTION model, ProgressBar<T>* progress)
{
    T meanWeight = 0;
    for (std::uint64_t i = 0; i < src.totalExposures(); ++i)
        meanWeight += src.getExposure(i).beamIntensityWeight();
    meanWeight /= src.totalExposures();
    src.setDirectionCosines({ -1, 0, 0, 0, 0, 1 });
    src.setPosition({ 0, 0, 0 });
    src.setScanLenght(src.collimation());
    src.setUseXCareFilter(false);
    std::uint64_t stat = CTDIPhantom<T>::ctdiMinHistories() / (src.exposuresPerRotatition() * src.historiesPerExposure());
    stat = std::max<std::uint64_t>(stat, 1);
    CTDIPhantom<T> world(src.ctdiPhantomDiameter());
    src.updateFromWorld(world);
    src.setHistoriesPerExposure(src.historiesPerExposure() * stat);
    src.validate();
    if (progress) {
        progress->setPlaneNormal(ProgressBar<T>::Axis::Z);
        progress->setPrefixMessage("CTDI calibration ");
    }
    Transport<T> transport;
    transport.setLowEnergyCorrectionModel(model);
    const auto prep = transport.prepare(world, &src);
    const std::string key = prep.valid ? CalibrationCache<T>::key(prep, static_cast<int>(model), src.ctdiPhantomDiameter()) : std::string();
    T ctdiw = 0;
    if (prep.valid && !CalibrationCache<T>::find(key, ctdiw)) {
        const auto result = transport.execute(world, &src, prep, progress, false);
        using Hole = typename CTDIPhantom<T>::HolePosition;
        const std::array<Hole, 5> holes = { Hole::Center, Hole::West, Hole::East, Hole::South, Hole::North };
        std::array<T, 5> d {};
        for (std::size_t i = 0; i < 5; ++i) {
            const auto& idx = world.holeIndices(holes[i]);
            for (const auto k : idx)
                d[i] += result.dose[k];
            d[i] /= idx.size();
        }
        const T peripheral = (d[1] + d[2] + d[3] + d[4]) / T { 4 };
        ctdiw = (d[0] + 2 * peripheral) / 3;
        ctdiw *= T { 100 } / src.collimation();
        if (result.numberOfHistories > 0 && ctdiw > 0)
            CalibrationCache<T>::store(key, ctdiw);
    }
    return src.ctdiVol() / ctdiw / meanWeight;
}
}
