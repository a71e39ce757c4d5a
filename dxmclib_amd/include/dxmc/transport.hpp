// Transport: the reference's CPU transport call (transport.hpp:114-840) with
// the same interface, running the photon histories on an MI355X through the
// C ABI of include/dxmc_hip.h.
//
//   auto result = Transport<double>()(world, &source);   // plain FP64 method
//   auto result = Transport<float>()(world, &source);    // FP32 persistent path
//
// operator() validates the world and source, builds the attenuation tables
// and the exposure table on the host, uploads them, transports on the device,
// normalises the FP64 tallies on the device (eV/history or dose) and returns
// a Result<T>. Invalid inputs and cancellation return the reference's empty
// result (all zero, numberOfHistories = 0); device errors are reported
// through lastError() and also return an empty result.
#pragma once

#include "dxmc/attenuationlut.hpp"
#include "dxmc/devicetables.hpp"
#include "dxmc/dxmcrandom.hpp"
#include "dxmc/exposure.hpp"
#include "dxmc/floating.hpp"
#include "dxmc/lowenergycorrectionmodel.hpp"
#include "dxmc/particle.hpp"
#include "dxmc/progressbar.hpp"
#include "dxmc/world.hpp"
#include "dxmc_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <string_view>
#include <thread>
#include <type_traits>
#include <vector>

namespace dxmc {

template <Floating T = double>
struct Result {
    std::vector<T> dose;
    std::vector<std::uint32_t> nEvents;
    std::vector<T> variance;
    std::uint64_t numberOfHistories { 0 };
    std::chrono::duration<float> simulationTime { 0 };
    std::string_view dose_units = "";

    Result() { }
    Result(std::size_t size)
        : dose(size, 0)
        , nEvents(size, 0)
        , variance(size, 0)
    {
    }
    std::vector<T> relativeError() const
    {
        std::vector<T> e(dose.size());
        for (std::size_t i = 0; i < dose.size(); ++i)
            e[i] = dose[i] > 0 ? std::sqrt(variance[i]) / dose[i] : 0;
        return e;
    }
};

template <Floating T>
class Source;

// One HIP context per device, shared by all Transport objects of the process.
class DeviceContexts {
public:
    // the device's context, or nullptr with the reason in 'why'
    static dxmc_ctx* get(int device, std::string& why)
    {
        static std::mutex mutex;
        static std::map<int, std::unique_ptr<dxmc_ctx, int (*)(dxmc_ctx*)>> contexts;
        std::lock_guard lock(mutex);
        auto it = contexts.find(device);
        if (it != contexts.end())
            return it->second.get();
        dxmc_ctx* c = nullptr;
        if (dxmc_hip_ctx_create(device, &c) != DXMC_OK) {
            why = std::string("cannot create a HIP context: ") + dxmc_hip_last_error(nullptr);
            return nullptr;
        }
        contexts.emplace(device, std::unique_ptr<dxmc_ctx, int (*)(dxmc_ctx*)>(c, dxmc_hip_ctx_destroy));
        return c;
    }
};

template <Floating T = double>
class Transport {
public:
    enum class OUTPUTMODE {
        EV_PER_HISTORY,
        DOSE
    };

    Transport() = default;

    // kept for interface compatibility: the GPU schedules histories itself
    void setNumberOfWorkers(std::uint64_t n) { m_nThreads = std::max<std::uint64_t>(n, 1); }
    std::size_t numberOfWorkers() const { return m_nThreads; }
    void setLowEnergyCorrectionModel(LOWENERGYCORRECTION m) { m_lowenergyCorrection = m; }
    LOWENERGYCORRECTION lowEnergyCorrectionModel() const { return m_lowenergyCorrection; }
    void setOutputMode(OUTPUTMODE m) { m_outputmode = m; }
    OUTPUTMODE outputMode() const { return m_outputmode; }

    // device-specific settings
    void setDevice(int d) { m_device = d; }
    int device() const { return m_device; }
    void setSeed(std::uint64_t s) { m_seed = s; }
    std::uint64_t seed() const { return m_seed; }
    // DXMC_PRECISION_FP64 (default for double) or DXMC_PRECISION_FP32 (default for float)
    void setDevicePrecision(int p) { m_precision = p; }
    int devicePrecision() const { return m_precision; }
    void setRunConfiguration(const dxmc_run_cfg& cfg) { m_tuning = cfg; }
    const dxmc_run_stats& lastRunStats() const { return m_stats; }
    // wall time of the device phases of the last run [s]: world, tables and
    // tally to the device; transport and normalisation; tally to the host;
    // release of the device objects
    struct DeviceTiming {
        double upload = 0, transport = 0, download = 0, release = 0;
    };
    const DeviceTiming& lastDeviceTiming() const { return m_timing; }
    const std::string& lastError() const { return m_error; }

    // The host part of operator(): attenuation tables for the world's
    // materials and the source's exposure table aligned to the world basis.
    // No pass over the voxels: the Woodcock majorant needs each material's
    // maximum density, which the device computes when execute() uploads the
    // world; the majorant is then fitted on the host from those values.
    struct Prepared {
        bool valid = false;
        AttenuationLut<double> lut; // all tables but the majorant
        TableSet tables;
        std::vector<dxmc_exposure> exposures;
        std::uint64_t histories = 0;
    };

    template <typename U>
        requires std::is_base_of<World<T>, U>::value
    Prepared prepare(const U& world, Source<T>* source)
    {
        Prepared p;
        if (!world.isValid() || !source)
            return p;
        source->updateFromWorld(world);
        source->validate();
        if (!source->isValid())
            return p;
        p.histories = source->historiesPerExposure() * source->totalExposures();
        // host tables (always FP64; the device derives its FP32 copy)
        p.lut.generateWorldTables(world.materialMap(), static_cast<double>(source->maxPhotonEnergyProduced()), 1.0);
        p.tables.setMaterialTables(p.lut);
        ExposureTable<T> exposures(p.tables);
        const auto& basis = world.directionCosines();
        const std::uint64_t nExp = source->totalExposures();
        for (std::uint64_t i = 0; i < nExp; ++i) {
            auto e = source->getExposure(i);
            e.alignToDirectionCosines(basis);
            exposures.add(e);
        }
        p.exposures = exposures.records();
        p.valid = true;
        return p;
    }

    // The device part of operator() on a prepared run. 'context' selects the
    // HIP context (and so the stream) to run on; nullptr uses the device's
    // shared context.
    template <typename U>
        requires std::is_base_of<World<T>, U>::value
    Result<T> execute(const U& world, Source<T>* source, const Prepared& prep, ProgressBar<T>* progressbar = nullptr,
        bool useSourceDoseCalibration = true, dxmc_ctx* context = nullptr)
    {
        Result<T> result(world.size());
        m_error.clear();
        if (!prep.valid)
            return result;
        if (progressbar)
            progressbar->setDoseData(result.dose.data(), world.dimensions(), world.spacing());
        const auto start = std::chrono::steady_clock::now();
        // FP64 tallies go straight into the result when T is double
        std::vector<double> dose, variance;
        double* doseOut = nullptr;
        double* varianceOut = nullptr;
        if constexpr (std::is_same_v<T, double>) {
            doseOut = result.dose.data();
            varianceOut = result.variance.data();
        } else {
            dose.resize(world.size());
            variance.resize(world.size());
            doseOut = dose.data();
            varianceOut = variance.data();
        }
        result.numberOfHistories = executeInto(world, source, prep, doseOut, result.nEvents.data(), varianceOut, progressbar, useSourceDoseCalibration, context);
        result.simulationTime = std::chrono::steady_clock::now() - start;
        if (progressbar)
            progressbar->clearDoseData();
        if (result.numberOfHistories == 0) {
            std::fill(result.dose.begin(), result.dose.end(), T { 0 });
            std::fill(result.variance.begin(), result.variance.end(), T { 0 });
            std::fill(result.nEvents.begin(), result.nEvents.end(), 0);
            return result;
        }
        if constexpr (!std::is_same_v<T, double>) {
            for (std::size_t i = 0; i < dose.size(); ++i) {
                result.dose[i] = static_cast<T>(dose[i]);
                result.variance[i] = static_cast<T>(variance[i]);
            }
        }
        if (m_outputmode == OUTPUTMODE::DOSE)
            result.dose_units = useSourceDoseCalibration ? "mGy" : "keV/kg";
        else
            result.dose_units = "eV/history";
        return result;
    }

    // execute() into caller-owned host arrays of world.size() elements: FP64
    // dose (or energy imparted) and variance, and event counts. Returns the
    // number of histories, or 0 on failure or cancel (the arrays are then
    // unspecified). Lets a caller that runs many requests reuse its arrays.
    template <typename U>
        requires std::is_base_of<World<T>, U>::value
    std::uint64_t executeInto(const U& world, Source<T>* source, const Prepared& prep, double* dose, std::uint32_t* nEvents, double* variance,
        ProgressBar<T>* progressbar = nullptr, bool useSourceDoseCalibration = true, dxmc_ctx* context = nullptr)
    {
        m_error.clear();
        if (!prep.valid || !dose || !nEvents || !variance)
            return 0;
        if (progressbar)
            progressbar->setTotalExposures(prep.exposures.size());
        // the calibration factor (a nested transport for CT sources) is applied
        // by the device normalisation: dose = ei * (calibration / voxel mass)
        auto calibration = [&]() -> double {
            if (m_outputmode == OUTPUTMODE::DOSE && useSourceDoseCalibration)
                return static_cast<double>(source->getCalibrationValue(m_lowenergyCorrection, progressbar));
            return 1.0;
        };
        const bool ok = runOnDevice(world, prep, progressbar, calibration, dose, nEvents, variance, prep.histories, context);
        if (!ok || (progressbar && progressbar->cancel()))
            return 0;
        return prep.histories;
    }

    // Transport<T>::operator() of the reference (transport.hpp:135-202)
    template <typename U>
        requires std::is_base_of<World<T>, U>::value
    Result<T> operator()(const U& world, Source<T>* source, ProgressBar<T>* progressbar = nullptr, bool useSourceDoseCalibration = true)
    {
        m_error.clear();
        const Prepared prep = prepare(world, source);
        if (!prep.valid)
            return Result<T>(world.size());
        return execute(world, source, prep, progressbar, useSourceDoseCalibration);
    }

    const AttenuationLut<double>& attenuationLut() const { return m_attenuationLut; }

protected:
    template <typename Calibration>
    bool runOnDevice(const World<T>& world, const Prepared& prep, ProgressBar<T>* pb, Calibration calibration, double* dose, std::uint32_t* nEvents,
        double* variance, std::uint64_t nHist, dxmc_ctx* context)
    {
        dxmc_ctx* c = context ? context : DeviceContexts::get(m_device, m_error);
        if (!c)
            return false;
        dxmc_world_desc wd {};
        for (int i = 0; i < 3; ++i) {
            wd.dims[i] = world.dimensions()[i];
            wd.spacing[i] = static_cast<double>(world.spacing()[i]);
        }
        for (int i = 0; i < 6; ++i)
            wd.extent_safe[i] = static_cast<double>(world.matrixExtentSafe()[i]);
        wd.material = world.materialIndexArray()->data();
        if constexpr (std::is_same_v<T, double>)
            wd.density_f64 = world.densityArray()->data();
        else
            wd.density_f32 = world.densityArray()->data();
        wd.measurement = world.measurementMapArray() ? world.measurementMapArray()->data() : nullptr;
        wd.n_materials = static_cast<std::uint32_t>(world.materialMap().size());
        dxmc_world* dw = nullptr;
        dxmc_tables* dt = nullptr;
        dxmc_tally* tally = nullptr;
        auto cleanup = [&]() {
            if (tally)
                dxmc_hip_tally_destroy(c, tally);
            if (dt)
                dxmc_hip_tables_destroy(c, dt);
            if (dw)
                dxmc_hip_world_destroy(c, dw);
        };
        dxmc_run_cfg cfg = m_tuning;
        cfg.precision = m_precision >= 0 ? m_precision : (std::is_same_v<T, float> ? DXMC_PRECISION_FP32 : DXMC_PRECISION_FP64);
        cfg.low_energy_correction = static_cast<std::int32_t>(m_lowenergyCorrection);
        cfg.seed = m_seed;
        const int mode = m_outputmode == OUTPUTMODE::DOSE ? DXMC_OUTPUT_DOSE : DXMC_OUTPUT_EV_PER_HISTORY;
        // the majorant from the per-material maximum densities of the device scan
        TableSet tables;
        std::vector<double> maxDensity(wd.n_materials);
        using clock = std::chrono::steady_clock;
        const auto seconds = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const auto t0 = clock::now();
        auto t1 = t0, t2 = t0;
        bool ok = !(dxmc_hip_world_upload(c, &wd, &dw) || dxmc_hip_world_max_density(dw, maxDensity.data(), wd.n_materials));
        if (ok) {
            m_attenuationLut = prep.lut;
            m_attenuationLut.setMaxDensity(maxDensity);
            tables = prep.tables;
            tables.majorant = m_attenuationLut.interpolator().maxCoefficients();
            const dxmc_tables_desc td = tables.desc();
            ok = !(dxmc_hip_tables_upload(c, &td, &dt) || dxmc_hip_tally_create(c, world.size(), &tally));
            t1 = clock::now();
            // progress, cancel and the live dose preview go through the
            // progress bar, as the reference's workers report to it
            dxmc_run_monitor monitor {};
            if (pb) {
                monitor.cancel = pb->cancelFlag();
                monitor.exposures_done = pb->exposureCounter();
                monitor.preview = &ProgressBar<T>::previewCallback;
                monitor.preview_user = pb;
                monitor.preview_axis = pb->planeNormal() == ProgressBar<T>::Axis::X ? 0 : pb->planeNormal() == ProgressBar<T>::Axis::Y ? 1 : 2;
                monitor.preview_interval_seconds = static_cast<float>(pb->previewInterval());
            }
            ok = ok && !dxmc_hip_run_monitored(c, dw, dt, prep.exposures.data(), prep.exposures.size(), &cfg, tally, &monitor, &m_stats);
        }
        if (ok && !(pb && pb->cancel())) {
            // dose does not depend on the history count (transport.hpp:
            // 796-816); eV/history needs two or more histories (780-794)
            const bool normalize = mode == DXMC_OUTPUT_DOSE || nHist > 1;
            ok = !(normalize && dxmc_hip_tally_normalize(c, tally, dw, mode, nHist, mode == DXMC_OUTPUT_DOSE ? calibration() : 1.0))
                && !dxmc_hip_synchronize(c);
            t2 = clock::now();
            ok = ok && !dxmc_hip_tally_download(c, tally, dose, nEvents, variance);
        }
        const auto t3 = clock::now();
        if (!ok)
            m_error = dxmc_hip_last_error(c);
        cleanup();
        m_timing = { seconds(t0, t1), seconds(t1, t2), seconds(t2, t3), seconds(t3, clock::now()) };
        return ok;
    }

private:
    AttenuationLut<double> m_attenuationLut;
    std::uint64_t m_nThreads = 1;
    OUTPUTMODE m_outputmode = OUTPUTMODE::DOSE;
    LOWENERGYCORRECTION m_lowenergyCorrection = LOWENERGYCORRECTION::LIVERMORE;
    int m_device = 0;
    std::uint64_t m_seed = 42;
    int m_precision = -1;
    dxmc_run_cfg m_tuning {};
    dxmc_run_stats m_stats {};
    DeviceTiming m_timing {};
    std::string m_error;
};
}
