// Implementation of include/dxmc_host.h on top of the C++ host API.
#include "dxmc_host.h"
#include "host_internal.hpp"

#include "dxmc/attenuationlut.hpp"
#include "dxmc/devicetables.hpp"
#include "dxmc/rawio.hpp"
#include "dxmc/source.hpp"
#include "dxmc/transport.hpp"
#include "dxmc/world.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>

using namespace dxmc;

struct dxmc_host_tables {
    AttenuationLut<double> lut;
    TableSet set;
    // sources whose spectrum/heel/bowtie tables were appended must outlive nothing:
    // TableSet copies the data
};

namespace {
std::mutex gMutex;
std::string gError;


bool allZero(const double* a, int n)
{
    for (int i = 0; i < n; ++i)
        if (a[i] != 0)
            return false;
    return true;
}

void setupTube(Tube<double>& tube, const dxmc_source_spec& s)
{
    if (s.tube_voltage > 0)
        tube.setVoltage(s.tube_voltage);
    if (s.anode_angle_deg > 0)
        tube.setAnodeAngleDeg(s.anode_angle_deg);
    if (s.energy_resolution > 0)
        tube.setEnergyResolution(s.energy_resolution);
    if (s.al_mm < 0 || s.cu_mm < 0 || s.sn_mm < 0)
        tube.clearFiltrationMaterials();
    if (s.al_mm > 0)
        tube.setAlFiltration(s.al_mm);
    if (s.cu_mm > 0)
        tube.setCuFiltration(s.cu_mm);
    if (s.sn_mm > 0)
        tube.setSnFiltration(s.sn_mm);
}

template <typename S>
void setupCommon(S& src, const dxmc_source_spec& s)
{
    src.setPosition({ s.position[0], s.position[1], s.position[2] });
    if (!allZero(s.direction_cosines, 6))
        src.setDirectionCosines({ s.direction_cosines[0], s.direction_cosines[1], s.direction_cosines[2], s.direction_cosines[3], s.direction_cosines[4],
            s.direction_cosines[5] });
    if (s.histories_per_exposure > 0)
        src.setHistoriesPerExposure(s.histories_per_exposure);
}

template <typename S>
void setupCT(S& src, const dxmc_source_spec& s)
{
    setupTube(src.tube(), s);
    if (s.sdd > 0)
        src.setSourceDetectorDistance(s.sdd);
    if (s.collimation > 0)
        src.setCollimation(s.collimation);
    if (s.fov > 0)
        src.setFieldOfView(s.fov);
    src.setStartAngleDeg(s.start_angle_deg);
    src.setGantryTiltAngleDeg(s.gantry_tilt_deg);
    if constexpr (requires { src.setExposureAngleStepDeg(1.0); }) // not the topogram
        if (s.exposure_angle_step_deg > 0)
            src.setExposureAngleStepDeg(s.exposure_angle_step_deg);
    if (s.ctdi_vol > 0)
        src.setCtdiVol(s.ctdi_vol);
    src.setModelHeelEffect(s.heel_effect != 0);
    if (s.bowtie_n > 0) {
        std::vector<double> a(s.bowtie_angles, s.bowtie_angles + s.bowtie_n), w(s.bowtie_weights, s.bowtie_weights + s.bowtie_n);
        src.setBowTieFilter(std::make_shared<BowTieFilter<double>>(a, w));
    }
}

// XCare settings of CT axial, spiral and dual sources
template <typename S>
void setupXCare(S& src, const dxmc_source_spec& s)
{
    src.setUseXCareFilter(s.xcare != 0);
    auto& f = src.xcareFilter();
    if (s.xcare_filter_angle_deg != 0)
        f.setFilterAngleDeg(s.xcare_filter_angle_deg);
    if (s.xcare_span_angle_deg > 0)
        f.setSpanAngleDeg(s.xcare_span_angle_deg);
    if (s.xcare_ramp_angle_deg > 0)
        f.setRampAngleDeg(s.xcare_ramp_angle_deg);
    if (s.xcare_low_weight > 0)
        f.setLowWeight(s.xcare_low_weight);
}

// tube B and the tube currents of a dual-source CT
template <typename S>
void setupDual(S& src, const dxmc_source_spec& s)
{
    dxmc_source_spec b = s;
    b.tube_voltage = s.tube_b_voltage;
    b.anode_angle_deg = s.tube_b_anode_angle_deg;
    b.al_mm = s.tube_b_al_mm;
    b.cu_mm = s.tube_b_cu_mm;
    b.sn_mm = s.tube_b_sn_mm;
    setupTube(src.tubeB(), b);
    if (s.sdd_b > 0)
        src.setSourceDetectorDistanceB(s.sdd_b);
    if (s.fov_b > 0)
        src.setFieldOfViewB(s.fov_b);
    src.setStartAngleB(s.start_angle_b_set ? s.start_angle_b_deg * DEG_TO_RAD<double>() : src.startAngle() + PI_VAL<double>() / 2);
    if (s.mas_a > 0)
        src.setTubeAmas(s.mas_a);
    if (s.mas_b > 0)
        src.setTubeBmas(s.mas_b);
    if (s.bowtie_b_n > 0) {
        std::vector<double> a(s.bowtie_b_angles, s.bowtie_b_angles + s.bowtie_b_n), w(s.bowtie_b_weights, s.bowtie_b_weights + s.bowtie_b_n);
        src.setBowTieFilterB(std::make_shared<BowTieFilter<double>>(a, w));
    }
}

std::array<double, 6> cosinesOrIdentity(const double* c)
{
    if (!c || allZero(c, 6))
        return { 1, 0, 0, 0, 1, 0 };
    return { c[0], c[1], c[2], c[3], c[4], c[5] };
}

}

namespace dxmc_host_detail {

int fail(int status, const std::string& msg)
{
    std::lock_guard lock(gMutex);
    gError = msg;
    return status;
}

std::unique_ptr<Source<double>> makeSource(const dxmc_source_spec& s)
{
    switch (s.type) {
    case DXMC_SOURCE_PENCIL: {
        auto p = std::make_unique<PencilSource<double>>();
        setupCommon(*p, s);
        if (s.photon_energy > 0)
            p->setPhotonEnergy(s.photon_energy);
        if (s.total_exposures > 0)
            p->setTotalExposures(s.total_exposures);
        if (s.air_dose > 0)
            p->setAirDose(s.air_dose);
        return p;
    }
    case DXMC_SOURCE_ISOTROPIC:
    case DXMC_SOURCE_ISOTROPIC_CT: {
        std::unique_ptr<IsotropicSource<double>> p;
        if (s.type == DXMC_SOURCE_ISOTROPIC)
            p = std::make_unique<IsotropicSource<double>>();
        else
            p = std::make_unique<IsotropicCTSource<double>>();
        setupCommon(*p, s);
        if (s.total_exposures > 0)
            p->setTotalExposures(s.total_exposures);
        p->setCollimationAngles(s.collimation_angles[0], s.collimation_angles[1], s.collimation_angles[2], s.collimation_angles[3]);
        if (s.spectrum_n > 0)
            p->setSpecter(std::vector<double>(s.spectrum_weights, s.spectrum_weights + s.spectrum_n),
                std::vector<double>(s.spectrum_energies, s.spectrum_energies + s.spectrum_n));
        return p;
    }
    case DXMC_SOURCE_DX: {
        auto p = std::make_unique<DXSource<double>>();
        setupTube(p->tube(), s);
        if (s.sdd > 0)
            p->setSourceDetectorDistance(s.sdd);
        if (s.field_size[0] > 0 && s.field_size[1] > 0)
            p->setFieldSize({ s.field_size[0], s.field_size[1] });
        if (s.source_angles_deg[0] != 0 || s.source_angles_deg[1] != 0)
            p->setSourceAnglesDeg(s.source_angles_deg[0], s.source_angles_deg[1]);
        setupCommon(*p, s);
        if (s.total_exposures > 0)
            p->setTotalExposures(s.total_exposures);
        if (s.dap > 0)
            p->setDap(s.dap);
        p->setModelHeelEffect(s.heel_effect != 0);
        return p;
    }
    case DXMC_SOURCE_CBCT: {
        auto p = std::make_unique<CBCTSource<double>>();
        setupTube(p->tube(), s);
        if (s.sdd > 0)
            p->setSourceDetectorDistance(s.sdd);
        if (s.field_size[0] > 0 && s.field_size[1] > 0)
            p->setFieldSize({ s.field_size[0], s.field_size[1] });
        setupCommon(*p, s);
        if (s.step_angle_deg > 0)
            p->setStepAngleDeg(s.step_angle_deg);
        if (s.span_angle_deg > 0)
            p->setSpanAngleDeg(s.span_angle_deg);
        if (s.dap > 0)
            p->setDap(s.dap);
        p->setModelHeelEffect(s.heel_effect != 0);
        return p;
    }
    case DXMC_SOURCE_CT_AXIAL: {
        auto p = std::make_unique<CTAxialSource<double>>();
        setupCommon(*p, s);
        setupCT(*p, s);
        setupXCare(*p, s);
        if (s.axial_step > 0)
            p->setStep(s.axial_step);
        if (s.scan_length > 0)
            p->setScanLenght(s.scan_length);
        return p;
    }
    case DXMC_SOURCE_CT_SPIRAL: {
        auto p = std::make_unique<CTSpiralSource<double>>();
        setupCommon(*p, s);
        setupCT(*p, s);
        setupXCare(*p, s);
        if (s.pitch > 0)
            p->setPitch(s.pitch);
        if (s.scan_length > 0)
            p->setScanLenght(s.scan_length);
        return p;
    }
    case DXMC_SOURCE_CT_AXIAL_DUAL: {
        auto p = std::make_unique<CTAxialDualSource<double>>();
        setupCommon(*p, s);
        setupCT(*p, s);
        setupXCare(*p, s);
        setupDual(*p, s);
        if (s.axial_step > 0)
            p->setStep(s.axial_step);
        if (s.scan_length > 0)
            p->setScanLenght(s.scan_length);
        return p;
    }
    case DXMC_SOURCE_CT_SPIRAL_DUAL: {
        auto p = std::make_unique<CTSpiralDualSource<double>>();
        setupCommon(*p, s);
        setupCT(*p, s);
        setupXCare(*p, s);
        setupDual(*p, s);
        if (s.pitch > 0)
            p->setPitch(s.pitch);
        if (s.scan_length > 0)
            p->setScanLenght(s.scan_length);
        return p;
    }
    case DXMC_SOURCE_CT_TOPOGRAM: {
        auto p = std::make_unique<CTTopogramSource<double>>();
        setupCommon(*p, s);
        setupCT(*p, s);
        if (s.scan_length > 0)
            p->setScanLenght(s.scan_length);
        return p;
    }
    default:
        return nullptr;
    }
}

// fn(b, e) over [0, n) in parts on up to 8 threads (parts of at least 2^21
// elements; a thread that cannot start has its part run on this one)
template <typename F>
void parallelFor(std::size_t n, F fn)
{
    constexpr std::size_t MIN_CHUNK = std::size_t { 1 } << 21;
    const unsigned cores = std::max(1u, std::thread::hardware_concurrency());
    const unsigned threads = static_cast<unsigned>(std::min<std::size_t>({ std::size_t { 8 }, cores, std::max<std::size_t>(1, n / MIN_CHUNK) }));
    if (threads <= 1) {
        fn(std::size_t { 0 }, n);
        return;
    }
    const std::size_t per = (n + threads - 1) / threads;
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    for (unsigned t = 1; t < threads; ++t) {
        const std::size_t b = std::min(n, t * per), e = std::min(n, b + per);
        if (b >= e)
            continue;
        try {
            pool.emplace_back([=] { fn(b, e); });
        } catch (...) {
            fn(b, e);
        }
    }
    fn(std::size_t { 0 }, std::min(n, per));
    for (auto& th : pool)
        th.join();
}

// dst[i] = src[i] over large arrays on several threads: one core streams
// about 9 GB/s, so the FP64 density copy of a 240^3 world took 19 ms alone
template <typename D, typename S>
void parallelCopy(D* dst, const S* src, std::size_t n)
{
    parallelFor(n, [=](std::size_t b, std::size_t e) { std::copy(src + b, src + e, dst + b); });
}

// v.resize(n) where the elements past v.size() are about to be overwritten:
// storage the vector has never used is first touched on several threads, so
// resize() zero-fills resident pages. A fresh 100 MB array otherwise takes
// its ~25000 page faults on one core (about 50 ms, against 10 ms for the
// parallel touch and 10 ms for the fill).
template <typename T>
void growForOverwrite(std::vector<T>& v, std::size_t n)
{
    static_assert(std::is_trivially_copyable_v<T>);
    if (v.size() >= n) {
        v.resize(n);
        return;
    }
    v.reserve(n);
    const std::size_t from = v.size();
    unsigned char* base = reinterpret_cast<unsigned char*>(v.data() + from);
    parallelFor((n - from) * sizeof(T), [=](std::size_t b, std::size_t e) { std::memset(base + b, 0, e - b); });
    v.resize(n);
}

bool buildWorld(const dxmc_host_world& w, World<double>& world, std::string& why, const WorldArrays* arrays)
{
    if (!w.material || !w.density || !w.material_names || w.n_materials == 0)
        return why = "world arrays or material names missing", false;
    world.setDimensions({ w.dims[0], w.dims[1], w.dims[2] });
    world.setSpacing({ w.spacing[0], w.spacing[1], w.spacing[2] });
    world.setOrigin({ w.origin[0], w.origin[1], w.origin[2] });
    world.setDirectionCosines(cosinesOrIdentity(w.direction_cosines));
    const std::size_t n = w.dims[0] * w.dims[1] * w.dims[2];
    auto dens = arrays ? arrays->density(n) : std::make_shared<std::vector<double>>();
    auto mat = arrays ? arrays->bytes(n) : std::make_shared<std::vector<std::uint8_t>>();
    growForOverwrite(*dens, n);
    parallelCopy(dens->data(), w.density, n);
    growForOverwrite(*mat, n);
    parallelCopy(mat->data(), w.material, n);
    world.setDensityArray(dens);
    world.setMaterialIndexArray(mat);
    if (w.measurement || arrays) {
        auto meas = arrays ? arrays->bytes(n) : std::make_shared<std::vector<std::uint8_t>>();
        growForOverwrite(*meas, n);
        if (w.measurement)
            parallelCopy(meas->data(), w.measurement, n);
        else
            std::memset(meas->data(), 0, n);
        world.setMeasurementMapArray(meas);
    }
    for (uint32_t i = 0; i < w.n_materials; ++i) {
        Material m(w.material_names[i]);
        if (!m.isValid())
            return why = std::string("unknown material ") + w.material_names[i], false;
        world.addMaterialToMap(m);
    }
    world.makeValid();
    if (!world.isValid())
        return why = "world failed validation", false;
    return true;
}

}

using dxmc_host_detail::buildWorld;
using dxmc_host_detail::fail;
using dxmc_host_detail::makeSource;

extern "C" {

const char* dxmc_host_last_error(void)
{
    std::lock_guard lock(gMutex);
    return gError.c_str();
}

int dxmc_host_tables_create(uint32_t n, const char* const* names, const double* maxDensity, double maxEnergy, double minEnergy, dxmc_host_tables** out)
{
    if (!out || !names || !maxDensity || n == 0 || n > 256)
        return fail(DXMC_ERR_ARGUMENT, "tables_create: bad arguments");
    *out = nullptr;
    std::vector<Material> mats;
    for (uint32_t i = 0; i < n; ++i) {
        Material m(names[i]);
        if (!m.isValid())
            return fail(DXMC_ERR_ARGUMENT, std::string("tables_create: unknown material ") + names[i]);
        mats.push_back(m);
    }
    auto t = std::make_unique<dxmc_host_tables>();
    t->lut.generate(mats, std::vector<double>(maxDensity, maxDensity + n), maxEnergy, minEnergy);
    t->set.setMaterialTables(t->lut);
    *out = t.release();
    return DXMC_OK;
}

int dxmc_host_tables_desc(const dxmc_host_tables* t, dxmc_tables_desc* out)
{
    if (!t || !out)
        return DXMC_ERR_ARGUMENT;
    *out = t->set.desc();
    return DXMC_OK;
}

void dxmc_host_tables_destroy(dxmc_host_tables* t) { delete t; }

int dxmc_host_source_exposures(const dxmc_source_spec* spec, const double worldCosines[6], dxmc_host_tables* tables, dxmc_exposure* out,
    uint64_t capacity, uint64_t* nOut)
{
    if (!spec || !tables || !nOut)
        return fail(DXMC_ERR_ARGUMENT, "source_exposures: bad arguments");
    auto src = makeSource(*spec);
    if (!src)
        return fail(DXMC_ERR_ARGUMENT, "source_exposures: unknown source type");
    src->validate();
    if (!src->isValid())
        return fail(DXMC_ERR_ARGUMENT, "source_exposures: source failed validation");
    const uint64_t n = src->totalExposures();
    *nOut = n;
    if (!out)
        return DXMC_OK;
    if (capacity < n)
        return fail(DXMC_ERR_ARGUMENT, "source_exposures: output too small");
    ExposureTable<double> table(tables->set);
    const auto basis = cosinesOrIdentity(worldCosines);
    for (uint64_t i = 0; i < n; ++i) {
        auto e = src->getExposure(i);
        e.alignToDirectionCosines(basis);
        table.add(e);
    }
    std::copy(table.records().cbegin(), table.records().cend(), out);
    return DXMC_OK;
}

int dxmc_host_source_spectrum(const dxmc_source_spec* spec, double* energies, double* weights, uint64_t capacity, uint64_t* nOut)
{
    if (!spec || !nOut)
        return DXMC_ERR_ARGUMENT;
    Tube<double> tube;
    if (spec->type == DXMC_SOURCE_CT_AXIAL || spec->type == DXMC_SOURCE_CT_SPIRAL)
        tube.setAlFiltration(7.0);
    else
        tube.setAlFiltration(2.0);
    setupTube(tube, *spec);
    const auto s = tube.getSpecter(true);
    *nOut = s.size();
    if (!energies || !weights)
        return DXMC_OK;
    if (capacity < s.size())
        return fail(DXMC_ERR_ARGUMENT, "source_spectrum: output too small");
    for (std::size_t i = 0; i < s.size(); ++i) {
        energies[i] = s[i].first;
        weights[i] = s[i].second;
    }
    return DXMC_OK;
}

int dxmc_host_read_raw(const char* path, void* data, uint64_t bytes)
{
    if (!path || (!data && bytes))
        return fail(DXMC_ERR_ARGUMENT, "read_raw: NULL argument");
    const auto a = readBinaryArray<unsigned char>(path, bytes);
    if (!a)
        return fail(DXMC_ERR_ARGUMENT, std::string("read_raw: ") + path + " is missing or does not hold exactly " + std::to_string(bytes) + " bytes");
    std::copy(a->cbegin(), a->cend(), static_cast<unsigned char*>(data));
    return DXMC_OK;
}

int dxmc_host_write_raw(const char* path, const void* data, uint64_t bytes)
{
    if (!path || (!data && bytes))
        return fail(DXMC_ERR_ARGUMENT, "write_raw: NULL argument");
    const auto* p = static_cast<const unsigned char*>(data);
    if (!saveBinaryArray(std::vector<unsigned char>(p, p + bytes), path))
        return fail(DXMC_ERR_ARGUMENT, std::string("write_raw: cannot write ") + path);
    return DXMC_OK;
}

int dxmc_host_world_extent(const dxmc_host_world* w, double extent[6])
{
    if (!w || !extent)
        return DXMC_ERR_ARGUMENT;
    World<double> world;
    world.setDimensions({ w->dims[0], w->dims[1], w->dims[2] });
    world.setSpacing({ w->spacing[0], w->spacing[1], w->spacing[2] });
    world.setOrigin({ w->origin[0], w->origin[1], w->origin[2] });
    for (int i = 0; i < 6; ++i)
        extent[i] = world.matrixExtentSafe()[i];
    return DXMC_OK;
}

int dxmc_host_transport(const dxmc_host_world* w, const dxmc_source_spec* spec, const dxmc_host_transport_cfg* cfg, double* dose, uint32_t* nEvents,
    double* variance, dxmc_run_stats* stats, double* seconds, uint64_t* nHistories)
{
    if (!w || !spec || !cfg)
        return fail(DXMC_ERR_ARGUMENT, "transport: NULL argument");
    World<double> world;
    std::string why;
    if (!buildWorld(*w, world, why))
        return fail(DXMC_ERR_ARGUMENT, "transport: " + why);
    auto src = makeSource(*spec);
    if (!src)
        return fail(DXMC_ERR_ARGUMENT, "transport: unknown source type");
    try {
        Transport<double> transport;
        transport.setDevice(cfg->device);
        transport.setDevicePrecision(cfg->precision);
        transport.setLowEnergyCorrectionModel(static_cast<LOWENERGYCORRECTION>(cfg->low_energy_correction));
        transport.setOutputMode(cfg->output_mode == DXMC_OUTPUT_EV_PER_HISTORY ? Transport<double>::OUTPUTMODE::EV_PER_HISTORY
                                                                                 : Transport<double>::OUTPUTMODE::DOSE);
        transport.setSeed(cfg->seed);
        transport.setRunConfiguration(cfg->tuning);
        const auto r = transport(world, src.get(), nullptr, cfg->use_source_calibration != 0);
        if (!transport.lastError().empty())
            return fail(DXMC_ERR_DEVICE, "transport: " + transport.lastError());
        if (r.numberOfHistories == 0)
            return fail(DXMC_ERR_ARGUMENT, "transport: invalid world or source");
        const std::size_t n = world.size();
        if (dose)
            std::copy(r.dose.cbegin(), r.dose.cend(), dose);
        if (nEvents)
            std::copy(r.nEvents.cbegin(), r.nEvents.cend(), nEvents);
        if (variance)
            std::copy(r.variance.cbegin(), r.variance.cend(), variance);
        if (stats)
            *stats = transport.lastRunStats();
        if (seconds)
            *seconds = r.simulationTime.count();
        if (nHistories)
            *nHistories = r.numberOfHistories;
        (void)n;
    } catch (const std::exception& e) {
        return fail(DXMC_ERR_DEVICE, std::string("transport: ") + e.what());
    }
    return DXMC_OK;
}

int dxmc_host_ctdi_phantom(uint64_t diameter, uint64_t dims[3], double spacing[3], uint8_t* material, double* density, uint8_t* measurement,
    uint8_t* hole, uint64_t capacity)
{
    if (!dims || !spacing || diameter == 0 || diameter > 4096)
        return fail(DXMC_ERR_ARGUMENT, "ctdi_phantom: bad arguments");
    const CTDIPhantom<double> w(diameter);
    const auto& d = w.dimensions();
    const auto& sp = w.spacing();
    for (int i = 0; i < 3; ++i) {
        dims[i] = d[i];
        spacing[i] = sp[i];
    }
    if (!material && !density && !measurement && !hole)
        return DXMC_OK;
    if (capacity < w.size())
        return fail(DXMC_ERR_ARGUMENT, "ctdi_phantom: capacity below the voxel count");
    if (material)
        std::copy(w.materialIndexArray()->cbegin(), w.materialIndexArray()->cend(), material);
    if (density)
        std::copy(w.densityArray()->cbegin(), w.densityArray()->cend(), density);
    if (measurement)
        std::copy(w.measurementMapArray()->cbegin(), w.measurementMapArray()->cend(), measurement);
    if (hole) {
        std::fill(hole, hole + w.size(), 0);
        using H = CTDIPhantom<double>::HolePosition;
        const H order[5] = { H::Center, H::West, H::East, H::South, H::North };
        for (int i = 0; i < 5; ++i)
            for (const auto k : w.holeIndices(order[i]))
                hole[k] = static_cast<uint8_t>(i + 1);
    }
    return DXMC_OK;
}

int dxmc_host_material_data(const char* name, double E, double out[5], double* density)
{
    if (!name || !out)
        return DXMC_ERR_ARGUMENT;
    Material m(name);
    if (!m.isValid())
        return fail(DXMC_ERR_ARGUMENT, std::string("material_data: unknown material ") + name);
    out[0] = m.getPhotoelectricAttenuation(E);
    out[1] = m.getComptonAttenuation(E);
    out[2] = m.getRayleightAttenuation(E);
    out[3] = m.getTotalAttenuation(E);
    out[4] = m.getMassEnergyAbsorbtion(E);
    if (density)
        *density = m.standardDensity();
    return DXMC_OK;
}

int dxmc_host_material_scatter_factors(const char* name, std::uint64_t n, const double* q, double* ff2, double* sf)
{
    if (!name || (n && !q))
        return DXMC_ERR_ARGUMENT;
    Material m(name);
    if (!m.isValid())
        return fail(DXMC_ERR_ARGUMENT, std::string("material_scatter_factors: unknown material ") + name);
    for (std::uint64_t i = 0; i < n; ++i) {
        if (ff2)
            ff2[i] = m.getRayleightFormFactorSquared(q[i]);
        if (sf)
            sf[i] = m.getComptonNormalizedScatterFactor(q[i]);
    }
    return DXMC_OK;
}
}
