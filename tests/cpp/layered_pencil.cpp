// Drop-in check of the C++ host API: a program written the way a user of
// the reference's Transport writes one (World, Material, PencilSource,
// Transport<T>::operator()), built against dxmclib_amd/include and
// libdxmc_host.so. Three slabs along z (air, water, aluminium), 60 keV pencil
// beam through a voxel centre, eV/history output, per-history streams from
// seed 42.
//
// usage: layered_pencil <precision 0|1> <histories per exposure> <exposures>
// prints one JSON line: histories, steps, events, energy per slab, seconds
#include "dxmc.hpp"

#include <array>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <numeric>
#include <vector>

using namespace dxmc;

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s precision histories_per_exposure exposures\n", argv[0]);
        return 2;
    }
    const int precision = std::atoi(argv[1]);
    const std::uint64_t perExposure = std::strtoull(argv[2], nullptr, 10);
    const std::uint64_t exposures = std::strtoull(argv[3], nullptr, 10);

    const std::array<std::size_t, 3> dims = { 48, 48, 48 };
    World<double> world;
    world.setDimensions(dims);
    world.setSpacing({ 1.0, 1.0, 1.0 });
    const Material air("Air, Dry (near sea level)"), water("Water, Liquid"), aluminium(13);
    world.addMaterialToMap(air);
    world.addMaterialToMap(water);
    world.addMaterialToMap(aluminium);
    auto density = std::make_shared<std::vector<double>>(world.size());
    auto material = std::make_shared<std::vector<std::uint8_t>>(world.size());
    const std::size_t slice = dims[0] * dims[1];
    for (std::size_t i = 0; i < world.size(); ++i) {
        const std::size_t z = i / slice;
        const std::uint8_t m = z < dims[2] / 3 ? 0 : z < 2 * dims[2] / 3 ? 1 : 2;
        (*material)[i] = m;
        (*density)[i] = (m == 0 ? air : m == 1 ? water : aluminium).standardDensity();
    }
    world.setDensityArray(density);
    world.setMaterialIndexArray(material);
    world.makeValid();
    if (!world.isValid()) {
        std::fprintf(stderr, "world invalid\n");
        return 1;
    }

    PencilSource<double> pencil;
    pencil.setPhotonEnergy(60.0);
    pencil.setPosition(0.5, 0.5, -48.0);
    pencil.setDirectionCosines({ 1, 0, 0, 0, 1, 0 });
    pencil.setTotalExposures(exposures);
    pencil.setHistoriesPerExposure(perExposure);

    Transport<double> transport;
    transport.setOutputMode(Transport<double>::OUTPUTMODE::EV_PER_HISTORY);
    transport.setLowEnergyCorrectionModel(LOWENERGYCORRECTION::LIVERMORE);
    transport.setSeed(42);
    transport.setDevicePrecision(precision);
    const auto result = transport(world, &pencil);
    if (result.numberOfHistories == 0) {
        std::fprintf(stderr, "transport failed: %s\n", transport.lastError().c_str());
        return 1;
    }
    std::array<double, 3> slab { 0, 0, 0 };
    for (std::size_t i = 0; i < world.size(); ++i)
        slab[(*material)[i]] += result.dose[i];
    const auto& s = transport.lastRunStats();
    std::printf("{\"histories\": %llu, \"steps\": %llu, \"events\": %llu, \"air\": %.17g, \"water\": %.17g, \"aluminium\": %.17g, "
                "\"seconds\": %.6f}\n",
        static_cast<unsigned long long>(result.numberOfHistories), static_cast<unsigned long long>(s.steps),
        static_cast<unsigned long long>(s.events), slab[0], slab[1], slab[2], static_cast<double>(result.simulationTime.count()));
    return 0;
}
