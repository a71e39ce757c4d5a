// The reference's ProgressBar as a caller uses it while a transport runs on
// another thread (reference validation/validation.cpp:270-285 polls it so):
// exposure progress, and the dose preview computeDoseProgressImage(), which
// here is the device's projection of the tally so far. A DX source on a
// water block with a bone slab; launches of 20000 histories give many
// previews. Prints one JSON line: images seen, those from device previews
// during the run with their size and largest pixel, the exposures counted,
// the final number of histories.
//
// usage: progress_preview <axis 0|1|2>
#include "dxmc.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

using namespace dxmc;

int main(int argc, char** argv)
{
    const int axis = argc > 1 ? std::atoi(argv[1]) : 2;
    const std::array<std::size_t, 3> dims = { 40, 36, 32 }; // x, y, z
    World<double> world;
    world.setDimensions(dims);
    world.setSpacing({ 1.0, 1.0, 1.0 });
    const Material water("Water, Liquid"), bone("Bone, Cortical (ICRU)");
    world.addMaterialToMap(water);
    world.addMaterialToMap(bone);
    auto density = std::make_shared<std::vector<double>>(world.size());
    auto material = std::make_shared<std::vector<std::uint8_t>>(world.size());
    for (std::size_t i = 0; i < world.size(); ++i) {
        const std::size_t x = i % dims[0];
        const std::uint8_t m = (x >= 18 && x < 22) ? 1 : 0;
        (*material)[i] = m;
        (*density)[i] = m ? bone.standardDensity() : water.standardDensity();
    }
    world.setDensityArray(density);
    world.setMaterialIndexArray(material);
    world.makeValid();

    DXSource<double> src;
    src.tube().setVoltage(90.0);
    src.setFieldSize({ 30.0, 30.0 });
    src.setSourceDetectorDistance(1000.0);
    src.setTotalExposures(50);
    src.setHistoriesPerExposure(20000);

    Transport<double> transport;
    transport.setOutputMode(Transport<double>::OUTPUTMODE::EV_PER_HISTORY);
    transport.setSeed(7);
    transport.setDevicePrecision(DXMC_PRECISION_FP32);
    dxmc_run_cfg cfg {};
    cfg.histories_per_launch = 20000;
    transport.setRunConfiguration(cfg);

    ProgressBar<double> progress;
    progress.setPlaneNormal(axis == 0 ? ProgressBar<double>::Axis::X : axis == 1 ? ProgressBar<double>::Axis::Y : ProgressBar<double>::Axis::Z);
    progress.setPreviewInterval(0.0);

    std::atomic<bool> done { false };
    Result<double> result;
    std::thread worker([&] {
        result = transport(world, &src, &progress, false);
        done = true;
    });
    // An image with dose in it, taken while exposures are still running,
    // comes from a device preview: the host dose array is written only
    // after the last launch.
    std::size_t images = 0, midRun = 0, width = 0, height = 0;
    unsigned maxPixel = 0;
    std::uint64_t lastExposures = 0;
    bool monotone = true;
    while (!done) {
        const std::uint64_t e = progress.completedExposures();
        const std::uint64_t total = progress.totalExposures();
        monotone = monotone && e >= lastExposures;
        lastExposures = e;
        if (auto img = progress.computeDoseProgressImage()) {
            ++images;
            unsigned mx = 0;
            for (auto v : img->image)
                mx = std::max<unsigned>(mx, v);
            if (mx > 0 && total > 0 && e < total) {
                ++midRun;
                width = img->dimensions[0];
                height = img->dimensions[1];
                maxPixel = std::max(maxPixel, mx);
            }
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    worker.join();
    if (result.numberOfHistories == 0) {
        std::fprintf(stderr, "transport failed: %s\n", transport.lastError().c_str());
        return 1;
    }
    std::printf("{\"images\": %zu, \"mid_run\": %zu, \"width\": %zu, \"height\": %zu, \"max_pixel\": %u, \"exposures\": %llu, \"total_exposures\": %llu, "
                "\"monotone\": %s, \"histories\": %llu}\n",
        images, midRun, width, height, maxPixel, static_cast<unsigned long long>(progress.completedExposures()),
        static_cast<unsigned long long>(progress.totalExposures()), monotone ? "true" : "false",
        static_cast<unsigned long long>(result.numberOfHistories));
    return 0;
}
