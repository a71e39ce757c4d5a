// Two behaviours of the C++ host API with more than one host thread:
//
// cancel      A transport runs on a worker thread; once the ProgressBar counts
//             a finished exposure, the main thread calls setCancel(true). The
//             result is the reference's cancelled Result (progressbar.hpp:
//             77-84, transport.hpp:176-185): numberOfHistories 0 and every
//             dose, variance and event count 0.
// concurrent  Two Transport objects on two threads at once (the device's
//             shared context), then the same two calls one after the other.
//             Prints the step/event counts and totals of each, which must
//             agree.
//
// usage: threads_cancel cancel | concurrent; prints one JSON line
#include "dxmc.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

using namespace dxmc;

namespace {

World<double> waterBox(std::size_t n, double boneFraction)
{
    World<double> world;
    world.setDimensions({ n, n, n });
    world.setSpacing({ 1.0, 1.0, 1.0 });
    const Material water("Water, Liquid"), bone("Bone, Cortical (ICRU)");
    world.addMaterialToMap(water);
    world.addMaterialToMap(bone);
    auto density = std::make_shared<std::vector<double>>(world.size());
    auto material = std::make_shared<std::vector<std::uint8_t>>(world.size());
    for (std::size_t i = 0; i < world.size(); ++i) {
        const std::uint8_t m = (i % n) < static_cast<std::size_t>(boneFraction * n) ? 1 : 0;
        (*material)[i] = m;
        (*density)[i] = m ? bone.standardDensity() : water.standardDensity();
    }
    world.setDensityArray(density);
    world.setMaterialIndexArray(material);
    world.makeValid();
    return world;
}

struct Summary {
    std::uint64_t histories = 0, steps = 0, events = 0;
    double dose = 0, variance = 0;
    std::uint64_t nEvents = 0;
    bool zero = true;
};

Summary summarise(const Result<double>& r, const Transport<double>& t)
{
    Summary s;
    s.histories = r.numberOfHistories;
    s.steps = t.lastRunStats().steps;
    s.events = t.lastRunStats().events;
    for (std::size_t i = 0; i < r.dose.size(); ++i) {
        s.dose += r.dose[i];
        s.variance += r.variance[i];
        s.nEvents += r.nEvents[i];
        s.zero = s.zero && r.dose[i] == 0 && r.variance[i] == 0 && r.nEvents[i] == 0;
    }
    return s;
}

void print(const char* name, const Summary& s, bool last)
{
    std::printf("\"%s\": {\"histories\": %llu, \"steps\": %llu, \"events\": %llu, \"dose\": %.17g, \"variance\": %.17g, \"n_events\": %llu, "
                "\"zero\": %s}%s",
        name, static_cast<unsigned long long>(s.histories), static_cast<unsigned long long>(s.steps), static_cast<unsigned long long>(s.events),
        s.dose, s.variance, static_cast<unsigned long long>(s.nEvents), s.zero ? "true" : "false", last ? "" : ", ");
}

Transport<double> makeTransport(std::uint64_t seed, std::uint64_t perLaunch)
{
    Transport<double> t;
    t.setOutputMode(Transport<double>::OUTPUTMODE::EV_PER_HISTORY);
    t.setSeed(seed);
    t.setDevicePrecision(DXMC_PRECISION_FP32);
    dxmc_run_cfg cfg {};
    cfg.flags = 1; // device step / event counters
    cfg.histories_per_launch = perLaunch;
    t.setRunConfiguration(cfg);
    return t;
}

int cancelRun()
{
    auto world = waterBox(40, 0.2);
    DXSource<double> src;
    src.tube().setVoltage(100.0);
    src.setFieldSize({ 30.0, 30.0 });
    src.setSourceDetectorDistance(1000.0);
    src.setTotalExposures(400);
    src.setHistoriesPerExposure(20000);
    auto transport = makeTransport(11, 20000);
    ProgressBar<double> progress;
    std::atomic<bool> done { false };
    Result<double> result;
    std::thread worker([&] {
        result = transport(world, &src, &progress, false);
        done = true;
    });
    std::uint64_t seenAtCancel = 0;
    while (!done) {
        const std::uint64_t e = progress.completedExposures();
        if (e >= 1) {
            seenAtCancel = e;
            progress.setCancel(true);
            break;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    worker.join();
    const Summary s = summarise(result, transport);
    std::printf("{\"exposures_at_cancel\": %llu, \"total_exposures\": %llu, \"completed\": %llu, \"size\": %zu, ",
        static_cast<unsigned long long>(seenAtCancel), static_cast<unsigned long long>(progress.totalExposures()),
        static_cast<unsigned long long>(progress.completedExposures()), result.dose.size());
    print("result", s, true);
    std::printf("}\n");
    return 0;
}

int concurrent()
{
    auto worldA = waterBox(48, 0.3);
    auto worldB = waterBox(36, 0.5);
    auto source = [](double kv, std::uint64_t n) {
        auto s = std::make_unique<DXSource<double>>();
        s->tube().setVoltage(kv);
        s->setFieldSize({ 40.0, 40.0 });
        s->setSourceDetectorDistance(1000.0);
        s->setTotalExposures(20);
        s->setHistoriesPerExposure(n);
        return s;
    };
    auto srcA = source(80.0, 200000);
    auto srcB = source(120.0, 150000);
    // small launches so that the two runs' launches interleave
    auto tA = makeTransport(3, 1 << 18), tB = makeTransport(5, 1 << 18);
    Result<double> rA, rB;
    std::thread a([&] { rA = tA(worldA, srcA.get(), nullptr, false); });
    std::thread b([&] { rB = tB(worldB, srcB.get(), nullptr, false); });
    a.join();
    b.join();
    const Summary cA = summarise(rA, tA), cB = summarise(rB, tB);
    auto sA = makeTransport(3, 1 << 18), sB = makeTransport(5, 1 << 18);
    const auto qA = sA(worldA, srcA.get(), nullptr, false);
    const Summary oA = summarise(qA, sA);
    const auto qB = sB(worldB, srcB.get(), nullptr, false);
    const Summary oB = summarise(qB, sB);
    // largest relative per-voxel difference of the dose (FP64 atomics sum in
    // a different order in different runs)
    auto maxRel = [](const Result<double>& x, const Result<double>& y) {
        double m = 0;
        for (std::size_t i = 0; i < x.dose.size(); ++i) {
            const double d = std::abs(x.dose[i] - y.dose[i]);
            const double s = std::max(std::abs(x.dose[i]), std::abs(y.dose[i]));
            if (s > 0)
                m = std::max(m, d / s);
            if (x.nEvents[i] != y.nEvents[i])
                return 1.0;
        }
        return m;
    };
    std::printf("{");
    print("concurrent_a", cA, false);
    print("concurrent_b", cB, false);
    print("serial_a", oA, false);
    print("serial_b", oB, false);
    std::printf("\"max_rel_a\": %.3g, \"max_rel_b\": %.3g}\n", maxRel(rA, qA), maxRel(rB, qB));
    return 0;
}
}

int main(int argc, char** argv)
{
    if (argc == 2 && std::strcmp(argv[1], "cancel") == 0)
        return cancelRun();
    if (argc == 2 && std::strcmp(argv[1], "concurrent") == 0)
        return concurrent();
    std::fprintf(stderr, "usage: %s cancel | concurrent\n", argv[0]);
    return 2;
}
