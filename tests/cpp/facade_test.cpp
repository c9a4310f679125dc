// Exercises the C++ facade (rampack_amd/host/rampack_gpu.hpp) the way the reference's own validation tests drive
// Simulation (test/validation_tests/SimulationTest.cpp): build traits, samplers and a box scaler, integrate, read
// the packing back.  Needs a GPU; prints one line per check and exits non-zero on the first failure.
#include "rampack_gpu.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <cstdlib>

using namespace rampack_amd;

#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);     \
            std::exit(1);                                                  \
        }                                                                  \
        std::printf("ok   %s\n", #cond);                                   \
    } while (0)

template <typename E, typename F>
bool throws(F f) {
    try {
        f();
    } catch (const E &) {
        return true;
    }
    return false;
}

int main() {
    CHECK(throws<PreconditionException>([] { SpherocylinderTraits(-1, 1); }));
    CHECK(throws<PreconditionException>([] { RototranslationSampler(0.5, std::nullopt, 0.1); }));
    CHECK(throws<PreconditionException>([] { AnisotropicVolumeScaler(AnisotropicVolumeScaler::LOG, {0, 3, 0}, false); }));
    CHECK(TriclinicDeltaScaler(0.1, false).descriptor().one_side == 1);

    // 1000 hard spheres of radius 0.5 on a 10^3 sc lattice at packing fraction 0.36, NpT p = 5 with log moves
    SphereTraits traits(0.5);
    double side = std::cbrt(1000 * traits.getVolume() / 0.36);
    std::vector<Shape> shapes;
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 10; j++)
            for (int k = 0; k < 10; k++) {
                Shape s;
                s.position = {(i + 0.5) * side / 10, (j + 0.5) * side / 10, (k + 0.5) * side / 10};
                shapes.push_back(s);
            }
    Box box{side, 0, 0, 0, side, 0, 0, 0, side};
    auto packing = std::make_unique<Packing>(box, shapes, traits.getInteraction());
    std::vector<std::unique_ptr<MoveSampler>> samplers;
    samplers.push_back(std::make_unique<RototranslationSampler>(0.1, std::nullopt));
    auto scaler = std::make_unique<TriclinicAdapter>(std::make_unique<LogIsotropicVolumeScaler>(), 1e-4);
    Simulation sim(std::move(packing), std::move(samplers), 1234, std::move(scaler));
    std::vector<double> phis;
    sim.integrate(1.0, 5.0, 600, 300, 50, [&](const Packing &p, std::size_t) {
        phis.push_back(p.getPackingFraction(traits.getVolume()));
    });
    for (double phi : phis) std::printf("phi %.5f\n", phi);
    std::printf("moves %llu/%llu box %llu/%llu steps %g %g %g\n", (unsigned long long)sim.getCounters().moves_accepted,
                (unsigned long long)sim.getCounters().moves_attempted,
                (unsigned long long)sim.getCounters().box_accepted, (unsigned long long)sim.getCounters().box_attempted,
                sim.getCounters().trans_step, sim.getCounters().rot_step, sim.getCounters().box_step);
    CHECK(phis.size() == 6);
    CHECK(sim.getPacking().countTotalOverlaps(traits.getInteraction(), false) == 0);
    CHECK(sim.getCounters().moves_attempted == 900 * 1000);
    CHECK(phis.back() > 0.34 && phis.back() < 0.42);  // a hard-sphere fluid at p = 5 (CS: 0.393)
    CHECK(sim.getCounters().box_accepted > 0);
    CHECK(sim.getMoveAcceptanceRate() > 0.05 && sim.getMoveAcceptanceRate() < 0.6);

    // separate translation and rotation samplers (SimulationTest.cpp:134-146 uses them the same way)
    {
        auto p3 = std::make_unique<Packing>(box, shapes, traits.getInteraction());
        std::vector<std::unique_ptr<MoveSampler>> s4;
        s4.push_back(std::make_unique<TranslationSampler>(0.1));
        s4.push_back(std::make_unique<RotationSampler>(0.5));
        Simulation twoSamplers(std::move(p3), std::move(s4), 77,
                               std::make_unique<TriclinicAdapter>(std::make_unique<LogIsotropicVolumeScaler>(), 1e-4));
        twoSamplers.integrate(1.0, 5.0, 300, 300, 100);
        CHECK(twoSamplers.getCounters().moves_attempted == 2 * 600 * 1000);
        CHECK(twoSamplers.getPacking().countTotalOverlaps(traits.getInteraction(), false) == 0);
        CHECK(twoSamplers.getPacking().getPackingFraction(traits.getVolume()) > 0.34);
    }

    // Packing::tryScaling / revertScaling (Packing.cpp:172-208, 363-371) on the lattice
    {
        Packing pk(box, shapes, traits.getInteraction());
        Box grown = box, crushed = box;
        for (double &x : grown) x *= 1.05;
        for (double &x : crushed) x *= 0.8;  // lattice spacing below the sphere diameter
        CHECK(pk.tryScaling(grown, traits.getInteraction()) == 0);
        CHECK(pk.getBox() == grown);
        CHECK(std::isinf(pk.tryScaling(crushed, traits.getInteraction())));
        CHECK(pk.getBox() == crushed);
        pk.revertScaling();
        CHECK(pk.getBox() == grown);
        CHECK(pk.countTotalOverlaps(traits.getInteraction(), false) == 0);
    }

    // a RAMTRJ recording of a run, replayed into another packing (SimulationIOTest.cpp:37-98)
    {
        std::string path = "/tmp/rampack_amd_facade_test.ramtrj";
        auto p2 = std::make_unique<Packing>(box, shapes, traits.getInteraction());
        std::vector<std::unique_ptr<MoveSampler>> s3;
        s3.push_back(std::make_unique<RototranslationSampler>(0.1, std::nullopt));
        Simulation recSim(std::move(p2), std::move(s3), 99,
                          std::make_unique<TriclinicAdapter>(std::make_unique<LogIsotropicVolumeScaler>(), 1e-4));
        std::vector<std::unique_ptr<SimulationRecorder>> recorders;
        recorders.push_back(std::make_unique<RamtrjRecorder>(path, shapes.size(), 100, false));
        recSim.integrate(1.0, 5.0, 250, 250, 50, {}, 100, recorders);
        CHECK(recorders.front()->getLastCycleNumber() == 500);
        recorders.front()->close();
        Packing replay(box, shapes, traits.getInteraction());
        RamtrjPlayer::AutoFix autoFix(shapes.size());
        RamtrjPlayer player(path, autoFix);
        CHECK(!autoFix.wasFixingNeeded() && autoFix.wasFixingSuccessful());
        CHECK(player.getTotalCycles() == 500 && player.getCycleStep() == 100);
        int snapshots = 0;
        while (player.hasNext()) {
            player.nextSnapshot(replay);
            snapshots++;
        }
        CHECK(snapshots == 5);
        const Packing &live = recSim.getPacking();
        double worst = 0;
        bool exact = live.getBox() == replay.getBox();
        for (std::size_t i = 0; i < live.size(); i++) {
            exact = exact && live[i].position == replay[i].position;
            for (int k = 0; k < 9; k++)
                worst = std::max(worst, std::abs(live[i].orientation[k] - replay[i].orientation[k]));
        }
        CHECK(exact);
        CHECK(worst < 1e-12);
        player.jumpToSnapshot(replay, 200);
        CHECK(player.getCurrentSnapshotCycles() == 200);
        CHECK(throws<PreconditionException>([&] { player.jumpToSnapshot(replay, 250); }));
        std::remove(path.c_str());
    }

    // overlapping start: Simulation::integrate refuses it (Simulation.cpp:131-132)
    std::vector<Shape> bad(2);
    bad[0].position = {1, 1, 1};
    bad[1].position = {1.5, 1, 1};
    auto badPacking = std::make_unique<Packing>(Box{10, 0, 0, 0, 10, 0, 0, 0, 10}, bad, traits.getInteraction());
    std::vector<std::unique_ptr<MoveSampler>> s2;
    s2.push_back(std::make_unique<RototranslationSampler>(0.1, std::nullopt));
    Simulation badSim(std::move(badPacking), std::move(s2), 1, nullptr);
    CHECK(throws<ValidationException>([&] { badSim.integrate(1.0, 1.0, 10, 0, 0); }));
    std::printf("facade ok\n");
    return 0;
}
