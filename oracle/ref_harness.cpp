// ref_harness.cpp -- C entry points over the REFERENCE's own compiled hot path.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's source files (read in place under the reference tree) into
// oracle/_ref/libref.so.  Nothing of the reference is copied: this file only
// calls its public classes (Packing, SpherocylinderTraits, SphereTraits,
// PolysphereTraits, RototranslationSampler, TriclinicDeltaScaler,
// TriclinicAdapter, DomainDecomposition, NematicOrder) the way the reference's
// own Simulation and tests do.
//
// The reference's Simulation.cpp cannot be compiled here (it includes the
// absent ZipIterator submodule), so ref_mc_run restates its cycle loop
// (Simulation.cpp:325-531, 533-607) on top of the reference's Packing /
// samplers / scalers / DomainDecomposition.  Draw order from the per-domain
// std::mt19937 streams follows Simulation.cpp:401-404, 456-458, 483, 499-525.

#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <chrono>
#include <memory>
#include <numeric>
#include <random>
#include <vector>
#include <omp.h>

#include "core/Packing.h"
#include "core/PeriodicBoundaryConditions.h"
#include "core/DomainDecomposition.h"
#include "core/shapes/SphereTraits.h"
#include "core/shapes/SpherocylinderTraits.h"
#include "core/shapes/PolysphereTraits.h"
#include "core/interactions/RepulsiveLennardJonesInteraction.h"
#include "core/interactions/LennardJonesInteraction.h"
#include "core/interactions/SquareInverseCoreInteraction.h"
#include "core/interactions/CompoundInteraction.h"
#include "core/move_samplers/RototranslationSampler.h"
#include "core/volume_scalers/TriclinicDeltaScaler.h"
#include "core/volume_scalers/TriclinicAdapter.h"
#include "core/volume_scalers/AnisotropicVolumeScaler.h"
#include "core/volume_scalers/LogScalingFactorSampler.h"
#include "core/volume_scalers/DeltaVolumeScaler.h"
#include "core/observables/NematicOrder.h"
#include "core/io/RamsnapReader.h"
#include "core/io/RamsnapWriter.h"
#include "core/io/RamtrjRecorder.h"
#include "core/io/RamtrjPlayer.h"
#include <fstream>
#include <sstream>

namespace {

// The BoundaryConditions the reference's Packing uses for neighbour-grid pairs (a fixed cell translation); the
// reference keeps its own copy in an anonymous namespace, so the harness supplies the same behaviour through the
// public BoundaryConditions interface.
class FixedTranslation : public BoundaryConditions {
    Vector<3> t;
public:
    explicit FixedTranslation(const Vector<3> &t) : t{t} {}
    void setBox(const TriclinicBox &) override {}
    [[nodiscard]] Vector<3> getCorrection(const Vector<3> &) const override { return {}; }
    [[nodiscard]] Vector<3> getTranslation(const Vector<3> &, const Vector<3> &) const override { return t; }
    [[nodiscard]] double getDistance2(const Vector<3> &p1, const Vector<3> &p2) const override {
        return (p2 + this->t - p1).norm2();
    }
};

std::unique_ptr<ShapeTraits> makeShapeTraits(int kind, double radius, double length, int nCentres,
                                             const double *centres, const double *radii, int soft, double eps,
                                             double sigma) {
    if (kind == 0) {
        if (soft == 1)
            return std::make_unique<SphereTraits>(radius, std::make_shared<RepulsiveLennardJonesInteraction>(eps, sigma));
        if (soft == 3)
            return std::make_unique<SphereTraits>(radius, std::make_shared<LennardJonesInteraction>(eps, sigma));
        if (soft == 4)
            return std::make_unique<SphereTraits>(radius, std::make_shared<SquareInverseCoreInteraction>(eps, sigma));
        return std::make_unique<SphereTraits>(radius);
    }
    if (kind == 1)
        return std::make_unique<SpherocylinderTraits>(length, radius);
    std::vector<PolysphereTraits::SphereData> data;
    for (int i = 0; i < nCentres; i++)
        data.emplace_back(Vector<3>{centres[3 * i], centres[3 * i + 1], centres[3 * i + 2]}, radii[i]);
    return std::make_unique<PolysphereTraits>(PolysphereTraits::PolysphereGeometry(data));
}

// The shape's traits and the interaction the harness runs with: the traits' own, or, for a soft kind carrying the
// hard-core flag 0x100, CompoundInteraction (CompoundInteraction.cpp:9-60) of the hard traits' interaction and the
// soft traits' interaction (spheres only: the reference has no WCA on the segment distance).
struct Model {
    std::unique_ptr<ShapeTraits> traits, softTraits;
    std::unique_ptr<CompoundInteraction> compound;
    const Interaction &getInteraction() const { return compound ? *compound : traits->getInteraction(); }
    const ShapeGeometry &getGeometry() const { return traits->getGeometry(); }
};

std::unique_ptr<Model> makeTraits(int kind, double radius, double length, int nCentres, const double *centres,
                                  const double *radii, int soft, double eps, double sigma) {
    auto m = std::make_unique<Model>();
    if ((soft & 0x100) && (soft & 0xff)) {
        m->traits = makeShapeTraits(kind, radius, length, nCentres, centres, radii, 0, eps, sigma);
        m->softTraits = makeShapeTraits(kind, radius, length, nCentres, centres, radii, soft & 0xff, eps, sigma);
        m->compound = std::make_unique<CompoundInteraction>(m->traits->getInteraction(),
                                                            m->softTraits->getInteraction());
    } else {
        m->traits = makeShapeTraits(kind, radius, length, nCentres, centres, radii, soft & 0xff, eps, sigma);
    }
    return m;
}

Matrix<3, 3> mat(const double *m) {
    return Matrix<3, 3>{m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8]};
}

std::vector<Shape> makeShapes(std::size_t n, const double *pos, const double *rot) {
    std::vector<Shape> shapes;
    shapes.reserve(n);
    for (std::size_t i = 0; i < n; i++)
        shapes.emplace_back(Vector<3>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, mat(rot + 9 * i));
    return shapes;
}

// Simulation::Counter (Simulation.cpp:610-640)
struct Counter {
    std::size_t moves = 0, accepted = 0, sinceEval = 0, acceptedSinceEval = 0;
    void increment(bool a) { moves++; sinceEval++; if (a) { accepted++; acceptedSinceEval++; } }
    double rate() const { return sinceEval ? (double)acceptedSinceEval / (double)sinceEval : 0; }
    void resetCurrent() { sinceEval = 0; acceptedSinceEval = 0; }
    Counter &operator+=(const Counter &o) {
        moves += o.moves; accepted += o.accepted; sinceEval += o.sinceEval; acceptedSinceEval += o.acceptedSinceEval;
        return *this;
    }
};

}  // namespace

extern "C" {

int ref_pair_decisions(int kind, double radius, double length, std::size_t n, const double *rec, std::uint8_t *out) {
    auto traits = makeTraits(kind, radius, length, 0, nullptr, nullptr, 0, 1, 1);
    const Interaction &inter = traits->getInteraction();
    for (std::size_t i = 0; i < n; i++) {
        const double *q = rec + 27 * i;
        Vector<3> p1{q[0], q[1], q[2]}, p2{q[12], q[13], q[14]};
        FixedTranslation bc(Vector<3>{q[24], q[25], q[26]});
        out[i] = inter.overlapBetween(p1, mat(q + 3), 0, p2, mat(q + 15), 0, bc) ? 1 : 0;
    }
    return 0;
}

// Interaction::calculateEnergyBetween through the reference's RepulsiveLennardJonesInteraction on a sphere
// (CentralInteraction.h:58-67, RepulsiveLennardJonesInteraction.cpp:8-16) with a fixed cell translation.  Only
// soft == 1 (WCA), 3 (Lennard-Jones) and 4 (square-inverse core) exist in the reference; other values return -1.
int ref_pair_energies(int kind, double radius, double length, int soft, double eps, double sigma, std::size_t n,
                      const double *rec, double *out) {
    if (kind != 0 || (soft != 1 && soft != 3 && soft != 4)) return -1;
    auto traits = makeTraits(kind, radius, length, 0, nullptr, nullptr, soft, eps, sigma);
    const Interaction &inter = traits->getInteraction();
    for (std::size_t i = 0; i < n; i++) {
        const double *q = rec + 27 * i;
        Vector<3> p1{q[0], q[1], q[2]}, p2{q[12], q[13], q[14]};
        FixedTranslation bc(Vector<3>{q[24], q[25], q[26]});
        out[i] = inter.calculateEnergyBetween(p1, mat(q + 3), 0, p2, mat(q + 15), 0, bc);
    }
    return 0;
}

int ref_snapshot_stats(int kind, double radius, double length, int nCentres, const double *centres,
                       const double *radii, int soft, double eps, double sigma, const double *box, std::size_t n,
                       const double *pos, const double *rot, int earlyExit, std::uint64_t *overlaps,
                       double *energy) {
    try {
        auto traits = makeTraits(kind, radius, length, nCentres, centres, radii, soft, eps, sigma);
        const Interaction &inter = traits->getInteraction();
        Packing packing(TriclinicBox(mat(box)), makeShapes(n, pos, rot), std::make_unique<PeriodicBoundaryConditions>(),
                        inter, 1, 1);
        *overlaps = inter.hasHardPart() ? packing.countTotalOverlaps(inter, earlyExit != 0) : 0;
        if (energy) *energy = packing.getTotalEnergy(inter);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// TriclinicBox (TriclinicBox.h:49-70, TriclinicBox.cpp:27-51): the inverse the box keeps (Matrix::inverse), read
// column by column through absoluteToRelative of the unit vectors (inv * e_j adds only exact zeros to inv(i, j)),
// the heights and the volume.
void ref_box_info(const double *box, double *inv, double *heights, double *volume) {
    TriclinicBox b(mat(box));
    for (int j = 0; j < 3; j++) {
        Vector<3> e{};
        e[j] = 1;
        Vector<3> c = b.absoluteToRelative(e);
        for (int i = 0; i < 3; i++) inv[3 * i + j] = c[i];
    }
    auto h = b.getHeights();
    for (int i = 0; i < 3; i++) heights[i] = h[i];
    *volume = b.getVolume();
}

// Shape::translate with the periodic correction (Shape.cpp:9-12, PeriodicBoundaryConditions.cpp:11-22), the step of
// Packing::tryTranslation / tryMove (Packing.cpp:84-164), for n positions and translations.
void ref_translate(const double *box, std::size_t n, const double *pos, const double *t, double *out) {
    PeriodicBoundaryConditions pbc(TriclinicBox(mat(box)));
    for (std::size_t i = 0; i < n; i++) {
        Shape s(Vector<3>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
        s.translate(Vector<3>{t[3 * i], t[3 * i + 1], t[3 * i + 2]}, pbc);
        for (int d = 0; d < 3; d++) out[3 * i + d] = s.getPosition()[d];
    }
}

// Matrix<3, 3>::rotation (Matrix.tpp:438-450) for n (axis, angle) pairs; normalize != 0 passes axis.normalized() as
// RototranslationSampler::sampleMove does (RototranslationSampler.cpp:39-43).
void ref_rotation(std::size_t n, const double *axis, const double *angle, int normalize, double *out) {
    for (std::size_t i = 0; i < n; i++) {
        Vector<3> a{axis[3 * i], axis[3 * i + 1], axis[3 * i + 2]};
        Matrix<3, 3> R = Matrix<3, 3>::rotation(normalize ? a.normalized() : a, angle[i]);
        for (int e = 0; e < 9; e++) out[9 * i + e] = R(e / 3, e % 3);
    }
}

// Shape::rotate (Shape.cpp:14-16): O' = R O, for n pairs.
void ref_rotate(std::size_t n, const double *R, const double *O, double *out) {
    for (std::size_t i = 0; i < n; i++) {
        Shape s(Vector<3>{0, 0, 0}, mat(O + 9 * i));
        s.rotate(mat(R + 9 * i));
        for (int e = 0; e < 9; e++) out[9 * i + e] = s.getOrientation()(e / 3, e % 3);
    }
}

// Packing::tryScaling (Packing.cpp:172-208) on hard spheres of radius `radius` (kind 0) or spherocylinders (kind 1):
// the positions it leaves (newBox.relativeToAbsolute(oldBox.absoluteToRelative(pos)), Packing.cpp:183-184) and its
// return value (+inf for an overlap, else the energy change, 0 for hard shapes).
int ref_try_scaling(int kind, double radius, double length, const double *oldBox, const double *newBox, std::size_t n,
                    const double *pos, const double *rot, double *outPos, double *dE) {
    try {
        auto traits = makeTraits(kind, radius, length, 0, nullptr, nullptr, 0, 1, 1);
        const Interaction &inter = traits->getInteraction();
        Packing packing(TriclinicBox(mat(oldBox)), makeShapes(n, pos, rot),
                        std::make_unique<PeriodicBoundaryConditions>(), inter, 1, 1);
        *dE = packing.tryScaling(TriclinicBox(mat(newBox)), inter);
        for (std::size_t i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) outPos[3 * i + d] = packing[i].getPosition()[d];
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// PackingFraction::calculate (PackingFraction.h:19-23) = Packing::getPackingFraction (Packing.cpp:225-232) with the
// shape's ShapeGeometry::getVolume and TriclinicBox::getVolume (TriclinicBox.cpp:35-38).
double ref_packing_fraction(int kind, double radius, double length, const double *box, std::size_t n) {
    auto traits = makeTraits(kind, radius, length, 0, nullptr, nullptr, 0, 1, 1);
    std::vector<double> pos(3 * n, 0.0), rot(9 * n, 0.0);
    for (std::size_t i = 0; i < n; i++) rot[9 * i] = rot[9 * i + 4] = rot[9 * i + 8] = 1;
    TriclinicBox b(mat(box));
    auto side = b.getSides();
    for (std::size_t i = 0; i < n; i++)  // any in-box position: only the count matters
        for (int d = 0; d < 3; d++) pos[3 * i + d] = 0.5 * side[0][d] + 0.25 * side[1][d] + 0.125 * side[2][d];
    Packing packing(b, makeShapes(n, pos.data(), rot.data()), std::make_unique<PeriodicBoundaryConditions>(),
                    traits->getInteraction(), 1, 1);
    return packing.getPackingFraction(traits->getGeometry().getVolume());
}

// Packing::fixRotationMatrix (Packing.cpp:1211-1216, private) through the public
// Packing::renormalizeOrientations(interaction, false) on particles 10 apart, where no fix can create an overlap, so
// every fix is accepted; n row-major matrices in place.
int ref_fix_rotation(std::size_t n, double *rot) {
    try {
        SpherocylinderTraits traits(1, 0.5);
        std::size_t side = 1;
        while (side * side * side < n) side++;
        const double L = 10.0 * (double)side;
        std::vector<double> pos(3 * n);
        for (std::size_t i = 0; i < n; i++) {
            pos[3 * i] = 10.0 * (double)(i % side) + 5;
            pos[3 * i + 1] = 10.0 * (double)((i / side) % side) + 5;
            pos[3 * i + 2] = 10.0 * (double)(i / (side * side)) + 5;
        }
        Packing packing(TriclinicBox(std::array<double, 3>{L, L, L}), makeShapes(n, pos.data(), rot),
                        std::make_unique<PeriodicBoundaryConditions>(), traits.getInteraction(), 1, 1);
        if (packing.renormalizeOrientations(traits.getInteraction(), false) != 0) return -2;
        for (std::size_t i = 0; i < n; i++) {
            const auto &O = packing[i].getOrientation();
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) rot[9 * i + 3 * a + b] = O(a, b);
        }
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// RamsnapWriter::write of a packing of spheres with the given state; aux_text as "key value" lines.
int ref_ramsnap_write(const char *path, const double *box, std::size_t n, const double *pos, const double *rot,
                      const char *aux_text) {
    try {
        SphereTraits traits(0.5);
        Packing packing(TriclinicBox(mat(box)), makeShapes(n, pos, rot), std::make_unique<PeriodicBoundaryConditions>(),
                        traits.getInteraction(), 1, 1);
        std::map<std::string, std::string> aux;
        std::istringstream in(aux_text ? aux_text : "");
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string key, value;
            if (!(ls >> key)) continue;
            ls >> std::ws;
            std::getline(ls, value);
            aux[key] = value;
        }
        std::ofstream f(path);
        RamsnapWriter().write(f, packing, aux);
        return f ? 0 : -1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ref_ramsnap_write: %s\n", e.what());
        return -1;
    }
}

// RamsnapReader::read into n-particle buffers; returns the particle count, or -1 on a broken file.
long ref_ramsnap_read(const char *path, double *box, std::size_t n, double *pos, double *rot) {
    try {
        SphereTraits traits(0.5);
        Packing packing(std::make_unique<PeriodicBoundaryConditions>());
        std::ifstream f(path);
        RamsnapReader().read(f, packing, traits.getInteraction());
        const auto &d = packing.getBox().getDimensions();
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) box[3 * a + b] = d(a, b);
        for (std::size_t i = 0; i < packing.size() && i < n; i++) {
            const auto &p = packing[i].getPosition();
            const auto &O = packing[i].getOrientation();
            for (int a = 0; a < 3; a++) pos[3 * i + a] = p[a];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) rot[9 * i + 3 * a + b] = O(a, b);
        }
        return (long)packing.size();
    } catch (const std::exception &) {
        return -1;
    }
}

// RamtrjRecorder over a file opened as SimulationRecorderFactory does: snapshots k = 0..n_snap-1 (boxes 9 doubles,
// positions n x 3, orientations n x 9 each, stacked) recorded at cycles first_cycle + k * cycle_step; append
// continues an existing file.  The recorder is closed (header rewritten) on return.  0, or -1 on an exception.
int ref_ramtrj_write(const char *path, std::size_t n, std::size_t cycle_step, std::size_t n_snap, const double *boxes,
                     const double *pos, const double *rot, int append, std::size_t first_cycle) {
    try {
        SphereTraits traits(0.5);
        auto mode = std::ios_base::in | std::ios_base::out | std::ios_base::binary;
        auto f = std::make_unique<std::fstream>(path, append ? mode : mode | std::ios_base::trunc);
        if (!*f) return -1;
        RamtrjRecorder recorder(std::move(f), n, cycle_step, append != 0);
        for (std::size_t k = 0; k < n_snap; k++) {
            Packing packing(TriclinicBox(mat(boxes + 9 * k)), makeShapes(n, pos + 3 * n * k, rot + 9 * n * k),
                            std::make_unique<PeriodicBoundaryConditions>(), traits.getInteraction(), 1, 1);
            recorder.recordSnapshot(packing, first_cycle + k * cycle_step);
        }
        recorder.close();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ref_ramtrj_write: %s\n", e.what());
        return -1;
    }
}

// RamtrjPlayer over the file (expected_n 0: strict constructor; else with AutoFix(expected_n)); jump_cycle 0 reads
// every snapshot in order with nextSnapshot, otherwise jumpToSnapshot(jump_cycle) once.  Up to max_snap snapshots
// go to boxes/pos/rot (stacked as in ref_ramtrj_write); info receives particles, snapshots, cycle step, total
// cycles, fixing needed, fixing successful, inferred snapshots, bytes remainder.  Returns the number of snapshots
// read, or -1 on an exception (info is still filled from the AutoFix report).
long ref_ramtrj_read(const char *path, std::size_t expected_n, std::size_t jump_cycle, std::size_t n,
                     std::size_t max_snap, double *boxes, double *pos, double *rot, unsigned long long *info) {
    std::unique_ptr<RamtrjPlayer::AutoFix> autoFix;
    if (expected_n) autoFix = std::make_unique<RamtrjPlayer::AutoFix>(expected_n);
    auto report = [&]() {
        if (!autoFix) return;
        info[4] = autoFix->wasFixingNeeded();
        info[5] = autoFix->wasFixingSuccessful();
        info[6] = autoFix->getInferredSnapshots();
        info[7] = autoFix->getBytesRemainder();
    };
    try {
        SphereTraits traits(0.5);
        auto in = std::make_unique<std::ifstream>(path, std::ios_base::in | std::ios_base::binary);
        std::unique_ptr<RamtrjPlayer> player;
        if (autoFix) player = std::make_unique<RamtrjPlayer>(std::move(in), *autoFix);
        else player = std::make_unique<RamtrjPlayer>(std::move(in));
        info[0] = player->getNumMolecules();
        info[2] = player->getCycleStep();
        info[3] = player->getTotalCycles();
        info[1] = info[2] ? info[3] / info[2] : 0;
        report();
        // nextSnapshot expects a packing of the trajectory's size (RamtrjPlayer.cpp:56); its state is replaced
        std::vector<double> p0(3 * n, 0.5), r0(9 * n, 0.0);
        for (std::size_t i = 0; i < n; i++) r0[9 * i] = r0[9 * i + 4] = r0[9 * i + 8] = 1;
        Packing packing(TriclinicBox(std::array<double, 3>{1, 1, 1}), makeShapes(n, p0.data(), r0.data()),
                        std::make_unique<PeriodicBoundaryConditions>(), traits.getInteraction(), 1, 1);
        long got = 0;
        auto store = [&]() {
            if ((std::size_t)got >= max_snap) return;
            const auto &d = packing.getBox().getDimensions();
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) boxes[9 * got + 3 * a + b] = d(a, b);
            for (std::size_t i = 0; i < packing.size() && i < n; i++) {
                const auto &p = packing[i].getPosition();
                const auto &O = packing[i].getOrientation();
                for (int a = 0; a < 3; a++) pos[3 * n * got + 3 * i + a] = p[a];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) rot[9 * n * got + 9 * i + 3 * a + b] = O(a, b);
            }
            got++;
        };
        if (jump_cycle) {
            player->jumpToSnapshot(packing, traits.getInteraction(), jump_cycle);
            store();
        } else {
            while (player->hasNext()) {
                player->nextSnapshot(packing, traits.getInteraction());
                store();
            }
        }
        return got;
    } catch (const std::exception &e) {
        report();
        std::fprintf(stderr, "ref_ramtrj_read: %s\n", e.what());
        return -1;
    }
}

double ref_nematic_p2(std::size_t n, const double *rot) {
    SpherocylinderTraits traits(1, 1);
    std::vector<double> pos(3 * n, 0.5);
    Packing packing(TriclinicBox(std::array<double, 3>{1, 1, 1}), makeShapes(n, pos.data(), rot),
                    std::make_unique<PeriodicBoundaryConditions>(), traits.getInteraction(), 1, 1);
    NematicOrder q;
    q.calculate(packing, 1, 1, traits);
    return q.getIntervalValues().front();
}

// One NpT (or NVT) run of the reference's hot path with the restated cycle loop.  box_kind: 0 none (NVT), 1
// delta_triclinic(scale_together=True), 2 log(spec="isotropic"), 3 delta_v.  soft 1: WCA sphere (eps, sigma).
// rot_step <= 0 selects the automatic rotation step (RototranslationSampler.cpp:95-114).  Samples of (packing
// fraction, P2 (spherocylinders, else 0), total energy / N) are taken every avg_every cycles of the averaging
// phase.  The return value is the number of samples, or a negative number on error.
// Bead data for polyspheres (kind 2) in ref_mc_run, which takes no centres of its own.
static std::vector<double> gBeadCentres, gBeadRadii;
void ref_set_beads(int nCentres, const double *centres, const double *radii) {
    gBeadCentres.assign(centres, centres + 3 * nCentres);
    gBeadRadii.assign(radii, radii + nCentres);
}

long ref_mc_run(int kind, double radius, double length, int soft, double eps, double sigma, const double *box,
                std::size_t n, const double *pos, const double *rot, double temperature, double pressure,
                double transStep, double rotStep, int boxKind, double boxStep, unsigned long seed, int ddx, int ddy,
                int ddz, int threads, std::size_t thermCycles, std::size_t avgCycles, std::size_t avgEvery,
                double *samples, double *outBox, double *outPos, double *outRot, double *seconds,
                double *finalSteps) {
    try {
        std::array<std::size_t, 3> dd{(std::size_t)ddx, (std::size_t)ddy, (std::size_t)ddz};
        std::size_t numDomains = dd[0] * dd[1] * dd[2];
        auto traits = makeTraits(kind, radius, length, kind == 2 ? (int)gBeadRadii.size() : 0, gBeadCentres.data(),
                                 gBeadRadii.data(), soft, eps, sigma);
        const Interaction &inter = traits->getInteraction();
        auto packing = std::make_unique<Packing>(TriclinicBox(mat(box)), makeShapes(n, pos, rot),
                                                 std::make_unique<PeriodicBoundaryConditions>(), inter,
                                                 (std::size_t)threads, (std::size_t)threads);
        RototranslationSampler sampler(transStep, rotStep > 0 ? std::optional<double>(rotStep) : std::nullopt, 0);
        sampler.setupForShapeTraits(*traits->traits);
        std::unique_ptr<TriclinicBoxScaler> scaler;
        if (boxKind == 1) {
            scaler = std::make_unique<TriclinicDeltaScaler>(boxStep, true);
        } else if (boxKind == 2) {
            auto dir = AnisotropicVolumeScaler::X & AnisotropicVolumeScaler::Y & AnisotropicVolumeScaler::Z;
            scaler = std::make_unique<TriclinicAdapter>(
                std::make_unique<AnisotropicVolumeScaler>(std::make_unique<LogScalingFactorSampler>(), dir, false),
                boxStep);
        } else if (boxKind == 3) {
            scaler = std::make_unique<TriclinicAdapter>(std::make_unique<DeltaVolumeScaler>(), boxStep);
        }

        std::vector<std::mt19937> mts;
        for (std::size_t i = 0; i < numDomains; i++) mts.emplace_back(seed + i);
        std::uniform_real_distribution<double> unit;
        std::vector<std::size_t> all(n);
        std::iota(all.begin(), all.end(), 0);
        Counter moveCounter, scalingCounter;
        double volume0 = traits->getGeometry().getVolume();

        // Simulation::tryMove (Simulation.cpp:447-491), one sampler
        auto tryMove = [&](const std::vector<std::size_t> &idxs, Counter &counter, std::size_t numMoves,
                           std::optional<ActiveDomain> bounds) {
            auto &mt = mts[omp_get_thread_num()];
            std::uniform_int_distribution<std::size_t> moveDistribution(0, numMoves - 1);
            (void)moveDistribution(mt);
            auto move = sampler.sampleMove(*packing, idxs, mt);
            double dE = packing->tryMove(move.particleIdx, move.translation, move.rotation, inter, bounds);
            if (unit(mt) <= std::exp(-dE / temperature)) {
                packing->acceptMove();
                counter.increment(true);
            } else {
                counter.increment(false);
            }
        };

        // Simulation::performMoves (Simulation.cpp:365-445)
        auto performMoves = [&]() {
            if (numDomains == 1) {
                for (std::size_t i = 0; i < n; i++) tryMove(all, moveCounter, n, std::nullopt);
                return;
            }
            while (true) {
                try {
                    const auto &pbox = packing->getBox();
                    auto &mt = mts[0];
                    Vector<3> origin{unit(mt), unit(mt), unit(mt)};
                    origin = pbox.relativeToAbsolute(origin);
                    auto ngDiv = packing->getNeighbourGridCellDivisions();
                    DomainDecomposition dec(*packing, inter, dd, ngDiv, origin);
                    std::vector<Counter> temp((std::size_t)omp_get_max_threads());
                    std::size_t avgN = n / numDomains;
                    #pragma omp parallel for collapse(3) num_threads(packing->getMoveThreads())
                    for (std::size_t i = 0; i < dd[0]; i++)
                        for (std::size_t j = 0; j < dd[1]; j++)
                            for (std::size_t k = 0; k < dd[2]; k++) {
                                std::array<std::size_t, 3> c{i, j, k};
                                const auto &idxs = dec.getParticlesInRegion(c);
                                auto active = dec.getActiveDomainBounds(c);
                                if (idxs.empty()) continue;
                                for (std::size_t x = 0; x < avgN; x++)
                                    tryMove(idxs, temp[omp_get_thread_num()], avgN, active);
                            }
                    for (auto &c : temp) moveCounter += c;
                    break;
                } catch (const TooNarrowDomainException &ex) {
                    dd[ex.getCoord()]--;
                    numDomains = dd[0] * dd[1] * dd[2];
                    if (numDomains == 1) {
                        for (std::size_t i = 0; i < n; i++) tryMove(all, moveCounter, n, std::nullopt);
                        break;
                    }
                }
            }
        };

        // Simulation::tryScaling (Simulation.cpp:493-531), hard interaction
        auto tryScaling = [&]() -> bool {
            auto &mt = mts.front();
            TriclinicBox oldBox = packing->getBox();
            TriclinicBox newBox = scaler->updateBox(oldBox, mt);
            double oldV = std::abs(oldBox.getVolume()), newV = std::abs(newBox.getVolume());
            double deltaV = newV - oldV, factor = newV / oldV;
            double N = (double)packing->size();
            if (inter.hasSoftPart()) {
                double dE = packing->tryScaling(newBox, inter);
                double exponent = N * log(factor) - dE / temperature - pressure * deltaV / temperature;
                if (unit(mt) <= std::exp(exponent)) return true;
                packing->revertScaling();
                return false;
            }
            double exponent = N * log(factor) - pressure * deltaV / temperature;
            if (unit(mt) > std::exp(exponent)) return false;
            if (packing->tryScaling(newBox, inter) > 0) {
                packing->revertScaling();
                return false;
            }
            return true;
        };

        // Simulation::evaluateCounters (Simulation.cpp:533-607)
        auto evaluate = [&]() {
            if (moveCounter.sinceEval >= 100 * n) {
                double r = moveCounter.rate();
                moveCounter.resetCurrent();
                if (r > 0.2) sampler.increaseStepSize();
                else if (r < 0.1) sampler.decreaseStepSize();
            }
            if (scaler && scalingCounter.sinceEval >= 100) {
                double r = scalingCounter.rate();
                scalingCounter.resetCurrent();
                if (r > 0.2) scaler->increaseStepSize();
                else if (r < 0.1) scaler->decreaseStepSize();
            }
        };

        packing->toggleOverlapCounting(false, inter);
        if (packing->countTotalOverlaps(inter) != 0) return -2;

        long nSamples = 0;
        std::size_t totalCycles = 0;
        auto start = std::chrono::high_resolution_clock::now();
        for (std::size_t cyc = 0; cyc < thermCycles + avgCycles; cyc++) {
            bool therm = cyc < thermCycles;
            performMoves();
            if (scaler) scalingCounter.increment(tryScaling());
            if (therm) evaluate();
            totalCycles++;
            if (totalCycles % 10000 == 0) packing->renormalizeOrientations(inter, false);
            if (!therm && avgEvery && (totalCycles - thermCycles) % avgEvery == 0 && samples) {
                samples[3 * nSamples] = packing->getPackingFraction(volume0);
                samples[3 * nSamples + 2] = packing->getTotalEnergy(inter) / (double)n;
                samples[3 * nSamples + 1] = 0;
                if (kind != 1) {
                    nSamples++;
                    continue;
                }
                std::vector<double> r(9 * n);
                for (std::size_t i = 0; i < n; i++) {
                    const auto &O = (*packing)[i].getOrientation();
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) r[9 * i + 3 * a + b] = O(a, b);
                }
                NematicOrder q;
                q.calculate(*packing, temperature, pressure, *traits->traits);
                samples[3 * nSamples + 1] = q.getIntervalValues().front();
                nSamples++;
            }
        }
        auto end = std::chrono::high_resolution_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(end - start).count();
        if (outBox) {
            const auto &d = packing->getBox().getDimensions();
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) outBox[3 * a + b] = d(a, b);
        }
        for (std::size_t i = 0; outPos && i < n; i++) {
            const auto &p = (*packing)[i].getPosition();
            for (int a = 0; a < 3; a++) outPos[3 * i + a] = p[a];
        }
        for (std::size_t i = 0; outRot && i < n; i++) {
            const auto &O = (*packing)[i].getOrientation();
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) outRot[9 * i + 3 * a + b] = O(a, b);
        }
        if (finalSteps) {
            auto steps = sampler.getStepSizes();
            finalSteps[0] = steps[0].second;
            finalSteps[1] = steps[1].second;
            finalSteps[2] = scaler ? scaler->getStepSize() : 0;
        }
        return nSamples;
    } catch (const std::exception &) {
        return -1;
    }
}

}  // extern "C"
