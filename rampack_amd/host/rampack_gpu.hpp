// rampack_gpu.hpp -- C++17 facade over the C ABI (include/rampack_amd.h) with the reference's plug-in seams.
//
// The classes keep the names, constructor arguments and error behaviour of the reference types they stand in
// for, so a caller written against rampack's Simulation / Packing / ShapeTraits / MoveSampler /
// TriclinicBoxScaler API reads the same here:
//   ShapeTraits, Interaction, ShapeGeometry           src/core/ShapeTraits.h:20-42, src/core/Interaction.h:18-150
//   SphereTraits(radius)                              src/core/shapes/SphereTraits.h:60
//   SpherocylinderTraits(length, radius)              src/core/shapes/SpherocylinderTraits.h:53
//   RototranslationSampler(trans, rot, maxTrans)      src/core/move_samplers/RototranslationSampler.h:34-35
//   TriclinicDeltaScaler(step, scaleTogether)         src/core/volume_scalers/TriclinicDeltaScaler.h:25
//   TriclinicAdapter(volumeScaler, step)              src/core/volume_scalers/TriclinicAdapter.h:22-24
//   Packing(box, shapes, interaction)                 src/core/Packing.h:154-160
//   Simulation(packing, moveSamplers, seed, scaler)   src/core/Simulation.h:281-283
//   Simulation::integrate(T, p, therm, avg, every)    src/core/Simulation.h:321-340
//   RamtrjRecorder(path, n, cycleStep, append)        src/core/io/RamtrjRecorder.h:30-50 (a file path in place
//                                                     of the stream, opened as SimulationRecorderFactory.cpp:13-31)
//   RamtrjPlayer(path[, autoFix])                     src/core/io/RamtrjPlayer.h:19-120
// Physics plug-ins cannot run per pair across a device boundary, so every plug-in exports a plain descriptor and
// the whole sweep runs behind the C ABI.  Errors map to the reference's exception types (src/utils/Exceptions.h).
#pragma once

#include "rampack_amd.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rampack_amd {

struct PreconditionException : std::logic_error {
    using std::logic_error::logic_error;
};
struct ValidationException : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceException : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int code) {
    if (code == RP_OK) return;
    std::string msg = rp_last_error();
    if (code == RP_ERR_PRECONDITION) throw PreconditionException(msg);
    if (code == RP_ERR_VALIDATION) throw ValidationException(msg);
    throw DeviceException(msg);
}

// ---------------------------------------------------------------------------------------------- shapes / physics

class Interaction {
public:
    virtual ~Interaction() = default;
    [[nodiscard]] virtual bool hasHardPart() const = 0;
    [[nodiscard]] virtual bool hasSoftPart() const = 0;
    [[nodiscard]] virtual bool hasWallPart() const = 0;
    [[nodiscard]] virtual bool isConvex() const = 0;
    [[nodiscard]] virtual double getRangeRadius() const = 0;
    [[nodiscard]] virtual double getTotalRangeRadius() const { return this->getRangeRadius(); }
    // The device form of this interaction.
    [[nodiscard]] virtual rp_shape_desc descriptor() const = 0;
};

class ShapeGeometry {
public:
    virtual ~ShapeGeometry() = default;
    [[nodiscard]] virtual double getVolume() const = 0;
};

class ShapeTraits {
public:
    virtual ~ShapeTraits() = default;
    [[nodiscard]] virtual const Interaction &getInteraction() const = 0;
    [[nodiscard]] virtual const ShapeGeometry &getGeometry() const = 0;
};

class SphereTraits : public ShapeTraits, public Interaction, public ShapeGeometry {
    double radius;

public:
    explicit SphereTraits(double radius = 1) : radius{radius} {
        if (!(radius > 0)) throw PreconditionException("SphereTraits: radius > 0");
    }
    [[nodiscard]] const Interaction &getInteraction() const override { return *this; }
    [[nodiscard]] const ShapeGeometry &getGeometry() const override { return *this; }
    [[nodiscard]] bool hasHardPart() const override { return true; }
    [[nodiscard]] bool hasSoftPart() const override { return false; }
    [[nodiscard]] bool hasWallPart() const override { return true; }
    [[nodiscard]] bool isConvex() const override { return true; }
    [[nodiscard]] double getRangeRadius() const override { return 2 * this->radius; }
    [[nodiscard]] double getVolume() const override { return 4. / 3 * M_PI * std::pow(this->radius, 3); }
    [[nodiscard]] rp_shape_desc descriptor() const override { return {RP_SHAPE_SPHERE, this->radius, 0}; }
};

class SpherocylinderTraits : public ShapeTraits, public Interaction, public ShapeGeometry {
    double length, radius;

public:
    SpherocylinderTraits() : SpherocylinderTraits(1, 1) {}
    SpherocylinderTraits(double length, double radius) : length{length}, radius{radius} {
        if (!(length >= 0)) throw PreconditionException("SpherocylinderTraits: length >= 0");
        if (!(radius > 0)) throw PreconditionException("SpherocylinderTraits: radius > 0");
    }
    [[nodiscard]] const Interaction &getInteraction() const override { return *this; }
    [[nodiscard]] const ShapeGeometry &getGeometry() const override { return *this; }
    [[nodiscard]] bool hasHardPart() const override { return true; }
    [[nodiscard]] bool hasSoftPart() const override { return false; }
    [[nodiscard]] bool hasWallPart() const override { return true; }
    [[nodiscard]] bool isConvex() const override { return true; }
    [[nodiscard]] double getRangeRadius() const override { return 2 * this->radius + this->length; }
    [[nodiscard]] double getVolume() const override {
        return M_PI * this->radius * this->radius * this->length + 4. / 3 * M_PI * std::pow(this->radius, 3);
    }
    [[nodiscard]] rp_shape_desc descriptor() const override {
        return {RP_SHAPE_SPHEROCYLINDER, this->radius, this->length};
    }
};

// ---------------------------------------------------------------------------------------------- moves / box moves

class MoveSampler {
public:
    virtual ~MoveSampler() = default;
    [[nodiscard]] virtual std::string getName() const = 0;
    [[nodiscard]] virtual std::size_t getNumOfRequestedMoves(std::size_t numParticles) const = 0;
    [[nodiscard]] virtual std::vector<std::pair<std::string, double>> getStepSizes() const = 0;
    [[nodiscard]] virtual rp_move_desc descriptor() const = 0;
    virtual void updateStepSizes(double translation, double rotation) = 0;
};

class RototranslationSampler : public MoveSampler {
    double translationStepSize;
    std::optional<double> rotationStepSize;
    double maxTranslationStepSize;

public:
    RototranslationSampler(double translationStepSize, std::optional<double> rotationStepSize,
                           double maxTranslationStepSize = 0)
            : translationStepSize{translationStepSize}, rotationStepSize{rotationStepSize},
              maxTranslationStepSize{maxTranslationStepSize} {
        if (!(maxTranslationStepSize >= 0)) throw PreconditionException("maxTranslationStepSize >= 0");
        if (maxTranslationStepSize > 0 && translationStepSize > maxTranslationStepSize)
            throw PreconditionException("translationStepSize <= maxTranslationStepSize");
        if (!(translationStepSize > 0)) throw PreconditionException("translationStepSize > 0");
        if (rotationStepSize.has_value() && !(*rotationStepSize > 0))
            throw PreconditionException("rotationStepSize > 0");
    }
    [[nodiscard]] std::string getName() const override { return "rototranslation"; }
    [[nodiscard]] std::size_t getNumOfRequestedMoves(std::size_t numParticles) const override { return numParticles; }
    [[nodiscard]] std::vector<std::pair<std::string, double>> getStepSizes() const override {
        return {{"translation", this->translationStepSize}, {"rotation", this->rotationStepSize.value_or(0)}};
    }
    [[nodiscard]] rp_move_desc descriptor() const override {
        return {this->translationStepSize, this->rotationStepSize.value_or(0), this->maxTranslationStepSize,
                RP_MOVE_ROTOTRANSLATION};
    }
    void updateStepSizes(double translation, double rotation) override {
        this->translationStepSize = translation;
        this->rotationStepSize = rotation;
    }
};

// TranslationSampler (src/core/move_samplers/TranslationSampler.cpp:8-59)
class TranslationSampler : public MoveSampler {
    double translationStepSize;
    double maxTranslationStepSize;

public:
    explicit TranslationSampler(double translationStepSize, double maxTranslationStepSize = 0)
            : translationStepSize{translationStepSize}, maxTranslationStepSize{maxTranslationStepSize} {
        if (!(maxTranslationStepSize >= 0)) throw PreconditionException("maxTranslationStepSize >= 0");
        if (maxTranslationStepSize > 0 && translationStepSize > maxTranslationStepSize)
            throw PreconditionException("translationStepSize <= maxTranslationStepSize");
        if (!(translationStepSize > 0)) throw PreconditionException("translationStepSize > 0");
    }
    [[nodiscard]] std::string getName() const override { return "translation"; }
    [[nodiscard]] std::size_t getNumOfRequestedMoves(std::size_t numParticles) const override { return numParticles; }
    [[nodiscard]] std::vector<std::pair<std::string, double>> getStepSizes() const override {
        return {{"translation", this->translationStepSize}};
    }
    [[nodiscard]] rp_move_desc descriptor() const override {
        return {this->translationStepSize, 0, this->maxTranslationStepSize, RP_MOVE_TRANSLATION};
    }
    void updateStepSizes(double translation, double) override { this->translationStepSize = translation; }
};

// RotationSampler (src/core/move_samplers/RotationSampler.cpp:8-60)
class RotationSampler : public MoveSampler {
    double rotationStepSize;

public:
    explicit RotationSampler(double rotationStepSize) : rotationStepSize{rotationStepSize} {
        if (!(rotationStepSize > 0)) throw PreconditionException("rotationStepSize > 0");
    }
    [[nodiscard]] std::string getName() const override { return "rotation"; }
    [[nodiscard]] std::size_t getNumOfRequestedMoves(std::size_t numParticles) const override { return numParticles; }
    [[nodiscard]] std::vector<std::pair<std::string, double>> getStepSizes() const override {
        return {{"rotation", this->rotationStepSize}};
    }
    [[nodiscard]] rp_move_desc descriptor() const override { return {0, this->rotationStepSize, 0, RP_MOVE_ROTATION}; }
    void updateStepSizes(double, double rotation) override { this->rotationStepSize = rotation; }
};

class TriclinicBoxScaler {
protected:
    double stepSize{};

public:
    explicit TriclinicBoxScaler(double stepSize) : stepSize{stepSize} {
        if (!(stepSize > 0)) throw PreconditionException("TriclinicBoxScaler: stepSize > 0");
    }
    virtual ~TriclinicBoxScaler() = default;
    [[nodiscard]] double getStepSize() const { return this->stepSize; }
    void setStepSize(double s) {
        if (!(s > 0)) throw PreconditionException("TriclinicBoxScaler: stepSize > 0");
        this->stepSize = s;
    }
    [[nodiscard]] virtual rp_box_desc descriptor() const = 0;
};

// TriclinicDeltaScaler (src/core/volume_scalers/TriclinicDeltaScaler.cpp:7-25): all nine entries, or one box side
// per move when scaleTogether is false
class TriclinicDeltaScaler : public TriclinicBoxScaler {
    bool scaleTogether;

public:
    TriclinicDeltaScaler(double stepSize, bool scaleTogether)
            : TriclinicBoxScaler(stepSize), scaleTogether{scaleTogether} {}
    [[nodiscard]] rp_box_desc descriptor() const override {
        rp_box_desc d{RP_BOX_DELTA_TRICLINIC, this->stepSize};
        d.one_side = this->scaleTogether ? 0 : 1;
        return d;
    }
};

// VolumeScaler kinds accepted by TriclinicAdapter.
class VolumeScaler {
public:
    virtual ~VolumeScaler() = default;
    [[nodiscard]] virtual int kind() const = 0;
    virtual void describe(rp_box_desc &) const {}
};

// AnisotropicVolumeScaler (src/core/volume_scalers/AnisotropicVolumeScaler.cpp:13-31) with the log or the linear
// factor sampler: groups[a] is the scaling group of box side a (sides of one group share a factor); independent:
// one group per move instead of every group
class AnisotropicVolumeScaler : public VolumeScaler {
    int sampler;
    std::array<int, 3> groups;
    bool independent;

public:
    enum FactorSampler { LOG = RP_BOX_LOG_ISOTROPIC, LINEAR = RP_BOX_LINEAR };
    AnisotropicVolumeScaler(FactorSampler sampler, std::array<int, 3> groups, bool independent)
            : sampler{sampler}, groups{groups}, independent{independent} {
        for (int g : groups)
            if (g < 0 || g > 2) throw PreconditionException("AnisotropicVolumeScaler: groups are 0, 1 or 2");
    }
    [[nodiscard]] int kind() const override { return this->sampler; }
    void describe(rp_box_desc &d) const override {
        for (int a = 0; a < 3; a++) d.groups[a] = this->groups[a];
        d.independent = this->independent ? 1 : 0;
    }
};
class LogIsotropicVolumeScaler : public VolumeScaler {  // log(spec="isotropic"), AnisotropicVolumeScaler + LogScaling
public:
    [[nodiscard]] int kind() const override { return RP_BOX_LOG_ISOTROPIC; }
};
class DeltaVolumeScaler : public VolumeScaler {
public:
    [[nodiscard]] int kind() const override { return RP_BOX_DELTA_V; }
};

class TriclinicAdapter : public TriclinicBoxScaler {
    std::unique_ptr<VolumeScaler> volumeScaler;

public:
    TriclinicAdapter(std::unique_ptr<VolumeScaler> volumeScaler, double stepSize)
            : TriclinicBoxScaler(stepSize), volumeScaler{std::move(volumeScaler)} {}
    [[nodiscard]] rp_box_desc descriptor() const override {
        rp_box_desc d{this->volumeScaler->kind(), this->stepSize};
        this->volumeScaler->describe(d);
        return d;
    }
};

// ---------------------------------------------------------------------------------------------- packing

struct Shape {
    std::array<double, 3> position{};
    std::array<double, 9> orientation{1, 0, 0, 0, 1, 0, 0, 0, 1};  // row-major
};

// Box as a row-major matrix whose columns are the box sides (TriclinicBox.h:49-58).
using Box = std::array<double, 9>;

class Packing {
    Box box;
    std::vector<Shape> shapes;
    rp_shape_desc shape;
    rp_ctx *ctx = nullptr;
    uint64_t seed = 0;
    bool stale = false;

    void upload(uint64_t seed_) {
        this->seed = seed_;
        std::vector<double> pos(3 * shapes.size()), rot(9 * shapes.size());
        for (std::size_t i = 0; i < shapes.size(); i++) {
            for (int k = 0; k < 3; k++) pos[3 * i + k] = shapes[i].position[k];
            for (int k = 0; k < 9; k++) rot[9 * i + k] = shapes[i].orientation[k];
        }
        rp_ctx *c = nullptr;
        check(rp_create(&c, &shape, box.data(), shapes.size(), pos.data(), rot.data(), this->seed, RP_PREC_FP64));
        if (ctx) rp_destroy(ctx);
        ctx = c;
    }

public:
    Packing(const Box &box, std::vector<Shape> shapes, const Interaction &interaction)
            : box{box}, shapes{std::move(shapes)}, shape{interaction.descriptor()} {
        if (this->shapes.empty()) throw PreconditionException("Packing: no shapes");
        this->upload(0);
    }
    Packing(const Packing &) = delete;
    Packing &operator=(const Packing &) = delete;
    ~Packing() {
        if (ctx) rp_destroy(ctx);
    }

    // Re-keys the device random streams (Simulation's seed, Simulation.cpp:66-68).
    void reseed(uint64_t seed) {
        this->refresh();
        this->upload(seed);
    }
    void refresh() const {
        if (!stale) return;
        auto *self = const_cast<Packing *>(this);
        std::vector<double> pos(3 * shapes.size()), rot(9 * shapes.size());
        check(rp_download(ctx, self->box.data(), pos.data(), rot.data()));
        for (std::size_t i = 0; i < shapes.size(); i++) {
            for (int k = 0; k < 3; k++) self->shapes[i].position[k] = pos[3 * i + k];
            for (int k = 0; k < 9; k++) self->shapes[i].orientation[k] = rot[9 * i + k];
        }
        self->stale = false;
    }
    // Packing::reset (src/core/Packing.cpp:67-75): a new state for the same shape, e.g. a trajectory snapshot
    void reset(std::vector<Shape> newShapes, const Box &newBox) {
        if (newShapes.empty()) throw PreconditionException("Packing: no shapes");
        this->shapes = std::move(newShapes);
        this->box = newBox;
        this->stale = false;
        this->upload(this->seed);
    }
    void markStale() { stale = true; }
    [[nodiscard]] rp_ctx *context() const { return ctx; }
    [[nodiscard]] std::size_t size() const { return shapes.size(); }
    [[nodiscard]] const Shape &operator[](std::size_t i) const {
        if (i >= shapes.size()) throw PreconditionException("Packing: index out of range");
        refresh();
        return shapes[i];
    }
    [[nodiscard]] const Box &getBox() const {
        refresh();
        return box;
    }
    [[nodiscard]] double getVolume() const {
        const Box &b = getBox();
        double det = b[0] * b[4] * b[8] + b[1] * b[5] * b[6] + b[2] * b[3] * b[7] - b[2] * b[4] * b[6] -
                     b[1] * b[3] * b[8] - b[0] * b[5] * b[7];
        return std::abs(det);
    }
    [[nodiscard]] double getPackingFraction(double shapeVolume) const {
        return (double)this->size() / this->getVolume() * shapeVolume;
    }
    // Packing::tryScaling(newBox, interaction) (Packing.cpp:172-208): +inf on an overlap, else the energy change;
    // the scaled state is current until revertScaling
    double tryScaling(const Box &newBox, const Interaction &) {
        double dE = 0;
        check(rp_try_scaling(ctx, newBox.data(), &dE));
        stale = true;
        return dE;
    }
    // Packing::revertScaling (Packing.cpp:363-371)
    void revertScaling() {
        check(rp_revert_scaling(ctx));
        stale = true;
    }
    // Packing::countTotalOverlaps (Packing.cpp:519-559)
    [[nodiscard]] std::size_t countTotalOverlaps(const Interaction &, bool earlyExit = true) const {
        uint64_t n = 0;
        check(rp_count_total_overlaps(ctx, earlyExit ? 1 : 0, &n));
        return (std::size_t)n;
    }
};

// ---------------------------------------------------------------------------------------------- trajectories

// SimulationRecorder (src/core/SimulationRecorder.h:15-33)
class SimulationRecorder {
public:
    virtual ~SimulationRecorder() = default;
    virtual void recordSnapshot(const Packing &packing, std::size_t cycle) = 0;
    [[nodiscard]] virtual std::size_t getLastCycleNumber() const = 0;
    virtual void close() = 0;
};

// RamtrjRecorder (src/core/io/RamtrjRecorder.cpp:8-66): snapshots of the device-resident state in RAMTRJ format
class RamtrjRecorder : public SimulationRecorder {
    rp_ramtrj_recorder *rec = nullptr;

public:
    RamtrjRecorder(const std::string &path, std::size_t numParticles, std::size_t cycleStep, bool append) {
        check(rp_ramtrj_recorder_open(&rec, path.c_str(), numParticles, cycleStep, append ? 1 : 0));
    }
    RamtrjRecorder(const RamtrjRecorder &) = delete;
    RamtrjRecorder &operator=(const RamtrjRecorder &) = delete;
    ~RamtrjRecorder() override {
        if (rec) rp_ramtrj_recorder_close(rec);
    }
    void recordSnapshot(const Packing &packing, std::size_t cycle) override {
        if (!rec) throw PreconditionException("RAMTRJ recorder is closed");
        check(rp_ramtrj_record_packing(rec, packing.context(), cycle));
    }
    [[nodiscard]] std::size_t getLastCycleNumber() const override {
        std::size_t c = 0;
        if (rec) check(rp_ramtrj_recorder_last_cycle(rec, &c));
        return c;
    }
    void close() override {
        if (!rec) return;
        rp_ramtrj_recorder *r = rec;
        rec = nullptr;
        check(rp_ramtrj_recorder_close(r));
    }
};

// RamtrjPlayer (src/core/io/RamtrjPlayer.cpp:11-128); AutoFix as RamtrjPlayer::AutoFix (RamtrjPlayer.h:29-77)
class RamtrjPlayer {
    rp_ramtrj_player *player = nullptr;
    rp_ramtrj_header header{};

    void open(const std::string &path, std::size_t expected, rp_ramtrj_fix *fix) {
        check(rp_ramtrj_player_open(&player, path.c_str(), expected, fix));
        check(rp_ramtrj_player_header(player, &header));
    }
    void load(Packing &packing, int (*read)(RamtrjPlayer &, Box &, std::vector<double> &, std::vector<double> &)) {
        if (packing.size() != header.num_particles) throw PreconditionException("RAMTRJ player: packing size");
        Box box{};
        std::vector<double> pos(3 * packing.size()), rot(9 * packing.size());
        check(read(*this, box, pos, rot));
        std::vector<Shape> shapes(packing.size());
        for (std::size_t i = 0; i < shapes.size(); i++) {
            for (int k = 0; k < 3; k++) shapes[i].position[k] = pos[3 * i + k];
            for (int k = 0; k < 9; k++) shapes[i].orientation[k] = rot[9 * i + k];
        }
        packing.reset(std::move(shapes), box);
        check(rp_ramtrj_player_header(player, &header));
    }

public:
    class AutoFix {
        std::size_t expected;
        rp_ramtrj_fix fix{};
        friend class RamtrjPlayer;

    public:
        explicit AutoFix(std::size_t expectedNumMolecules) : expected{expectedNumMolecules} {
            if (expected == 0) throw PreconditionException("AutoFix: expected number of molecules must be > 0");
        }
        [[nodiscard]] bool wasFixingSuccessful() const { return fix.fixing_successful != 0; }
        [[nodiscard]] bool wasFixingNeeded() const { return fix.fixing_needed != 0; }
        [[nodiscard]] std::size_t getInferredSnapshots() const { return fix.inferred_snapshots; }
        [[nodiscard]] std::size_t getBytesRemainder() const { return fix.bytes_remainder; }
    };

    explicit RamtrjPlayer(const std::string &path) { open(path, 0, nullptr); }
    RamtrjPlayer(const std::string &path, AutoFix &autoFix) { open(path, autoFix.expected, &autoFix.fix); }
    RamtrjPlayer(const RamtrjPlayer &) = delete;
    RamtrjPlayer &operator=(const RamtrjPlayer &) = delete;
    ~RamtrjPlayer() { close(); }

    [[nodiscard]] bool hasNext() const { return player && header.current_snapshot < header.num_snapshots; }
    void nextSnapshot(Packing &packing) {
        if (!player) throw PreconditionException("RAMTRJ player is closed");
        load(packing, [](RamtrjPlayer &p, Box &b, std::vector<double> &x, std::vector<double> &r) {
            return rp_ramtrj_next_snapshot(p.player, b.data(), x.size() / 3, x.data(), r.data());
        });
    }
    void jumpToSnapshot(Packing &packing, std::size_t cycleNumber) {
        if (!player) throw PreconditionException("RAMTRJ player is closed");
        this->jumpCycle = cycleNumber;
        load(packing, [](RamtrjPlayer &p, Box &b, std::vector<double> &x, std::vector<double> &r) {
            return rp_ramtrj_jump_to_snapshot(p.player, p.jumpCycle, b.data(), x.size() / 3, x.data(), r.data());
        });
    }
    void lastSnapshot(Packing &packing) { jumpToSnapshot(packing, header.num_snapshots * header.cycle_step); }
    [[nodiscard]] std::size_t getCurrentSnapshotCycles() const { return header.current_snapshot * header.cycle_step; }
    [[nodiscard]] std::size_t getTotalCycles() const { return header.cycle_step * header.num_snapshots; }
    [[nodiscard]] std::size_t getCycleStep() const { return header.cycle_step; }
    [[nodiscard]] std::size_t getNumMolecules() const { return header.num_particles; }
    void close() {
        if (player) rp_ramtrj_player_close(player);
        player = nullptr;
    }

private:
    std::size_t jumpCycle = 0;
};

// ---------------------------------------------------------------------------------------------- simulation

class Simulation {
    std::unique_ptr<Packing> packing;
    std::vector<std::unique_ptr<MoveSampler>> moveSamplers;
    std::unique_ptr<TriclinicBoxScaler> boxScaler;
    rp_counters counters{};
    double temperature = 1, pressure = 0;

public:
    Simulation(std::unique_ptr<Packing> packing, std::vector<std::unique_ptr<MoveSampler>> moveSamplers,
               unsigned long seed, std::unique_ptr<TriclinicBoxScaler> boxScaler)
            : packing{std::move(packing)}, moveSamplers{std::move(moveSamplers)}, boxScaler{std::move(boxScaler)} {
        if (!this->packing) throw PreconditionException("Simulation: no packing");
        if (this->moveSamplers.empty() || this->moveSamplers.size() > RP_MAX_MOVE_SAMPLERS)
            throw PreconditionException("Simulation: 1 to " + std::to_string(RP_MAX_MOVE_SAMPLERS) + " move samplers");
        for (const auto &sampler : this->moveSamplers)
            if (!sampler) throw PreconditionException("Simulation: null move sampler");
        this->packing->reseed(seed);
    }

    // Simulation::integrate (Simulation.cpp:98-199): thermalisation with step adaptation, then averaging; the
    // observer is called every `averagingEvery` averaging cycles with the current packing.
    // With snapshotEvery > 0 every recorder takes a snapshot whenever the cycle count is a multiple of it, in both
    // phases (Simulation.cpp:146-150, 175-179).
    void integrate(double temperature_, double pressure_, std::size_t thermalisationCycles,
                   std::size_t averagingCycles, std::size_t averagingEvery,
                   const std::function<void(const Packing &, std::size_t)> &observer = {},
                   std::size_t snapshotEvery = 0,
                   const std::vector<std::unique_ptr<SimulationRecorder>> &recorders = {}) {
        if (thermalisationCycles == 0 && averagingCycles == 0) throw PreconditionException("no cycles");
        if (averagingCycles > 0 && (averagingEvery == 0 || averagingEvery > averagingCycles))
            throw PreconditionException("averagingEvery");
        this->temperature = temperature_;
        this->pressure = pressure_;
        std::vector<rp_move_desc> md;
        for (const auto &sampler : this->moveSamplers) md.push_back(sampler->descriptor());
        rp_box_desc bd{RP_BOX_NONE, 0};
        if (this->boxScaler) bd = this->boxScaler->descriptor();
        check(rp_set_environment(this->packing->context(), temperature_, pressure_, md.data(), &bd));
        check(rp_set_move_samplers(this->packing->context(), md.data(), md.size()));
        auto advance = [&](std::size_t cycles, bool adapt) {
            while (cycles > 0) {
                std::size_t chunk = cycles;
                if (snapshotEvery) chunk = std::min(chunk, snapshotEvery - this->counters.cycles % snapshotEvery);
                this->run(chunk, adapt);
                cycles -= chunk;
                if (snapshotEvery && this->counters.cycles % snapshotEvery == 0)
                    for (const auto &recorder : recorders) recorder->recordSnapshot(*this->packing, this->counters.cycles);
            }
        };
        if (thermalisationCycles > 0)
            advance(thermalisationCycles, true);
        for (std::size_t done = 0; done < averagingCycles; done += averagingEvery) {
            std::size_t chunk = std::min(averagingEvery, averagingCycles - done);
            advance(chunk, false);
            if (observer && chunk == averagingEvery) observer(*this->packing, this->counters.cycles);
        }
    }

    [[nodiscard]] const Packing &getPacking() const { return *this->packing; }
    [[nodiscard]] const rp_counters &getCounters() const { return this->counters; }
    [[nodiscard]] double getMoveAcceptanceRate() const {
        return this->counters.moves_attempted ? (double)this->counters.moves_accepted / this->counters.moves_attempted
                                              : 0;
    }

private:
    void run(std::size_t cycles, bool adapt) {
        check(rp_run_cycles(this->packing->context(), cycles, adapt ? 1 : 0, &this->counters));
        this->packing->markStale();
        std::vector<rp_move_desc> steps(this->moveSamplers.size());
        check(rp_get_move_samplers(this->packing->context(), steps.data(), steps.size(), nullptr));
        for (std::size_t i = 0; i < steps.size(); i++)
            this->moveSamplers[i]->updateStepSizes(steps[i].trans_step, steps[i].rot_step);
        if (this->boxScaler) this->boxScaler->setStepSize(this->counters.box_step);
    }
};

}  // namespace rampack_amd
