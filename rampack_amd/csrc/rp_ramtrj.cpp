// RAMTRJ trajectories (rampack's binary trajectory format, docs/output-formats.md): host-side recorder and player
// behind the C ABI.
//
// Layout (RamtrjIO.cpp:16-105, all numbers little-endian as the x86-64 host writes them):
//   header, 33 bytes: "RAMTRJ\n", minor version byte, major version byte (1.1), then three 64-bit unsigned
//     numbers: particles, snapshots, cycle step;
//   each snapshot: the box matrix as 9 doubles (row-major), then per particle 3 doubles of position and the 3
//     extrinsic XYZ Tait-Bryan angles (R = Rz Ry Rx) of its orientation.
// The snapshot count in the header is written when the recorder is closed; a file whose recorder never closed
// holds 0 there and is read back with the auto-fix, which infers the count from the file size
// (RamtrjPlayer.cpp:22-44, 165-189).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "rampack_amd.h"

namespace rp {
int set_error(int code, const std::string &msg);  // rp_capi.cpp
}

namespace {

constexpr char MAGIC[7] = {'R', 'A', 'M', 'T', 'R', 'J', '\n'};
constexpr uint8_t VERSION_MAJOR = 1, VERSION_MINOR = 1;
constexpr size_t HEADER_BYTES = 7 + 1 + 1 + 3 * 8;
constexpr size_t BOX_BYTES = 9 * sizeof(double);
constexpr size_t PARTICLE_BYTES = 6 * sizeof(double);
// EulerAngles::EPSILON (src/geometry/EulerAngles.h:30)
constexpr double EULER_EPS = 1e-12;

struct Header {
    uint8_t major = VERSION_MAJOR, minor = VERSION_MINOR;
    uint64_t particles = 0, snapshots = 0, cycleStep = 0;
};

size_t snapshotBytes(const Header &h) { return BOX_BYTES + (size_t)h.particles * PARTICLE_BYTES; }

int invalid(const std::string &what) { return rp::set_error(RP_ERR_VALIDATION, what); }
int precondition(const std::string &what) { return rp::set_error(RP_ERR_PRECONDITION, what); }

void putHeader(unsigned char *b, const Header &h) {
    std::memcpy(b, MAGIC, 7);
    b[7] = h.minor;  // the minor version precedes the major one (RamtrjIO.cpp:40-47)
    b[8] = h.major;
    std::memcpy(b + 9, &h.particles, 8);
    std::memcpy(b + 17, &h.snapshots, 8);
    std::memcpy(b + 25, &h.cycleStep, 8);
}

// RamtrjIO::readHeader (RamtrjIO.cpp:16-37): versions up to 1.1; from 1.1 on, particles and cycle step are > 0
int readHeader(std::istream &in, Header &h) {
    unsigned char b[HEADER_BYTES];
    in.read(reinterpret_cast<char *>(b), 7);
    if (!in || std::memcmp(b, MAGIC, 7) != 0) return invalid("RAMTRJ read error: magic");
    in.read(reinterpret_cast<char *>(b + 7), 2);
    h.minor = b[7];
    h.major = b[8];
    if (!in || h.major != 1 || h.minor > 1) return invalid("RAMTRJ: only versions up to 1.1 are supported");
    const char *what[3] = {"RAMTRJ read error: num particles", "RAMTRJ read error: num snapshots",
                           "RAMTRJ read error: cycle step"};
    uint64_t *field[3] = {&h.particles, &h.snapshots, &h.cycleStep};
    for (int i = 0; i < 3; i++) {
        in.read(reinterpret_cast<char *>(field[i]), 8);
        if (!in) return invalid(what[i]);
    }
    if (h.minor >= 1) {
        if (h.particles == 0) return invalid(what[0]);
        if (h.cycleStep == 0) return invalid(what[2]);
    }
    return RP_OK;
}

// (R R^T - I) entries below 1e-12 in magnitude (EulerAngles::isRotationMatrix, src/geometry/EulerAngles.cpp:58-65),
// with the products summed as Matrix's operator* does (src/geometry/Matrix.tpp:105-113)
bool isRotation(const double *R) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += R[3 * i + k] * R[3 * j + k];
            if (std::abs(s - (i == j ? 1.0 : 0.0)) >= EULER_EPS) return false;
        }
    return true;
}

// The first of the two angle triples of EulerAngles::forMatrix (src/geometry/EulerAngles.cpp:12-45), the one
// RamtrjIO::writeShape stores.  The reference guards its general branch with `R31 > -1 + eps || R31 < 1 - eps`,
// which every number satisfies, so the gimbal-lock branch is reached only by a NaN; both are kept so that the
// same orientation always gives the same bytes.
void eulerAngles(const double *R, double out[3]) {
    double R11 = R[0], R12 = R[1], R13 = R[2], R21 = R[3], R31 = R[6], R32 = R[7], R33 = R[8];
    if (R31 > -1 + EULER_EPS || R31 < 1 - EULER_EPS) {
        double theta = -std::asin(R31);
        double c = std::cos(theta);
        out[0] = std::atan2(R32 / c, R33 / c);
        out[1] = theta;
        out[2] = std::atan2(R21 / c, R11 / c);
    } else if (R31 < 0) {
        out[0] = std::atan2(R12, R13);
        out[1] = M_PI / 2;
        out[2] = 0;
    } else {
        out[0] = std::atan2(-R12, -R13);
        out[1] = -M_PI / 2;
        out[2] = 0;
    }
}

// Matrix<3,3>::rotation(ax, ay, az) = Rz(az) Ry(ay) Rx(ax) (src/geometry/Matrix.tpp:408-433), entry by entry.
// gcc compiles the sin and cos of one angle there into a single sincos call, and glibc's sincos can differ from
// its separate sin and cos in the last bit, so sincos is called here too: the matrices are then the bits the
// reference's gcc build reads.
void rotationFromAngles(const double a[3], double *R) {
    double sx, sy, sz, cx, cy, cz;
    sincos(a[0], &sx, &cx);
    sincos(a[1], &sy, &cy);
    sincos(a[2], &sz, &cz);
    R[0] = cy * cz;
    R[1] = sx * sy * cz - cx * sz;
    R[2] = cx * sy * cz + sx * sz;
    R[3] = cy * sz;
    R[4] = sx * sy * sz + cx * cz;
    R[5] = cx * sy * sz - sx * cz;
    R[6] = -sy;
    R[7] = sx * cy;
    R[8] = cx * cy;
}

double det3(const double *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

}  // namespace

struct rp_ramtrj_recorder {
    std::fstream f;
    Header h;
};

struct rp_ramtrj_player {
    std::ifstream f;
    Header h;
    uint64_t current = 0;  // snapshots read since the start or the last jump target
};

extern "C" {

int rp_ramtrj_recorder_open(rp_ramtrj_recorder **out, const char *path, size_t num_particles, size_t cycle_step,
                            int append) {
    if (!out || !path) return precondition("null argument");
    *out = nullptr;
    // RamtrjRecorder::RamtrjRecorder (src/core/io/RamtrjRecorder.cpp:8-36)
    if (num_particles == 0) return precondition("RAMTRJ recorder: the number of particles must be > 0");
    if (cycle_step == 0) return precondition("RAMTRJ recorder: the cycle step must be > 0");
    auto *r = new rp_ramtrj_recorder;
    r->h.particles = num_particles;
    r->h.cycleStep = cycle_step;
    // SimulationRecorderFactory.cpp:16-28: a continuation opens the file for update, a new recording truncates it
    auto mode = std::ios::in | std::ios::out | std::ios::binary;
    r->f.open(path, append ? mode : mode | std::ios::trunc);
    if (!r->f) {
        delete r;
        return invalid(std::string("Could not open ") + path + " to store RAMTRJ trajectory");
    }
    if (append) {
        Header old;
        r->f.seekg(0);
        int e = readHeader(r->f, old);
        if (!e && (old.particles != num_particles || old.cycleStep != cycle_step))
            e = invalid("RAMTRJ append error: unmatching number of molecules and/or cycle step");
        if (!e) {
            r->h.snapshots = old.snapshots;
            r->f.seekp(0, std::ios::end);
            if ((uint64_t)r->f.tellp() != HEADER_BYTES + old.snapshots * snapshotBytes(old))
                e = invalid("RAMTRJ append error: broken snapshot structure");
        }
        if (e) {
            delete r;
            return e;
        }
    } else {
        unsigned char b[HEADER_BYTES];
        putHeader(b, r->h);
        r->f.write(reinterpret_cast<char *>(b), HEADER_BYTES);
        if (!r->f) {
            delete r;
            return invalid("RAMTRJ write error: header");
        }
    }
    *out = r;
    return RP_OK;
}

// RamtrjRecorder::recordSnapshot (src/core/io/RamtrjRecorder.cpp:42-53) with RamtrjIO::writeBox/writeShape
// (RamtrjIO.cpp:57-83)
int rp_ramtrj_record(rp_ramtrj_recorder *r, const double box[9], size_t n, const double *pos, const double *rot,
                     size_t cycle) {
    if (!r || !box || (n && (!pos || !rot))) return precondition("null argument");
    if (!r->f.is_open()) return precondition("RAMTRJ recorder is closed");
    if (cycle == 0 || cycle != (r->h.snapshots + 1) * r->h.cycleStep)
        return precondition("RAMTRJ recorder: expected cycle " + std::to_string((r->h.snapshots + 1) * r->h.cycleStep) +
                            ", got " + std::to_string(cycle));
    if (n != r->h.particles)
        return precondition("RAMTRJ recorder: " + std::to_string(n) + " particles, the trajectory holds " +
                            std::to_string(r->h.particles));
    std::vector<double> buf(9 + 6 * n);
    std::memcpy(buf.data(), box, BOX_BYTES);
    for (size_t i = 0; i < n; i++) {
        const double *R = rot + 9 * i;
        // EulerAngles::EulerAngles expects a rotation matrix (EulerAngles.cpp:47-53)
        if (!isRotation(R))
            return precondition("RAMTRJ recorder: orientation of particle " + std::to_string(i) +
                                " is not a rotation matrix");
        double *o = buf.data() + 9 + 6 * i;
        o[0] = pos[3 * i];
        o[1] = pos[3 * i + 1];
        o[2] = pos[3 * i + 2];
        eulerAngles(R, o + 3);
    }
    r->f.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
    if (!r->f) return invalid("RAMTRJ write error: snapshot data");
    r->h.snapshots++;
    return RP_OK;
}

int rp_ramtrj_recorder_last_cycle(rp_ramtrj_recorder *r, size_t *cycle) {
    if (!r || !cycle) return precondition("null argument");
    *cycle = r->h.snapshots * r->h.cycleStep;  // RamtrjRecorder.h:47
    return RP_OK;
}

// RamtrjRecorder::close0 (src/core/io/RamtrjRecorder.cpp:55-66): the header is rewritten with the snapshot count
int rp_ramtrj_recorder_close(rp_ramtrj_recorder *r) {
    if (!r) return RP_OK;
    int e = RP_OK;
    if (r->f.is_open()) {
        unsigned char b[HEADER_BYTES];
        putHeader(b, r->h);
        r->f.seekp(0);
        r->f.write(reinterpret_cast<char *>(b), HEADER_BYTES);
        r->f.close();
        if (!r->f) e = invalid("RAMTRJ write error: header");
    }
    delete r;
    return e;
}

// RamtrjPlayer::RamtrjPlayer (src/core/io/RamtrjPlayer.cpp:11-44) and IO::loadRamtrjPlayer (src/frontend/IO.cpp:
// 32-64): expected_particles 0 reads strictly (a file whose size disagrees with its header is an error); otherwise
// a truncated file is repaired in memory as RamtrjPlayer::AutoFix does and *fix reports what was done.
int rp_ramtrj_player_open(rp_ramtrj_player **out, const char *path, size_t expected_particles, rp_ramtrj_fix *fix) {
    if (!out || !path) return precondition("null argument");
    *out = nullptr;
    if (fix) std::memset(fix, 0, sizeof *fix);
    auto *p = new rp_ramtrj_player;
    p->f.open(path, std::ios::in | std::ios::binary);
    if (!p->f) {
        delete p;
        return invalid(std::string("Could not open ") + path + " to read trajectory");
    }
    auto fail = [&](int e) {  // *fix stays zeroed: not fixed, not successful (AutoFix::reportError)
        delete p;
        return e;
    };
    int e = readHeader(p->f, p->h);
    if (e) return fail(e);
    p->f.seekg(0, std::ios::end);
    uint64_t size = (uint64_t)p->f.tellg();
    uint64_t expected = HEADER_BYTES + p->h.snapshots * snapshotBytes(p->h);
    if (expected_particles == 0) {
        if (size != expected) return fail(invalid("RAMTRJ read error: broken snapshot structure"));
    } else if (size == expected) {
        if (fix) {  // AutoFix::reportNofix (RamtrjPlayer.cpp:151-158)
            fix->fixing_successful = 1;
            fix->header_snapshots = fix->inferred_snapshots = p->h.snapshots;
            fix->bytes_per_snapshot = snapshotBytes(p->h);
        }
    } else {
        // AutoFix::tryFixing (RamtrjPlayer.cpp:166-189)
        if (p->h.particles == 0 || p->h.cycleStep == 0)
            return fail(invalid("The header does not contain information about the number of particles or cycle step"
                                " - it was not always stored in RAMTRJ v1.0."));
        if (p->h.particles != expected_particles)
            return fail(invalid("Expected number of molecules (" + std::to_string(expected_particles) +
                                ") != number of molecules in the input (" + std::to_string(p->h.particles) + ")"));
        uint64_t per = snapshotBytes(p->h), body = size - HEADER_BYTES;
        if (fix) {
            fix->fixing_needed = 1;
            fix->fixing_successful = 1;
            fix->header_snapshots = p->h.snapshots;
            fix->inferred_snapshots = body / per;
            fix->bytes_remainder = body % per;
            fix->bytes_per_snapshot = per;
        }
        p->h.snapshots = body / per;
    }
    p->f.seekg(HEADER_BYTES);  // RamtrjPlayer::reset (RamtrjPlayer.cpp:124-128)
    p->current = 0;
    *out = p;
    return RP_OK;
}

int rp_ramtrj_player_header(rp_ramtrj_player *p, rp_ramtrj_header *out) {
    if (!p || !out) return precondition("null argument");
    out->version_major = p->h.major;
    out->version_minor = p->h.minor;
    out->num_particles = p->h.particles;
    out->num_snapshots = p->h.snapshots;
    out->cycle_step = p->h.cycleStep;
    out->current_snapshot = p->current;
    return RP_OK;
}

// RamtrjPlayer::nextSnapshot (src/core/io/RamtrjPlayer.cpp:54-68) with RamtrjIO::readBox/readShape
// (RamtrjIO.cpp:50-72)
int rp_ramtrj_next_snapshot(rp_ramtrj_player *p, double box[9], size_t n, double *pos, double *rot) {
    if (!p || !box || (n && (!pos || !rot))) return precondition("null argument");
    if (p->current >= p->h.snapshots) return precondition("RAMTRJ player: no more snapshots");
    if (n != p->h.particles)
        return precondition("RAMTRJ player: " + std::to_string(n) + " particles, the trajectory holds " +
                            std::to_string(p->h.particles));
    std::vector<double> buf(9 + 6 * n);
    p->f.read(reinterpret_cast<char *>(buf.data()), BOX_BYTES);
    if (!p->f) return invalid("RAMTRJ read error: snapshot box data");
    p->f.read(reinterpret_cast<char *>(buf.data() + 9), (std::streamsize)(n * PARTICLE_BYTES));
    if (!p->f) return invalid("RAMTRJ read error: snapshot particle data");
    // TriclinicBox inverts the box matrix, which fails for a singular one (src/geometry/Matrix.tpp)
    if (det3(buf.data()) == 0) return invalid("RAMTRJ read error: zero box volume");
    std::memcpy(box, buf.data(), BOX_BYTES);
    for (size_t i = 0; i < n; i++) {
        const double *o = buf.data() + 9 + 6 * i;
        pos[3 * i] = o[0];
        pos[3 * i + 1] = o[1];
        pos[3 * i + 2] = o[2];
        rotationFromAngles(o + 3, rot + 9 * i);
    }
    p->current++;
    return RP_OK;
}

// RamtrjPlayer::jumpToSnapshot (src/core/io/RamtrjPlayer.cpp:74-89): the snapshot recorded at `cycle`
int rp_ramtrj_jump_to_snapshot(rp_ramtrj_player *p, size_t cycle, double box[9], size_t n, double *pos, double *rot) {
    if (!p) return precondition("null argument");
    if (p->h.snapshots == 0) return precondition("RAMTRJ player: the trajectory is empty");
    if (cycle % p->h.cycleStep != 0)
        return precondition("RAMTRJ player: cycle " + std::to_string(cycle) + " is not a multiple of the cycle step " +
                            std::to_string(p->h.cycleStep));
    uint64_t s = cycle / p->h.cycleStep;
    if (s == 0 || s > p->h.snapshots)
        return precondition("RAMTRJ player: no snapshot for cycle " + std::to_string(cycle));
    p->current = s - 1;
    p->f.clear();
    p->f.seekg((std::streamoff)(HEADER_BYTES + p->current * snapshotBytes(p->h)));
    return rp_ramtrj_next_snapshot(p, box, n, pos, rot);
}

void rp_ramtrj_player_close(rp_ramtrj_player *p) { delete p; }

}  // extern "C"
