// rp_capi.cpp -- host side of the C ABI (include/rampack_amd.h): device buffers, the cycle driver, box moves and
// step adaptation.  The reference's Simulation/Packing orchestration (src/core/Simulation.cpp:325-607,
// src/core/Packing.cpp:172-208, 858-899) is restated here on top of the gfx950 kernels of rp_kernels.hip.
#include "rampack_amd.h"
#include "rp_kernels.h"
#include "rp_mlp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <limits>
#include <cstring>
#include <string>
#include <vector>

using namespace rp;

namespace {

thread_local std::string g_lastError;

int fail(int code, const std::string &msg) {
    g_lastError = msg;
    return code;
}

#define RP_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(RP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HostBox {
    double dims[9], inv[9];
    void set(const double *d) {
        std::memcpy(dims, d, sizeof dims);
        inverse3(dims, inv);
    }
    // TriclinicBox::getSides / getVolume / getHeights (TriclinicBox.cpp:27-51)
    V3 side(int i) const { return mulv(dims, v3(i == 0, i == 1, i == 2)); }
    double volume() const { return std::fabs(dot(cross(side(0), side(1)), side(2))); }
    void heights(double h[3]) const {
        V3 s0 = side(0), s1 = side(1), s2 = side(2);
        double vol = volume();
        V3 c12 = cross(s1, s2), c20 = cross(s2, s0), c01 = cross(s0, s1);
        h[0] = vol / std::sqrt(dot(c12, c12));
        h[1] = vol / std::sqrt(dot(c20, c20));
        h[2] = vol / std::sqrt(dot(c01, c01));
    }
};

// Packing::isBoxUpscaled (Packing.cpp:1155-1176)
bool isBoxUpscaled(const HostBox &o, const HostBox &nb) {
    const double EPS = 100 * 2.220446049250313e-16;
    for (int i = 0; i < 3; i++) {
        V3 os = o.side(i), ns = nb.side(i);
        double d = dot(os, ns), n2o = dot(os, os), n2n = dot(ns, ns);
        if (std::fabs(d * d / n2o / n2n - 1) > EPS) return false;
        if (d * d / n2o / n2o < 1 + EPS) return false;
    }
    return true;
}

struct DeviceParticles {
    Particles p{};
    void *mem = nullptr;
    size_t bytes = 0;  // the whole block; the arrays sit at the same offsets in every block of n particles
};

}  // namespace

struct rp_ctx {
    hipStream_t stream = nullptr;
    rp_shape_desc shape{};
    PairConsts k{};
    double range = 0;
    size_t n = 0;
    HostBox box;
    DeviceParticles cur, alt;
    int *cellStart = nullptr, *cellCount = nullptr;
    int ncellCap = 0;
    int *tmpSlot = nullptr, *tmpOrder = nullptr;
    double *tmpPos = nullptr;
    Proposals prop{};
    size_t propCap = 0;
    void *propMem = nullptr;
    double *obsMem = nullptr;            // Q-tensor partial sums
    double *cellEnergy = nullptr;        // per-cell soft energy partials
    int cellEnergyCap = 0;
    double *sumScratch = nullptr;        // fixed-order sum of the partials
    unsigned long long *dFixCounters = nullptr;  // counters of the orientation renormalisation pass
    unsigned long long *dSlots = nullptr;        // move-count slots of the sweep, SLOT_WORDS per grid cell
    int *scanTmp = nullptr;                      // tile sums of the rebuild's scan on large grids
    size_t scanTmpCap = 0;
    size_t slotsCap = 0;                         // cells the slots cover
    int slotsDirty = 0;                          // cells of the grid whose slots hold counts not yet folded (0: none)
    int scanHint = 0;                            // a cell where the last overlap count found an overlap
    uint64_t fixEvery = 10000;           // cycles between orientation renormalisations (0: never)
    uint64_t fixRejected = 0;            // renormalisations skipped because they would create an overlap
    bool counting = false;               // overlap-counting mode (Packing::toggleOverlapCounting)
    long long numOverlaps = 0;           // cached overlap count in counting mode (Packing::getCachedNumberOfOverlaps)
    unsigned long long *dCounters = nullptr;
    unsigned long long *dPhases = nullptr;  // sweep phase timers (RAMPACK_SWEEP_PHASES=1)
    int device = 0;                         // the HIP device the context was created on
    int sweepTuning = 0;                    // grouped sweep wave priorities (RAMPACK_SWEEP_TUNING)
    std::string tracePath;                  // per-workgroup trace of one sweep launch (RAMPACK_SWEEP_TRACE)
    uint64_t traceLaunch = 0, sweepLaunches = 0;
    unsigned long long *hCounters = nullptr;  // pinned
    uint64_t seed = 0;
    uint64_t cycle = 0;
    int precision = RP_PREC_FP64;        // RP_PREC_*: the stencil scan of the default hard sweep
    Philox hostRng{};

    // environment
    bool envSet = false;
    double temperature = 1, pressure = 0;
    // move samplers (Simulation.h:113-165 keeps a vector of them): kind, steps, maximum translation step
    struct Sampler {
        int kind = RP_MOVE_ROTOTRANSLATION;
        double transStep = 0.1, rotStep = 0.1, maxTransStep = 0;
        bool maxTransImplicit = true;
    };
    std::vector<Sampler> samplers;
    uint64_t typeAccBase[RP_MAX_MOVE_SAMPLERS] = {}, typeAttBase[RP_MAX_MOVE_SAMPLERS] = {};
    int boxKind = RP_BOX_NONE;
    double boxStep = 0;
    int boxGroups[3] = {0, 0, 0};  // scaling group of each box side (log / linear)
    int boxIndependent = 0, boxOneSide = 0;

    // grid the current arrays are sorted by, and whether it is ready for the next cycle
    Grid grid{};
    bool sortedForCycle = false;

    // polyspheres: bead table on the device, total range, bead items for overlap counts
    double *dBeads = nullptr;
    double totalRange = 0;
    double beadVolume = 0;  // sum of bead volumes (NaN when beads overlap)
    int *beadCellStart = nullptr, *beadCellCount = nullptr;  // cell arrays of bead grids (kept apart from the
    int beadNcellCap = 0;                                     // sweep grid's, which box moves must not disturb)
    DeviceParticles beadA, beadB;

    // state saved by rp_try_scaling for rp_revert_scaling (Packing::lastBox / lastShapes / lastScalingNumOverlaps)
    DeviceParticles last;
    bool scalingSaved = false;
    HostBox lastBox;
    long long lastNumOverlaps = 0;

    // counters
    uint64_t boxAttempted = 0, boxAccepted = 0, boxScans = 0, boxTooSmall = 0, cycles = 0;
    uint64_t accSinceEval = 0, attSinceEval = 0;      // host view used by the step adaptation
    uint64_t accBase = 0, attBase = 0;
    uint64_t boxSinceEval = 0, boxAccSinceEval = 0;

    // launch shape
    size_t ldsLimit = 64 * 1024 - 4096;  // dynamic LDS a workgroup may use
    int sweepWaves = 0;                  // wavefronts per active cell; 0 = chosen from the stencil size
    int colourPeriod = 0;                // checkerboard colour period (2 or 3); 0 = by shape
    int pipelined = -1;                  // staged sweep: 0 plain, 1 planner/tester split, 2 grouped moves, -1 auto
    // sweep workgroups launched, and their number when the counters were last read: the stencil cap that lets two
    // eight-wavefront workgroups share a compute unit is dropped once more than 1 % of them overflow it
    uint64_t blocksLaunched = 0, blocksAtRead = 0;
    bool pairCapOff = false;

    // learned candidate ordering: the network's parameters with Adam's moments and a gradient (4 x MLP_PARAMS), the
    // training pairs and the pairs of the last order statistics (x: 12 features, y: exact decision, meta, score)
    struct MlpData {
        float *x = nullptr, *y = nullptr, *score = nullptr;
        int *meta = nullptr;
        size_t cap = 0;
    };
    float *mlpTheta = nullptr;
    MlpData mlpSet, mlpEval;
    uint32_t mlpStream = 0;
    bool mlpTrained = false;

    // timing
    int timing = 0;  // 0 off, 1 every launch, 2 one sweep launch per cycle (RP_TIMING_SAMPLED)
    std::vector<hipEvent_t> evStart, evStop;
    std::vector<int> evKind;
    size_t evUsed = 0;
    rp_timing tstats{};
};

namespace {

int allocParticles(DeviceParticles &d, size_t n) {
    size_t bytes = n * (12 * sizeof(double) + 2 * sizeof(int) + 1) + 4096;
    RP_HIP(hipMalloc(&d.mem, bytes));
    d.bytes = bytes;
    char *q = static_cast<char *>(d.mem);
    auto take = [&](size_t b) {
        char *r = q;
        q += (b + 255) / 256 * 256;
        return r;
    };
    d.p.x = reinterpret_cast<double *>(take(n * 8));
    d.p.y = reinterpret_cast<double *>(take(n * 8));
    d.p.z = reinterpret_cast<double *>(take(n * 8));
    for (int c = 0; c < 9; c++) d.p.o[c] = reinterpret_cast<double *>(take(n * 8));
    d.p.tag = reinterpret_cast<int *>(take(n * 4));
    d.p.cell = reinterpret_cast<int *>(take(n * 4));
    d.p.wbits = reinterpret_cast<uint8_t *>(take(n));
    return RP_OK;
}

// sigma * pow(2, 1/6), as RepulsiveLennardJonesInteraction's constructor computes it
double wcaCut(const rp_shape_desc &s) { return s.sigma * std::pow(2., 1. / 6); }

// The soft kind of a descriptor without the hard-core flag, and whether the shape's hard test applies.
int softKind(const rp_shape_desc &s) { return s.soft & ~RP_SOFT_HARD_CORE; }
bool compound(const rp_shape_desc &s) { return softKind(s) != RP_SOFT_NONE && (s.soft & RP_SOFT_HARD_CORE) != 0; }
bool hasHardPart(const rp_shape_desc &s) { return softKind(s) == RP_SOFT_NONE || compound(s); }

PairConsts makeConsts(const rp_shape_desc &s) {
    PairConsts k{};
    k.kind = s.kind;
    k.soft = softKind(s);
    k.hardCore = compound(s) ? 1 : 0;
    if (s.kind == RP_SHAPE_SPHERE) {
        k.contact2 = std::pow(2 * s.radius, 2);              // SphereTraits.cpp:64
    } else {
        k.contact2 = 4 * s.radius * s.radius;                // SpherocylinderTraits.cpp:35, 44
        k.range2 = std::pow(2 * s.radius + s.length, 2);     // SpherocylinderTraits.cpp:37
        k.halfLen = 0.5 * s.length;                          // SpherocylinderTraits.cpp:24-26 (0.5 * 1 * L)
    }
    const int soft = softKind(s);
    if (soft) {
        const double cut = wcaCut(s);
        k.eps = s.epsilon;
        k.sigma2 = s.sigma * s.sigma;                        // RepulsiveLennardJonesInteraction.cpp:12
        k.cut2 = cut * cut;                                  // RepulsiveLennardJonesInteraction.cpp:9
        // segment distance >= centre distance - L: beyond (L + cut) the energy is 0; the margin keeps every pair
        // whose rounded segment distance could still fall below the cut-off
        k.softRange2 = soft == RP_SOFT_WCA_SEGMENT ? std::pow(s.length + cut, 2) * (1 + 1e-9) : k.cut2;
        if (soft == RP_SOFT_LJ) {
            // LennardJonesInteraction.cpp:9 (sigma * sigma / d2); truncated at the range 3 sigma
            k.cut2 = std::pow(3 * s.sigma, 2);
            k.softRange2 = k.cut2;
        } else if (soft == RP_SOFT_SIC) {
            k.sigma2 = std::pow(s.sigma, 2);                // SquareInverseCoreInteraction.cpp:14
            k.cut2 = k.sigma2;                              // zero beyond sigma
            k.softRange2 = k.cut2;
        }
    }
    return k;
}

// Interaction range: SphereTraits.h:34, SpherocylinderTraits.h:81; a WCA sphere the range of its interaction
// (RepulsiveLennardJonesInteraction.h:26); a WCA spherocylinder L + sigma 2^(1/6)
double hardRange(const rp_shape_desc &s);
// a compound's range is the larger of its parts' (CompoundInteraction.cpp:33-35)
double shapeRange(const rp_shape_desc &s) {
    const int soft = softKind(s);
    double r = 0;
    if (soft == RP_SOFT_WCA) r = wcaCut(s);
    else if (soft == RP_SOFT_LJ) r = 3 * s.sigma;   // LennardJonesInteraction.h:25
    else if (soft == RP_SOFT_SIC) r = s.sigma;      // SquareInverseCoreInteraction.h:22
    else if (soft == RP_SOFT_WCA_SEGMENT) r = s.length + wcaCut(s);
    else return hardRange(s);
    return compound(s) ? std::max(r, hardRange(s)) : r;
}

double hardRange(const rp_shape_desc &s) {
    if (s.kind == RP_SHAPE_POLYSPHERE) {
        // PolysphereTraits.cpp:116-121: twice the largest bead radius
        double m = 0;
        for (int i = 0; i < s.n_centres; i++) m = std::max(m, s.radii[i]);
        return 2 * m;
    }
    return s.kind == RP_SHAPE_SPHERE ? 2 * s.radius : 2 * s.radius + s.length;
}

// Interaction::getTotalRangeRadius (Interaction.cpp:55-70): 2 max |centre| + range for multi-centre shapes
double shapeTotalRange(const rp_shape_desc &s) {
    if (s.kind != RP_SHAPE_POLYSPHERE) return shapeRange(s);
    double maxR2 = 0;
    for (int i = 0; i < s.n_centres; i++) {
        const double *q = s.centres + 3 * i;
        double r2 = 0;
        r2 += q[0] * q[0];
        r2 += q[1] * q[1];
        r2 += q[2] * q[2];
        maxR2 = std::max(maxR2, r2);
    }
    return 2 * std::sqrt(maxR2) + shapeRange(s);
}

// Radius of the sphere about the particle centre that holds every bead: max |centre| + radius, raised by a relative
// 1e-12 so that the rounding of the square root cannot make it smaller than any bead's reach.
double shapeBound(const rp_shape_desc &s) {
    double bound = 0;
    for (int i = 0; i < s.n_centres; i++) {
        const double *q = s.centres + 3 * i;
        bound = std::max(bound, std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) + s.radii[i]);
    }
    return bound * (1 + 1e-12);
}

// Shape descriptor checks shared by rp_create and the pair entry points.
int validateShape(const rp_shape_desc &s) {
    if (s.kind == RP_SHAPE_POLYSPHERE) {
        if (s.n_centres < 1 || s.n_centres > RP_MAX_CENTRES)
            return fail(RP_ERR_PRECONDITION, "a polysphere needs 1 to " + std::to_string(RP_MAX_CENTRES) + " beads");
        if (!s.centres || !s.radii) return fail(RP_ERR_PRECONDITION, "polysphere centres or radii missing");
        for (int i = 0; i < s.n_centres; i++)  // SphereData constructor, PolysphereTraits.cpp:83-87
            if (!(s.radii[i] > 0)) return fail(RP_ERR_PRECONDITION, "bead radius must be > 0");
        if (s.soft != RP_SOFT_NONE) return fail(RP_ERR_UNSUPPORTED, "soft interaction not available for this shape");
        return RP_OK;
    }
    if (s.kind != RP_SHAPE_SPHERE && s.kind != RP_SHAPE_SPHEROCYLINDER)
        return fail(RP_ERR_UNSUPPORTED, "unsupported shape kind");
    if (!(s.radius > 0)) return fail(RP_ERR_PRECONDITION, "radius must be > 0");
    const int soft = softKind(s);
    if (s.soft & ~(RP_SOFT_HARD_CORE | 0xff)) return fail(RP_ERR_PRECONDITION, "unknown soft interaction flags");
    if (soft == RP_SOFT_NONE && s.soft != RP_SOFT_NONE)
        return fail(RP_ERR_PRECONDITION, "RP_SOFT_HARD_CORE needs a soft interaction");
    if (soft != RP_SOFT_NONE) {
        const bool sphereSoft = soft == RP_SOFT_WCA || soft == RP_SOFT_LJ || soft == RP_SOFT_SIC;
        if (!((sphereSoft && s.kind == RP_SHAPE_SPHERE) ||
              (soft == RP_SOFT_WCA_SEGMENT && s.kind == RP_SHAPE_SPHEROCYLINDER)))
            return fail(RP_ERR_UNSUPPORTED, "soft interaction not available for this shape");
        // RepulsiveLennardJonesInteraction.cpp:21-22, LennardJonesInteraction.cpp:16-17; the square-inverse core
        // only needs a nonzero epsilon (SquareInverseCoreInteraction.cpp:8-9)
        if (soft == RP_SOFT_SIC ? !(s.epsilon != 0 && std::isfinite(s.epsilon)) : !(s.epsilon > 0))
            return fail(RP_ERR_PRECONDITION, soft == RP_SOFT_SIC ? "epsilon must be nonzero" : "epsilon must be > 0");
        if (!(s.sigma > 0)) return fail(RP_ERR_PRECONDITION, "sigma must be > 0");
    }
    if (s.kind == RP_SHAPE_SPHEROCYLINDER && !(s.length >= 0))
        return fail(RP_ERR_PRECONDITION, "length must be >= 0");
    return RP_OK;
}

// Device copy of a polysphere's beads (body-frame x, y, z, radius per bead) for k.beads; nullptr otherwise.
int uploadBeads(const rp_shape_desc &s, double **dBeads) {
    *dBeads = nullptr;
    if (s.kind != RP_SHAPE_POLYSPHERE) return RP_OK;
    std::vector<double> h(4 * (size_t)s.n_centres);
    for (int i = 0; i < s.n_centres; i++) {
        for (int d = 0; d < 3; d++) h[4 * i + d] = s.centres[3 * i + d];
        h[4 * i + 3] = s.radii[i];
    }
    RP_HIP(hipMalloc(dBeads, sizeof(double) * h.size()));
    if (hipMemcpy(*dBeads, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(*dBeads);
        *dBeads = nullptr;
        return fail(RP_ERR_DEVICE, "bead upload failed");
    }
    return RP_OK;
}

void fillTranslations(Grid &g) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                V3 t = mulv(g.dims, v3((double)i - 1, (double)j - 1., (double)k - 1.));
                int idx = 3 * (i * 9 + j * 3 + k);
                g.T[idx] = t.x;
                g.T[idx + 1] = t.y;
                g.T[idx + 2] = t.z;
            }
}

// Cell arrays of at least ncell cells: the shared ones, or (beads) those of the bead grids.
int ensureCells(rp_ctx *c, int ncell, bool beads = false) {
    int *&start = beads ? c->beadCellStart : c->cellStart;
    int *&count = beads ? c->beadCellCount : c->cellCount;
    int &have = beads ? c->beadNcellCap : c->ncellCap;
    if (ncell <= have) return RP_OK;
    if (start) hipFree(start);
    if (count) hipFree(count);
    start = count = nullptr;
    int cap = ncell + ncell / 4 + 64;
    RP_HIP(hipMalloc(&start, sizeof(int) * (cap + 1)));
    RP_HIP(hipMalloc(&count, sizeof(int) * cap));
    RP_HIP(launch_fill_i32(c->stream, count, (size_t)cap, 0));  // rebuilds keep them zero from here on
    have = cap;
    const size_t scanInts = rebuild_scan_ints(cap);
    if (scanInts > c->scanTmpCap) {
        if (c->scanTmp) hipFree(c->scanTmp);
        c->scanTmp = nullptr;
        c->scanTmpCap = 0;
        RP_HIP(hipMalloc(&c->scanTmp, sizeof(int) * scanInts));
        c->scanTmpCap = scanInts;
    }
    return RP_OK;
}

// Move-count slots for a grid of ncell cells.  Every fold clears the slots it covers, so between cycles they are all
// zero and may be replaced.
// Folds move counts the slots still hold into the counters (the next cycle's proposal launch does it otherwise).
int foldSlots(rp_ctx *c) {
    if (!c->slotsDirty) return RP_OK;
    RP_HIP(launch_fold_slots(c->stream, c->dSlots, c->slotsDirty, c->dCounters));
    c->slotsDirty = 0;
    return RP_OK;
}

int ensureSlots(rp_ctx *c, int ncell) {
    if ((size_t)ncell <= c->slotsCap) return RP_OK;
    int r = foldSlots(c);
    if (r) return r;
    if (c->dSlots) hipFree(c->dSlots);
    c->dSlots = nullptr;
    c->slotsCap = 0;
    const size_t cap = (size_t)ncell + (size_t)ncell / 4 + 64;
    RP_HIP(hipMalloc(&c->dSlots, sizeof(unsigned long long) * SLOT_WORDS * cap));
    RP_HIP(launch_fill_u64(c->stream, c->dSlots, SLOT_WORDS * cap, 0));
    c->slotsCap = cap;
    return RP_OK;
}

int ensureProposals(rp_ctx *c, size_t need) {
    if (need <= c->propCap) return RP_OK;
    if (c->propMem) hipFree(c->propMem);
    c->propMem = nullptr;
    c->propCap = 0;
    size_t cap = need + need / 4 + 256;
    RP_HIP(hipMalloc(&c->propMem, cap * (sizeof(int) + 13 * sizeof(double) + 2)));
    char *q = static_cast<char *>(c->propMem);
    c->prop.t = reinterpret_cast<double *>(q);
    c->prop.R = c->prop.t + 3 * cap;
    c->prop.u = c->k.soft ? c->prop.R + 9 * cap : nullptr;
    c->prop.k = reinterpret_cast<int *>(c->prop.R + 10 * cap);
    c->prop.type = reinterpret_cast<uint8_t *>(c->prop.k + cap);
    c->prop.acc = c->prop.type + cap;
    c->prop.stride = (int)cap;
    c->propCap = cap;
    return RP_OK;
}

// Checkerboard grid with colour period P: cells at least range / (P - 1) high and, as in the reference's grid
// (Packing.cpp:862-866), at least a fifth of the mean interparticle spacing; a multiple of P cells per axis, at
// least 2 * reach + 1 of them, so that a stencil offset wraps around the box at most once.
int makeSweepGrid(rp_ctx *c, const HostBox &b, const double shiftUnit[3], Grid &g, int period) {
    double h[3];
    b.heights(h);
    std::memcpy(g.dims, b.dims, sizeof g.dims);
    std::memcpy(g.inv, b.inv, sizeof g.inv);
    const int reach = period - 1;
    // a particle-level grid: a polysphere reaches other particles within its total range
    const double gridRange = c->k.kind == RP_SHAPE_POLYSPHERE ? c->totalRange : c->range;
    double cellSize = std::max(gridRange / reach, std::cbrt(b.volume() / (double)c->n) / 5);
    for (int d = 0; d < 3; d++) {
        int n = period * (int)std::floor(h[d] / (period * cellSize));
        if (period == 2 && n < 2 && h[d] >= cellSize) {
            // one cell across this axis: it is its own neighbour under the images -1, 0, +1 (as in the reference's
            // grid with one cell per axis, NeighbourGrid.cpp:80-100), only colours of even parity on the axis
            // exist, and the origin shift is 0 so that image offsets stay in {-1, 0, 1}
            n = 1;
        } else if (n < period || (reach > 1 && n < 2 * reach + 1)) {
            return fail(RP_ERR_UNSUPPORTED, "box height " + std::to_string(h[d]) + " is too small for a checkerboard "
                                            "of period " + std::to_string(period));
        }
        g.n[d] = n;
        g.relCell[d] = 1 / (double)n;
        g.shift[d] = n == 1 ? 0.0 : shiftUnit[d] * g.relCell[d];
    }
    g.ncell = g.n[0] * g.n[1] * g.n[2];
    g.exact = 0;
    g.reach = reach;
    g.period = period;
    fillTranslations(g);
    int r = ensureCells(c, g.ncell);
    if (r) return r;
    g.cellStart = c->cellStart;
    g.cellCount = c->cellCount;
    return RP_OK;
}

// Whether the reference would run box b without its neighbour grid (Packing::rebuildNeighbourGrid,
// Packing.cpp:862-883): cells of the interaction range (at least a fifth of the mean spacing) with fewer than four
// across every direction, or wider than the box in one.  Such packings are run here through every pair and the
// nearest image, as the reference then does; the all-pairs path covers hard spheres and spherocylinders.
bool noGridBox(const rp_ctx *c, const HostBox &b) {
    if (c->k.soft || c->counting || c->k.kind == RP_SHAPE_POLYSPHERE) return false;
    double h[3];
    b.heights(h);
    double cell = c->range;
    const double minCell = std::cbrt(b.volume() / (double)c->n) / 5;
    if (c->range < minCell) cell = minCell;
    const double hmax = std::max(h[0], std::max(h[1], h[2])), hmin = std::min(h[0], std::min(h[1], h[2]));
    return cell * 4 > hmax || cell * (1 + 1e-12) > hmin;
}

// The grid of a packing without a neighbour grid: one cell that holds every particle, no origin shift.
int makeAllPairsGrid(rp_ctx *c, const HostBox &b, Grid &g) {
    std::memcpy(g.dims, b.dims, sizeof g.dims);
    std::memcpy(g.inv, b.inv, sizeof g.inv);
    for (int d = 0; d < 3; d++) {
        g.n[d] = 1;
        g.relCell[d] = 1;
        g.shift[d] = 0;
    }
    g.ncell = 1;
    g.exact = 1;
    g.reach = 1;
    g.period = 1;
    g.allPairs = 1;
    fillTranslations(g);
    int r = ensureCells(c, 1);
    if (r) return r;
    g.cellStart = c->cellStart;
    g.cellCount = c->cellCount;
    return RP_OK;
}

// Particles the all-pairs sweep stages in LDS: all of them when they fit, else none (it reads them in place).
int allPairsStaged(const rp_ctx *c) {
    return allpairs_lds_bytes((int)c->n) <= c->ldsLimit ? (int)c->n : 0;
}

// Colour period of the checkerboard: RAMPACK_COLOURING (2 or 3) when set; otherwise 3 for polyspheres and 2 for
// the other shapes.  Period 2 has 8 colours and cells >= range.  Period 3 has 27 colours and cells >= range / 2:
// at the same density that is 27 launches per cycle of workgroups with a few particles each and 125-cell
// stencils, which measured about four times slower per trial move on 1e5 rods.  A polysphere's sweep cell spans
// the whole molecule, so period 2 leaves few active cells per colour and stencils of thousands of beads; period 3
// measured 1.9x faster on 50 653 kinked 11-bead chains.  Period 3 falls back to 2 when the box is too small for it.
int chooseSweepGrid(rp_ctx *c, const HostBox &b, const double shiftUnit[3], Grid &g) {
    if (noGridBox(c, b)) return makeAllPairsGrid(c, b, g);
    const int period = c->colourPeriod ? c->colourPeriod : (c->k.kind == RP_SHAPE_POLYSPHERE ? 3 : 2);
    if (period == 3 && makeSweepGrid(c, b, shiftUnit, g, 3) == RP_OK) return RP_OK;
    return makeSweepGrid(c, b, shiftUnit, g, 2);
}

// The reference's neighbour grid (Packing.cpp:858-883, NeighbourGrid.cpp:178-201).  Returns 1 when the
// reference would run without a grid.
int makeExactGrid(rp_ctx *c, const HostBox &b, Grid &g, int &disabled, bool beads = false) {
    double h[3];
    b.heights(h);
    double cell = c->range;
    double minCell = std::cbrt(b.volume() / (double)c->n) / 5;
    if (c->range < minCell) cell = minCell;
    double hmax = std::max(h[0], std::max(h[1], h[2])), hmin = std::min(h[0], std::min(h[1], h[2]));
    disabled = (cell * 4 > hmax || cell * (1 + 1e-12) > hmin) ? 1 : 0;
    if (disabled) return RP_OK;
    std::memcpy(g.dims, b.dims, sizeof g.dims);
    std::memcpy(g.inv, b.inv, sizeof g.inv);
    for (int d = 0; d < 3; d++) {
        g.n[d] = (int)std::floor(h[d] / cell);
        g.relCell[d] = 1 / (double)g.n[d];
        g.shift[d] = 0;
    }
    g.ncell = g.n[0] * g.n[1] * g.n[2];
    g.exact = 1;
    g.reach = 1;
    g.period = 2;
    fillTranslations(g);
    int r = ensureCells(c, g.ncell, beads);
    if (r) return r;
    g.cellStart = beads ? c->beadCellStart : c->cellStart;
    g.cellCount = beads ? c->beadCellCount : c->cellCount;
    return RP_OK;
}

// Events for one timed launch of the given kind (0 sweep, 1 rebuild, 2 overlap scan, 3 proposals), or none when
// timing is off.  The launch itself records them (hipExtLaunchKernel), so timing adds no packets between kernels.
Timed timingSlot(rp_ctx *c, int kind) {
    if (!c->timing || (c->timing == RP_TIMING_SAMPLED && kind != 0)) return Timed{};
    if (c->evUsed == c->evStart.size()) {
        // timing only, read after a stream synchronisation: no system-scope fence (cache write-back and
        // invalidation) per record
        hipEvent_t a, b;
        hipEventCreateWithFlags(&a, hipEventDisableSystemFence);
        hipEventCreateWithFlags(&b, hipEventDisableSystemFence);
        c->evStart.push_back(a);
        c->evStop.push_back(b);
        c->evKind.push_back(0);
    }
    c->evKind[c->evUsed] = kind;
    Timed t;
    t.start = c->evStart[c->evUsed];
    t.stop = c->evStop[c->evUsed];
    c->evUsed++;
    return t;
}

int timingHarvest(rp_ctx *c) {
    if (!c->timing || c->evUsed == 0) return RP_OK;
    RP_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->evUsed; i++) {
        float ms = 0;
        RP_HIP(hipEventElapsedTime(&ms, c->evStart[i], c->evStop[i]));
        if (c->evKind[i] == 0) { c->tstats.sweep_launches++; c->tstats.sweep_ms += ms; }
        else if (c->evKind[i] == 1) { c->tstats.rebuild_launches++; c->tstats.rebuild_ms += ms; }
        else if (c->evKind[i] == 2) { c->tstats.scan_launches++; c->tstats.scan_ms += ms; }
        else { c->tstats.proposal_launches++; c->tstats.proposal_ms += ms; }
    }
    c->evUsed = 0;
    return RP_OK;
}

int rebuild(rp_ctx *c, const Grid &g, int rescale, const double *oldInv) {
    RP_HIP(launch_rebuild(c->stream, g, (int)c->n, c->cur.p, c->alt.p, c->tmpSlot, c->tmpOrder, c->tmpPos, rescale,
                          oldInv, timingSlot(c, 1), c->scanTmp));
    return RP_OK;
}

int readCounters(rp_ctx *c) {
    int r = foldSlots(c);
    if (r) return r;
    RP_HIP(hipMemcpyAsync(c->hCounters, c->dCounters, sizeof(unsigned long long) * CNT_NUM, hipMemcpyDeviceToHost,
                          c->stream));
    RP_HIP(hipStreamSynchronize(c->stream));
    c->blocksAtRead = c->blocksLaunched;
    return RP_OK;
}

// Largest multiple of 64 entries whose LDS footprint (perEntry bytes each) fits the workgroup limit.
int capForBytes(rp_ctx *c, double perEntry) { return std::max(64, (int)((double)c->ldsLimit / perEntry) / 64 * 64); }

// Largest stencil the overlap-count / energy kernels can stage within the LDS limit.
int countCapLimit(rp_ctx *c) {
    const size_t fixed = count_lds_bytes(c->k.kind, 0);
    const double per = (double)(count_lds_bytes(c->k.kind, 1024) - fixed) / 1024.0;
    return std::max(64, (int)((double)(c->ldsLimit - fixed) / per) / 64 * 64);
}

// Overlap count of `items` sorted by `g`, enqueued with the copy of the counters to the host (no synchronisation).
int scanLaunch(rp_ctx *c, const Grid &g, const Particles &items, size_t nItems, int earlyExit) {
    RP_HIP(launch_clear_count(c->stream, c->dCounters));
    double lambda = (double)nItems / (double)g.ncell;
    const int side = 2 * g.reach + 1;
    double expect = ((side * side * side - 1) / 2 + 1) * lambda;
    int cap = (int)std::ceil(expect + 8 * std::sqrt(expect) + 64);
    cap = std::min((cap + 63) / 64 * 64, countCapLimit(c));
    if (g.allPairs)
        RP_HIP(launch_count_allpairs(c->stream, g, (int)nItems, items, c->k, earlyExit, c->dCounters));
    else
        // an early-exit count starts a few hundred cells before the cell of the last overlap found: a box move that
        // is rejected for an overlap is most often followed by one that meets the same near pair
        RP_HIP(launch_count_overlaps(c->stream, g, (int)nItems, items, c->k, earlyExit, c->dCounters, cap,
                                     timingSlot(c, 2), earlyExit ? c->scanHint - 300 : 0));
    RP_HIP(hipMemcpyAsync(c->hCounters, c->dCounters, sizeof(unsigned long long) * CNT_NUM, hipMemcpyDeviceToHost,
                          c->stream));
    return RP_OK;
}

// The count of a scanLaunch, once the stream has been synchronised.
int scanResult(rp_ctx *c, uint64_t &count) {
    c->blocksAtRead = c->blocksLaunched;
    if (c->hCounters[CNT_STAGE_OVERFLOW])
        return fail(RP_ERR_UNSUPPORTED, "a grid cell holds more particles than the overlap count can stage");
    count = c->hCounters[CNT_OVERLAPS];
    if (count) c->scanHint = (int)c->hCounters[CNT_SCAN_HINT];
    return RP_OK;
}

// Early-exit overlap scan of the `alt` arrays sorted by `g`; sets found.
int scanItems(rp_ctx *c, const Grid &g, const Particles &items, size_t nItems, int earlyExit, uint64_t &count) {
    int r = scanLaunch(c, g, items, nItems, earlyExit);
    if (r) return r;
    RP_HIP(hipStreamSynchronize(c->stream));
    return scanResult(c, count);
}

int scanOverlaps(rp_ctx *c, const Grid &g, int earlyExit, uint64_t &count) {
    return scanItems(c, g, c->alt.p, c->n, earlyExit, count);
}

void swapBuffers(rp_ctx *c) { std::swap(c->cur, c->alt); }

double minHeight(const HostBox &b) {
    double h[3];
    b.heights(h);
    return std::min(h[0], std::min(h[1], h[2]));
}

// Simulation::tryScaling for a hard interaction (Simulation.cpp:493-531) + Packing::tryScaling
// (Packing.cpp:172-208).  On acceptance the rescaled arrays, already sorted by the next cycle's grid, become
// current.
// Packing::getTotalEnergy (Packing.cpp:690-709) of particles sorted by grid g.
int totalEnergy(rp_ctx *c, const Grid &g, const Particles &p, double &out);
int countBeads(rp_ctx *c, const Particles &src, const HostBox &box, int earlyExit, uint64_t &out);

// Simulation::tryScaling in overlap-counting mode (Simulation.cpp:508-517 with Packing::tryScaling's counting branch,
// Packing.cpp:190-198): the rescaled packing is always counted in full; fewer overlaps accept the move, more reject
// it, the same number leaves the ordinary Metropolis test.
int countingBoxMove(rp_ctx *c, const HostBox &nb, double exponent, double u, const double shiftUnit[3],
                    bool &accepted) {
    Grid g{};
    int r = chooseSweepGrid(c, nb, shiftUnit, g);
    if (r == RP_ERR_UNSUPPORTED) {
        c->boxTooSmall++;
        return RP_OK;
    }
    if (r) return r;
    r = rebuild(c, g, 1, c->box.inv);
    if (r) return r;
    uint64_t after = 0;
    c->boxScans++;
    r = c->k.kind == RP_SHAPE_POLYSPHERE ? countBeads(c, c->alt.p, nb, 0, after) : scanOverlaps(c, g, 0, after);
    if (r) return r;
    const long long delta = (long long)after - c->numOverlaps;
    if (delta > 0 || (delta == 0 && !(u <= std::exp(exponent)))) return RP_OK;
    c->box = nb;
    swapBuffers(c);
    c->grid = g;
    c->sortedForCycle = true;
    c->numOverlaps = (long long)after;
    accepted = true;
    return RP_OK;
}

// Simulation::tryScaling with a soft interaction (Simulation.cpp:508-517): the scaling is always tried, and
// accepted if u <= exp(N ln(V'/V) - dE/T - p dV/T) with dE = getTotalEnergy(new) - getTotalEnergy(old)
// (Packing.cpp:172-208).  The old energy is summed on the cycle's grid, which the colour passes keep valid.
int softBoxMove(rp_ctx *c, const HostBox &nb, double exponent, double u, const double shiftUnit[3], bool &accepted) {
    double eOld = 0, eNew = 0;
    int r = totalEnergy(c, c->grid, c->cur.p, eOld);
    if (r) return r;
    Grid g{};
    r = chooseSweepGrid(c, nb, shiftUnit, g);
    if (r == RP_ERR_UNSUPPORTED) {
        c->boxTooSmall++;
        return RP_OK;
    }
    if (r) return r;
    r = rebuild(c, g, 1, c->box.inv);
    if (r) return r;
    c->boxScans++;
    if (c->k.hardCore && !isBoxUpscaled(c->box, nb)) {
        // the compound's hard part: an overlap makes dE infinite (Packing.cpp:199-204); the compound is convex
        // (CompoundInteraction.cpp:25-31)
        uint64_t found = 0;
        r = scanOverlaps(c, g, 1, found);
        if (r) return r;
        if (found) return RP_OK;
    }
    r = totalEnergy(c, g, c->alt.p, eNew);
    if (r) return r;
    if (!(u <= std::exp(exponent - (eNew - eOld) / c->temperature))) return RP_OK;
    c->box = nb;
    swapBuffers(c);
    c->grid = g;
    c->sortedForCycle = true;
    accepted = true;
    return RP_OK;
}

int totalEnergy(rp_ctx *c, const Grid &g, const Particles &p, double &out) {
    if (g.ncell > c->cellEnergyCap) {
        if (c->cellEnergy) hipFree(c->cellEnergy);
        c->cellEnergy = nullptr;
        c->cellEnergyCap = 0;
        RP_HIP(hipMalloc(&c->cellEnergy, sizeof(double) * (size_t)g.ncell));
        c->cellEnergyCap = g.ncell;
    }
    if (!c->sumScratch) RP_HIP(hipMalloc(&c->sumScratch, sizeof(double) * fixed_sum_scratch_doubles()));
    double lambda = (double)c->n / (double)g.ncell;
    const int side = 2 * g.reach + 1;
    double expect = ((side * side * side - 1) / 2 + 1) * lambda;
    int cap = (int)std::ceil(expect + 8 * std::sqrt(expect) + 64);
    cap = std::min((cap + 63) / 64 * 64, countCapLimit(c));
    RP_HIP(launch_fill_u64(c->stream, c->dCounters + CNT_STAGE_OVERFLOW, 1, 0));
    RP_HIP(launch_total_energy(c->stream, g, (int)c->n, p, c->k, c->cellEnergy, c->sumScratch, c->dCounters, cap));
    RP_HIP(hipMemcpyAsync(&out, c->sumScratch + fixed_sum_scratch_doubles() - 1, sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    int r = readCounters(c);  // synchronises the stream
    if (r) return r;
    if (c->hCounters[CNT_STAGE_OVERFLOW])
        return fail(RP_ERR_UNSUPPORTED, "a grid cell holds more particles than the energy sum can stage");
    return RP_OK;
}

// A box move's draws from the context's host stream: the new box nb, the Metropolis exponent (without the energy
// change), the uniform u of the test and the origin shift of the new grid.
int drawBoxMove(rp_ctx *c, HostBox &nb, double &exponent, double &u, double shiftUnit[3]) {
    double nd[9];
    std::memcpy(nd, c->box.dims, sizeof nd);
    if (c->boxKind == RP_BOX_DELTA_TRICLINIC) {
        if (c->boxOneSide) {
            // TriclinicDeltaScaler.cpp:17-23: one side (column) drawn uniformly
            const int side = (int)philox_below(c->hostRng, 3);
            for (int i = 0; i < 3; i++) nd[3 * i + side] += philox_uniform(c->hostRng, -c->boxStep, c->boxStep);
        } else {
            for (int i = 0; i < 9; i++) nd[i] += philox_uniform(c->hostRng, -c->boxStep, c->boxStep);
        }
    } else if (c->boxKind == RP_BOX_LOG_ISOTROPIC || c->boxKind == RP_BOX_LINEAR) {
        // TriclinicAdapter::updateBox (TriclinicAdapter.h:27-33): side c (column c) times factor c, the factors from
        // AnisotropicVolumeScaler::sampleScalingFactors (AnisotropicVolumeScaler.cpp:13-31) with the log sampler
        // (exp(U), LogScalingFactorSampler.cpp:8-19) or the linear one ((h + U) / h, LinearScalingFactorSampler.cpp)
        double h[3];
        c->box.heights(h);
        int groupList[3], ngroups = 0;
        for (int a = 0; a < 3; a++) {
            if (c->boxGroups[a] == 3) continue;  // a side that is not scaled ("[y]" in the reference's spec)
            bool seen = false;
            for (int g = 0; g < ngroups; g++) seen = seen || groupList[g] == c->boxGroups[a];
            if (!seen) groupList[ngroups++] = c->boxGroups[a];
        }
        double factors[3] = {1, 1, 1};
        auto sampleGroup = [&](int group) {
            const double u = philox_uniform(c->hostRng, -c->boxStep, c->boxStep);
            const double logFactor = std::exp(u);
            for (int a = 0; a < 3; a++)
                if (c->boxGroups[a] == group)
                    factors[a] *= c->boxKind == RP_BOX_LOG_ISOTROPIC ? logFactor : (h[a] + u) / h[a];
        };
        if (c->boxIndependent && ngroups > 1) sampleGroup(groupList[philox_below(c->hostRng, (uint32_t)ngroups)]);
        else
            for (int g = 0; g < ngroups; g++) sampleGroup(groupList[g]);
        for (int a = 0; a < 3; a++)
            if (!(factors[a] > 0)) return fail(RP_ERR_PRECONDITION, "box scaling factor <= 0 (step too large)");
        for (int r = 0; r < 3; r++)
            for (int col = 0; col < 3; col++)
                if (factors[col] != 1) nd[3 * r + col] = nd[3 * r + col] * factors[col];
    } else {
        // DeltaVolumeScaler: factor = cbrt((dV + V) / V) with V the product of the heights
        double h[3];
        c->box.heights(h);
        double curV = 1.;
        curV = curV * h[0];
        curV = curV * h[1];
        curV = curV * h[2];
        double dV = (2 * philox_unit(c->hostRng) - 1) * c->boxStep;
        double f = std::cbrt((dV + curV) / curV);
        for (int i = 0; i < 9; i++) nd[i] = nd[i] * f;
    }
    nb.set(nd);
    double oldV = c->box.volume(), newV = nb.volume();
    if (newV == 0) return fail(RP_ERR_PRECONDITION, "box move produced a zero-volume box");
    double dV = newV - oldV, factor = newV / oldV;
    double N = (double)c->n;
    exponent = N * std::log(factor) - c->pressure * dV / c->temperature;
    u = philox_unit(c->hostRng);
    for (int d = 0; d < 3; d++) shiftUnit[d] = philox_unit(c->hostRng);
    return RP_OK;
}

// A hard packing's box move whose overlap scan is in flight: decided by boxMoveEnd after a stream synchronisation.
struct BoxTrial {
    bool pending = false;
    HostBox nb;
    Grid g{};
};

void acceptBox(rp_ctx *c, const HostBox &nb, const Grid &g, bool &accepted) {
    c->box = nb;
    swapBuffers(c);
    c->grid = g;
    c->sortedForCycle = true;
    accepted = true;
}

// Box move of a hard packing from its draws: the Metropolis test, the rescaled cell list and the overlap scan.
// With `defer` the scan is only launched (defer->pending); boxMoveEnd decides once the stream is synchronised.
int hardBoxMove(rp_ctx *c, const HostBox &nb, double exponent, double u, const double shiftUnit[3], bool &accepted,
                BoxTrial *defer) {
    if (u > std::exp(exponent)) return RP_OK;

    Grid g{};
    int r = chooseSweepGrid(c, nb, shiftUnit, g);
    if (r == RP_ERR_UNSUPPORTED) {
        // the checkerboard needs two cells per axis; the proposal is rejected and counted
        c->boxTooSmall++;
        return RP_OK;
    }
    if (r) return r;
    r = rebuild(c, g, 1, c->box.inv);
    if (r) return r;
    // spheres and spherocylinders are convex; polyspheres are not (PolysphereTraits.h:133) and are always scanned,
    // bead by bead on the rescaled particles
    const bool poly = c->k.kind == RP_SHAPE_POLYSPHERE;
    bool cannotOverlap = !poly && isBoxUpscaled(c->box, nb);
    if (!cannotOverlap) {
        uint64_t found = 0;
        c->boxScans++;
        if (defer && !poly) {
            r = scanLaunch(c, g, c->alt.p, c->n, 1);
            if (r) return r;
            defer->pending = true;
            defer->nb = nb;
            defer->g = g;
            return RP_OK;
        }
        r = poly ? countBeads(c, c->alt.p, nb, 1, found) : scanOverlaps(c, g, 1, found);
        if (r) return r;
        if (found) return RP_OK;
    }
    acceptBox(c, nb, g, accepted);
    return RP_OK;
}

int boxMoveEnd(rp_ctx *c, BoxTrial &t, bool &accepted) {
    if (!t.pending) return RP_OK;
    t.pending = false;
    uint64_t found = 0;
    int r = scanResult(c, found);
    if (r) return r;
    if (!found) acceptBox(c, t.nb, t.g, accepted);
    return RP_OK;
}

// One box move (Simulation::tryScaling, Simulation.cpp:508-517).  defer: see hardBoxMove (hard packings only).
int boxMove(rp_ctx *c, bool &accepted, BoxTrial *defer = nullptr) {
    accepted = false;
    HostBox nb;
    double exponent = 0, u = 0, shiftUnit[3] = {0, 0, 0};
    int r = drawBoxMove(c, nb, exponent, u, shiftUnit);
    if (r) return r;
    if (c->k.soft) return softBoxMove(c, nb, exponent, u, shiftUnit, accepted);
    if (c->counting) return countingBoxMove(c, nb, exponent, u, shiftUnit, accepted);
    return hardBoxMove(c, nb, exponent, u, shiftUnit, accepted, defer);
}

int prepareCycleGrid(rp_ctx *c) {
    if (c->sortedForCycle) return RP_OK;
    double shiftUnit[3] = {philox_unit(c->hostRng), philox_unit(c->hostRng), philox_unit(c->hostRng)};
    Grid g{};
    int r = chooseSweepGrid(c, c->box, shiftUnit, g);
    if (r) return r;
    r = rebuild(c, g, 0, nullptr);
    if (r) return r;
    swapBuffers(c);
    c->grid = g;
    c->sortedForCycle = true;
    return RP_OK;
}

// One sweep launch with the per-workgroup trace on, written to the context's trace file as lines of
// "block start_us end_us own stencil cell cu" (times from the first workgroup's start).
int traceSweep(rp_ctx *c, SweepParams sp, int waves) {
    const size_t blocks = (size_t)sp.na[0] * sp.na[1] * sp.na[2];
    unsigned long long *d = nullptr;
    RP_HIP(hipMalloc(&d, sizeof(unsigned long long) * 4 * blocks));
    std::vector<unsigned long long> h(4 * blocks);
    int r = RP_OK;
    if (launch_fill_u64(c->stream, d, 4 * blocks, 0) != hipSuccess) r = fail(RP_ERR_DEVICE, "trace fill failed");
    sp.trace = d;
    if (!r && launch_sweep(c->stream, sp, waves) != hipSuccess) r = fail(RP_ERR_DEVICE, "traced sweep failed");
    if (!r && (hipMemcpyAsync(h.data(), d, sizeof(unsigned long long) * 4 * blocks, hipMemcpyDeviceToHost,
                              c->stream) != hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess))
        r = fail(RP_ERR_DEVICE, "trace copy failed");
    hipFree(d);
    if (r) return r;
    int khz = 100000, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;
    unsigned long long first = ~0ull;
    for (size_t i = 0; i < blocks; i++)
        if (h[4 * i]) first = std::min(first, h[4 * i]);
    FILE *f = std::fopen(c->tracePath.c_str(), "w");
    if (!f) return fail(RP_ERR_PRECONDITION, "cannot write " + c->tracePath);
    const double perUs = khz / 1000.0;
    for (size_t i = 0; i < blocks; i++) {
        if (!h[4 * i]) continue;  // no particle in the cell
        std::fprintf(f, "%zu %.2f %.2f %llu %llu %llu %llu\n", i, (h[4 * i] - first) / perUs,
                     (h[4 * i + 1] - first) / perUs, h[4 * i + 2] & 0xffff, (h[4 * i + 2] >> 16) & 0xffff,
                     h[4 * i + 2] >> 32, h[4 * i + 3]);
    }
    std::fclose(f);
    return RP_OK;
}

// Guard bands of the guarded FP32 scan for grid g (SweepParams::g32Lim).  An FP32 squared centre distance x
// differs from the FP64 one the reference computes by at most E(threshold) whenever x is near that threshold:
//   coordinates: a staged position q (FP64, q = pos + T for translated cells) and a trial position p enter as
//   fl32(fl64(q - o)) with |q - o| <= M, so each carries an error <= 1.01 (u32 + u64) M; on cells with wrapped
//   particles the FP32 image translation adds u32 |T| and the FP32 sum u32 |q|, with M a box length;
//   d = fq - fp adds u32 |d|; x = (dx dx + dy dy) + dz dz rounds three times (3 u32 x);
//   the FP64 value rounds the difference and the sum (at most 8 u64 x);
//   |sum (a_i^2 - b_i^2)| <= 2 e sqrt(3) |d| + 3 e^2 for per-coordinate errors e.
// Lower limits are rounded down and upper limits up when converted to FP32.
void g32Limits(const rp_ctx *c, const Grid &g, float lim[2][4]) {
    const double u = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    double mCell = 0, mBox = 0;
    for (int i = 0; i < 3; i++) {
        double a = 0, b = 0;
        for (int d = 0; d < 3; d++) {
            a += std::fabs(g.dims[3 * i + d]) * 1.5 / g.n[d];
            b += std::fabs(g.dims[3 * i + d]);
        }
        mCell = std::max(mCell, a);
        mBox = std::max(mBox, b);
    }
    mCell *= 1.01;
    mBox *= 1.01;
    const double th[2] = {c->k.contact2, c->k.kind == RP_SHAPE_SPHEROCYLINDER ? c->k.range2 : 0};
    for (int set = 0; set < 2; set++) {
        const double eq = set == 0 ? 1.01 * (u + u64) * mCell : 1.02 * (3 * u + 2 * u64) * mBox;
        const double ep = set == 0 ? 1.01 * (u + u64) * mCell : 1.01 * (u + u64) * mBox;
        for (int t = 0; t < 2; t++) {
            if (th[t] == 0) {
                lim[set][2 * t] = lim[set][2 * t + 1] = std::numeric_limits<float>::infinity();
                continue;
            }
            const double D = std::sqrt(th[t] * 1.01);
            const double e = eq + ep + u * D * 1.01;
            const double E = 1.05 * (2 * e * std::sqrt(3.0) * D + 3 * e * e + 3.01 * u * th[t] * 1.01 +
                                     8 * u64 * th[t] * 1.01);
            float lo = (float)(th[t] - E), hi = (float)(th[t] + E);
            if ((double)lo > th[t] - E) lo = std::nextafter(lo, -std::numeric_limits<float>::infinity());
            if ((double)hi < th[t] + E) hi = std::nextafter(hi, std::numeric_limits<float>::infinity());
            lim[set][2 * t] = lo;
            lim[set][2 * t + 1] = hi;
        }
    }
}

// What one cycle's particle moves used: the colour order and the proposal layout (for decision records).
struct SweepInfo {
    int colours = 0;
    int order[27] = {};
    int moves = 0, mult = 1;
    size_t proposals = 0;
};

// Number of proposals the cycle about to run draws on the current (prepared) grid.
size_t cycleProposals(const rp_ctx *c) {
    const double lambda = (double)c->n / (double)c->grid.ncell;
    const int nSamplers = (int)c->samplers.size();
    const int moves = lambda >= 3 ? (int)std::lround(lambda) * nSamplers : 0;
    return moves > 0 ? (size_t)c->grid.ncell * moves : c->n * (size_t)nSamplers;
}

// One cycle's particle moves, planned: the launch parameters of every colour pass (all but the colour), the launch
// shape, the colour order and the number of proposals.
struct SweepPlan {
    SweepParams sp{};
    int waves = 1, g32 = 0, colours = 0;
    int order[27] = {};
    size_t proposals = 0;
};

// Whether colour `colour` of grid g has active cells: a colour with odd parity on a one-cell axis has none.
bool colourHasCells(const Grid &g, int colour) {
    const int par[3] = {colour % g.period, (colour / g.period) % g.period, colour / (g.period * g.period)};
    return !((g.n[0] == 1 && par[0]) || (g.n[1] == 1 && par[1]) || (g.n[2] == 1 && par[2]));
}

// The first part of a cycle's particle moves: the cycle's grid, the proposals (launched), the launch shape and the
// colour order (drawn from the context's host stream).  logDecisions: every proposal's acceptance flag is written
// (prop.acc) even with a single move sampler.
int planSweep(rp_ctx *c, bool logDecisions, SweepPlan &pl) {
    int r = prepareCycleGrid(c);
    if (r) return r;
    const Grid &g = c->grid;
    SweepParams &sp = pl.sp;
    sp = SweepParams{};
    sp.g = g;
    sp.p = c->cur.p;
    sp.k = c->k;
    // the samplers' steps in use: min(step, maximum step) for translations (TranslationSampler.cpp:29,
    // RototranslationSampler.cpp:31-33), min(step, pi) for rototranslation angles (RototranslationSampler.cpp:35);
    // a rotation sampler's step is capped at pi when it grows (RotationSampler.cpp:36-44)
    MoveSet ms{};
    ms.n = (int)c->samplers.size();
    for (int i = 0; i < ms.n; i++) {
        const auto &sm = c->samplers[i];
        const double maxT = sm.maxTransImplicit ? minHeight(c->box) / 2 : sm.maxTransStep;
        ms.kind[i] = sm.kind;
        ms.ts[i] = std::min(maxT, sm.transStep);
        ms.rs[i] = sm.kind == RP_MOVE_ROTATION ? sm.rotStep : std::min(sm.rotStep, M_PI);
    }
    for (int d = 0; d < 3; d++) sp.na[d] = std::max(1, g.n[d] / g.period);
    double lambda = (double)c->n / (double)g.ncell;
    // Dense cells make a fixed number of moves each (the reference's floor(N / domains) per domain,
    // Simulation.cpp:431-433), which balances the workgroups; sparse cells move as many times as they hold
    // particles, so a cycle is exactly N moves.
    // (with k samplers, k times as many: each requests N moves)
    sp.moves = lambda >= 3 ? (int)std::lround(lambda) * ms.n : 0;
    // without a grid: the reference's serial cycle, N moves per sampler, each on any particle (Simulation.cpp:365-445)
    if (g.allPairs) sp.moves = (int)c->n * ms.n;
    sp.mult = ms.n;
    pl.proposals = sp.moves > 0 ? (size_t)g.ncell * sp.moves : c->n * (size_t)ms.n;
    r = ensureProposals(c, pl.proposals);
    if (r) return r;
    sp.pr = c->prop;
    if (ms.n == 1) {
        sp.pr.type = nullptr;  // one sampler: the totals are its counters
        if (!logDecisions) sp.pr.acc = nullptr;
    }
    // the previous cycle's move counts are folded by the same launch
    const FoldSlots fold = c->slotsDirty ? FoldSlots{c->dSlots, c->slotsDirty, c->dCounters} : FoldSlots{};
    RP_HIP(launch_proposals(c->stream, g, sp.pr, sp.moves, c->cur.p.cell, (int)c->n, ms, c->seed,
                            (uint32_t)c->cycle, timingSlot(c, 3), fold));
    c->slotsDirty = 0;
    const int side = 2 * g.reach + 1;
    double expect = side * side * side * lambda;
    int cap = (int)std::ceil(expect + 8 * std::sqrt(expect) + 64);
    cap = (cap + 63) / 64 * 64;
    int waves = c->sweepWaves;
    // eight wavefronts with grouped moves for dense stencils (measured fastest on 1e5 rods), else fewer
    if (waves <= 0) waves = expect > 384 ? 8 : (expect > 128 ? 2 : 1);
    // polyspheres whose stencils hold over a thousand particles: sixteen (2.24e6 against 2.09e6 trial moves/s with
    // eight on 50 653 kinked chains)
    if (c->sweepWaves <= 0 && c->k.kind == RP_SHAPE_POLYSPHERE && expect > 1024) waves = 16;
    waves = waves >= 16 ? 16 : (waves >= 8 ? 8 : (waves >= 4 ? 4 : (waves == 2 ? 2 : 1)));
    // sixteen wavefronts only for the hard polysphere sweep (moves that scan thousands of stencil entries from HBM)
    if (waves == 16 && (c->k.kind != RP_SHAPE_POLYSPHERE || c->k.soft || c->counting)) waves = 8;
    // eight wavefronts only for the hard sweeps (polyspheres: 1.5x faster than four on 50 653 kinked chains, whose
    // moves scan up to 2 350 stencil entries from HBM)
    if (waves == 8 && (c->k.soft || c->counting)) waves = 4;
    // per-sampler acceptance flags are compiled into the kernels of at most four wavefronts only (the eight-wavefront
    // kernel has no registers to spare for them)
    if (waves >= 8 && ms.n > 1) waves = 4;
    sp.pipelined = c->pipelined >= 0 ? c->pipelined : (waves >= 8 ? 2 : 1);
    // guarded FP32 stencil scan: the default launch of hard spherocylinders (eight wavefronts, grouped moves); the
    // all-pairs path (no grid) tests in FP64
    const int g32 = c->precision != RP_PREC_FP64 && waves == 8 && sp.pipelined == 2 && !c->k.soft && !c->counting &&
                    c->k.kind == RP_SHAPE_SPHEROCYLINDER && !g.allPairs ? c->precision : 0;
    if (g32) g32Limits(c, g, sp.g32Lim);
    const double perEntry =
        (double)(sweep_lds_bytes(c->k.kind, 1024, waves, g32) - sweep_lds_bytes(c->k.kind, 0, waves, g32)) / 1024.0;
    const size_t fixed = sweep_lds_bytes(c->k.kind, 0, waves, g32);
    int maxCap = (int)((double)(c->ldsLimit - fixed) / perEntry) / 64 * 64;
    sp.cap = std::max(64, std::min(cap, maxCap));
    // workgroups that read their stencil from HBM so far (hCounters as of the last read)
    if (!c->pairCapOff && c->blocksAtRead >= 10000 && c->hCounters[CNT_UNSTAGED_BLOCKS] * 100 > c->blocksAtRead)
        c->pairCapOff = true;
    if (waves == 8 && !c->pairCapOff) {
        // two eight-wavefront workgroups share a compute unit (128 registers each) only if each takes at most half the
        // unit's LDS (ldsLimit is the unit's LDS less 4 KB; the kernel's static arrays take under 2 KB): a smaller
        // stencil area is worth it while it still holds the expected stencil with a margin of 3 standard deviations
        // (larger cells read HBM); states whose stencils vary more than that (layered smectics, whose cells cut
        // through one or two layers) overflow it, and once more than 1 % of the workgroups have, the full area is used
        const double half = (double)(c->ldsLimit + 4096) / 2 - 2048;
        const int pairCap = (int)((half - (double)fixed) / perEntry) / 64 * 64;
        if (pairCap < sp.cap && pairCap >= expect + 3 * std::sqrt(expect)) sp.cap = pairCap;
    }
    if (c->k.kind == RP_SHAPE_POLYSPHERE) sp.cap = 64;  // the polysphere sweep reads its stencil from HBM
    if (g.allPairs) sp.cap = allPairsStaged(c);
    r = ensureSlots(c, g.ncell);
    if (r) return r;
    sp.counters = c->dCounters;
    sp.slots = c->dSlots;
    sp.phases = c->dPhases;
    sp.tuning = c->sweepTuning;
    sp.temperature = c->temperature;
    sp.counting = c->counting && !c->k.soft ? 1 : 0;
    pl.waves = waves;
    pl.g32 = g32;
    pl.colours = g.period * g.period * g.period;
    for (int i = 0; i < pl.colours; i++) pl.order[i] = i;
    for (int i = pl.colours - 1; i > 0; i--) {
        int j = (int)philox_below(c->hostRng, (uint32_t)(i + 1));
        std::swap(pl.order[i], pl.order[j]);
    }
    c->tstats.lds_limit = c->ldsLimit;
    c->tstats.sweep_cap = (uint64_t)sp.cap;
    c->tstats.sweep_waves = (uint64_t)waves;
    c->tstats.sweep_period = (uint64_t)g.period;
    return RP_OK;
}

// The last part of a cycle's particle moves, after its colour passes: per-sampler counts and, in overlap-counting
// mode, the cached overlap count.
int finishSweep(rp_ctx *c, const SweepPlan &pl) {
    c->sortedForCycle = false;
    c->slotsDirty = pl.sp.g.ncell;  // folded by the next cycle's proposal launch or the next counter read
    RP_HIP(launch_type_counts(c->stream, pl.sp.pr, pl.proposals, c->dCounters));
    if (pl.sp.counting) {
        // Packing::acceptMove adds each accepted move's overlap change to the cached count (Packing.cpp:298-301)
        unsigned long long d = 0;
        RP_HIP(hipMemcpyAsync(&d, c->dCounters + CNT_OVERLAP_DELTA, sizeof d, hipMemcpyDeviceToHost, c->stream));
        RP_HIP(launch_fill_u64(c->stream, c->dCounters + CNT_OVERLAP_DELTA, 1, 0));
        RP_HIP(hipStreamSynchronize(c->stream));
        c->numOverlaps += (long long)d;
    }
    return RP_OK;
}

// The particle moves of one cycle: proposals, then the colour passes in random order.  logDecisions: every
// proposal's acceptance flag is written (prop.acc) even with a single move sampler.
int sweepCycle(rp_ctx *c, bool logDecisions, SweepInfo *info) {
    SweepPlan pl;
    int r = planSweep(c, logDecisions, pl);
    if (r) return r;
    if (info) {
        info->colours = pl.colours;
        std::memcpy(info->order, pl.order, sizeof pl.order);
        info->moves = pl.sp.moves;
        info->mult = pl.sp.mult;
        info->proposals = pl.proposals;
    }
    SweepParams sp = pl.sp;
    for (int i = 0; i < pl.colours; i++) {
        sp.colour = pl.order[i];
        if (!colourHasCells(sp.g, sp.colour)) continue;
        const bool trace = !c->tracePath.empty() && ++c->sweepLaunches == c->traceLaunch;
        if (trace) {
            r = traceSweep(c, sp, pl.waves);
            if (r) return r;
        } else {
            // sampled timing: the i-th colour pass of cycle c is timed when i == c mod (colour count)
            const bool timed = c->timing == 1 || (c->timing == RP_TIMING_SAMPLED && i == (int)(c->cycle % pl.colours));
            RP_HIP(launch_sweep(c->stream, sp, pl.waves, timed ? timingSlot(c, 0) : Timed{}, pl.g32));
        }
        c->blocksLaunched += (uint64_t)sp.na[0] * sp.na[1] * sp.na[2];
    }
    return finishSweep(c, pl);
}

// One cycle: the particle moves (colour passes), then one box move.
// The end of a cycle: box-move counters (when a box move was tried) and the cycle number.
void endCycle(rp_ctx *c, bool acc) {
    if (c->boxKind != RP_BOX_NONE) {
        c->boxAttempted++;
        c->boxSinceEval++;
        if (acc) {
            c->boxAccepted++;
            c->boxAccSinceEval++;
        }
    }
    c->cycle++;
    c->cycles++;
}

int runCycle(rp_ctx *c) {
    int r = sweepCycle(c, false, nullptr);
    if (r) return r;
    bool acc = false;
    if (c->boxKind != RP_BOX_NONE) {
        r = boxMove(c, acc);
        if (r) return r;
    }
    endCycle(c, acc);
    return RP_OK;
}

// Simulation::evaluateCounters during thermalisation (Simulation.cpp:533-607) with
// RototranslationSampler::increase/decreaseStepSize (RototranslationSampler.cpp:51-62, 87-93) and
// TriclinicBoxScaler (TriclinicBoxScaler.cpp:13-21).
int adaptSteps(rp_ctx *c) {
    if (c->cycles % 100 == 0) {
        int r = readCounters(c);
        if (r) return r;
        const bool several = c->samplers.size() > 1;
        for (size_t i = 0; i < c->samplers.size(); i++) {
            // one sampler: the move totals; several: each sampler's own counter (Simulation.cpp:561-607)
            const int ia = several ? CNT_TYPE_ACC + (int)i : CNT_ACCEPTED;
            const int it = several ? CNT_TYPE_ATT + (int)i : CNT_ATTEMPTED;
            uint64_t &accBase = several ? c->typeAccBase[i] : c->accBase;
            uint64_t &attBase = several ? c->typeAttBase[i] : c->attBase;
            const uint64_t acc = c->hCounters[ia] - accBase, att = c->hCounters[it] - attBase;
            accBase = c->hCounters[ia];
            attBase = c->hCounters[it];
            if (att == 0) continue;
            const double rate = (double)acc / (double)att;
            auto &sm = c->samplers[i];
            const double maxT = sm.maxTransImplicit ? minHeight(c->box) / 2 : sm.maxTransStep;
            if (rate > 0.2) {
                if (sm.kind == RP_MOVE_ROTATION) {
                    sm.rotStep = std::min(sm.rotStep * 1.1, M_PI);  // RotationSampler.cpp:36-44
                } else {
                    // TranslationSampler.cpp:48-54; RototranslationSampler.cpp:51-56 scales the rotation too
                    double old = sm.transStep;
                    if (!(maxT > 0 && sm.transStep >= maxT)) sm.transStep *= 1.1;
                    if (sm.kind == RP_MOVE_ROTOTRANSLATION) sm.rotStep *= sm.transStep / old;
                }
            } else if (rate < 0.1) {
                if (sm.kind != RP_MOVE_ROTATION) sm.transStep /= 1.1;
                if (sm.kind != RP_MOVE_TRANSLATION) sm.rotStep /= 1.1;
            }
        }
    }
    if (c->boxKind != RP_BOX_NONE && c->boxSinceEval >= 100) {
        double rate = (double)c->boxAccSinceEval / (double)c->boxSinceEval;
        c->boxSinceEval = c->boxAccSinceEval = 0;
        if (rate > 0.2) c->boxStep *= 1.1;
        else if (rate < 0.1) c->boxStep /= 1.1;
    }
    return RP_OK;
}

// Packing::countTotalOverlaps for polyspheres (Packing.cpp:519-559 over the grid of interaction centres,
// Packing.cpp:901-929): bead items of the particles `src` in box `box`, sorted by the reference's grid of bead range
// in the bead cell arrays (the sweep grid's arrays stay valid).  RP_ERR_UNSUPPORTED when the reference would count
// without a grid.
int countBeads(rp_ctx *c, const Particles &src, const HostBox &box, int earlyExit, uint64_t &out) {
    const size_t items = c->n * (size_t)c->k.nc;
    if (!c->beadA.mem) {
        int r = allocParticles(c->beadA, items);
        if (!r) r = allocParticles(c->beadB, items);
        if (r) return r;
        RP_HIP(hipMemsetAsync(c->beadA.mem, 0, c->beadA.bytes, c->stream));
        RP_HIP(hipMemsetAsync(c->beadB.mem, 0, c->beadB.bytes, c->stream));
    }
    Grid g{};
    int disabled = 0;
    int r = makeExactGrid(c, box, g, disabled, true);
    if (r) return r;
    if (disabled)
        return fail(RP_ERR_UNSUPPORTED, "the reference would count without a neighbour grid for this box; "
                                        "the all-pairs path is not implemented on the device");
    RP_HIP(launch_centres(c->stream, src, (int)c->n, c->k, g, c->beadA.p));
    RP_HIP(launch_rebuild(c->stream, g, (int)items, c->beadA.p, c->beadB.p, c->tmpSlot, c->tmpOrder, c->tmpPos, 0,
                          nullptr, timingSlot(c, 1), c->scanTmp));
    r = scanItems(c, g, c->beadB.p, items, earlyExit, out);
    if (r) return r;
    if (earlyExit && out) out = 1;
    return RP_OK;
}

int countExact(rp_ctx *c, int earlyExit, uint64_t &out) {
    if (c->k.kind == RP_SHAPE_POLYSPHERE) return countBeads(c, c->cur.p, c->box, earlyExit, out);
    c->sortedForCycle = false;  // the shared cell arrays are about to describe another grid
    Grid g{};
    int disabled = 0;
    int r = makeExactGrid(c, c->box, g, disabled);
    if (r) return r;
    if (disabled && noGridBox(c, c->box)) {
        // the reference counts every pair through the nearest image (Packing.cpp:547-556)
        r = makeAllPairsGrid(c, c->box, g);
        if (r) return r;
    } else if (disabled) {
        return fail(RP_ERR_UNSUPPORTED, "the reference would count without a neighbour grid for this box; "
                                        "the all-pairs path covers hard spheres and spherocylinders");
    }
    r = rebuild(c, g, 0, nullptr);
    if (r) return r;
    r = scanOverlaps(c, g, earlyExit, out);
    if (r) return r;
    if (earlyExit && out) out = 1;
    return RP_OK;
}


// Simulation::integrate's check before the first cycle (Simulation.cpp:127-132): no overlap on the reference's grid,
// or on the checkerboard grid where the reference's grid is off; nothing to check for a purely soft interaction or in
// overlap-counting mode.
int validateStart(rp_ctx *c) {
    uint64_t overl = 0;
    int r = RP_OK;
    if ((!c->k.soft || c->k.hardCore) && !c->counting) r = countExact(c, 1, overl);
    if (r == RP_ERR_UNSUPPORTED) {
        double zero[3] = {0, 0, 0};
        c->sortedForCycle = false;
        Grid g{};
        r = chooseSweepGrid(c, c->box, zero, g);
        if (r) return r;
        r = rebuild(c, g, 0, nullptr);
        if (r) return r;
        r = scanOverlaps(c, g, 1, overl);
    }
    if (r) return r;
    if (overl)
        return fail(RP_ERR_VALIDATION,
                    "Overlaps are present at the start of integration. Perform overlap reduction beforehand.");
    return RP_OK;
}

// After each cycle: the periodic orientation renormalisation (Simulation.cpp:144-145; in counting mode the
// reference's renormalisation keeps the orientations, Packing.cpp:1179-1190, so none is run) and the step
// adaptation.
int afterCycle(rp_ctx *c, bool adapt) {
    if (c->fixEvery && !c->counting && c->cycles % c->fixEvery == 0) {
        int r = rp_fix_orientations(c, nullptr);
        if (r) return r;
    }
    return adapt ? adaptSteps(c) : RP_OK;
}

int reportCounters(rp_ctx *c, rp_counters *out) {
    int r = readCounters(c);
    if (r) return r;
    if (!out) return RP_OK;
    out->cycles = c->cycles;
    out->moves_attempted = c->hCounters[CNT_ATTEMPTED];
    out->moves_accepted = c->hCounters[CNT_ACCEPTED];
    out->moves_left_cell = c->hCounters[CNT_LEFT_CELL];
    out->box_attempted = c->boxAttempted;
    out->box_accepted = c->boxAccepted;
    out->box_scans = c->boxScans;
    out->box_too_small = c->boxTooSmall;
    out->unstaged_blocks = c->hCounters[CNT_UNSTAGED_BLOCKS];
    out->pair_tests = c->hCounters[CNT_TESTS];
    // the first sampler that translates and the first that rotates
    out->trans_step = out->rot_step = 0;
    for (auto it = c->samplers.rbegin(); it != c->samplers.rend(); ++it) {
        const double maxT = it->maxTransImplicit ? minHeight(c->box) / 2 : it->maxTransStep;
        if (it->kind != RP_MOVE_ROTATION) out->trans_step = std::min(maxT, it->transStep);
        if (it->kind != RP_MOVE_TRANSLATION) out->rot_step = it->rotStep;
    }
    out->box_step = c->boxStep;
    return RP_OK;
}

// Device and pinned host memory of a replica run: the launch tables of one cycle's colour passes.
struct BatchMem {
    SweepParams *dTable = nullptr, *hTable = nullptr;
    int *dFirst = nullptr, *hFirst = nullptr;
    size_t tableCap = 0, firstCap = 0;
    hipEvent_t uploaded = nullptr;
    ~BatchMem() {
        if (dTable) hipFree(dTable);
        if (hTable) hipHostFree(hTable);
        if (dFirst) hipFree(dFirst);
        if (hFirst) hipHostFree(hFirst);
        if (uploaded) hipEventDestroy(uploaded);
    }
    int reserve(size_t tables, size_t firsts) {
        if (tables > tableCap) {
            if (dTable) hipFree(dTable);
            if (hTable) hipHostFree(hTable);
            dTable = hTable = nullptr;
            tableCap = 0;
            RP_HIP(hipMalloc(&dTable, sizeof(SweepParams) * tables));
            RP_HIP(hipHostMalloc(&hTable, sizeof(SweepParams) * tables, hipHostMallocDefault));
            tableCap = tables;
        }
        if (firsts > firstCap) {
            if (dFirst) hipFree(dFirst);
            if (hFirst) hipHostFree(hFirst);
            dFirst = hFirst = nullptr;
            firstCap = 0;
            RP_HIP(hipMalloc(&dFirst, sizeof(int) * firsts));
            RP_HIP(hipHostMalloc(&hFirst, sizeof(int) * firsts, hipHostMallocDefault));
            firstCap = firsts;
        }
        return RP_OK;
    }
};

// Every context's work on one stream for the run (the first replica's): one host thread orders all of it, and one
// synchronisation per cycle decides every replica's box move.  The contexts' own streams are restored at the end.
struct SharedStream {
    std::vector<rp_ctx *> ctxs;
    std::vector<hipStream_t> own;
    std::vector<int> timing;
    ~SharedStream() {
        if (!ctxs.empty()) hipStreamSynchronize(ctxs[0]->stream);
        for (size_t i = 0; i < ctxs.size(); i++) {
            ctxs[i]->stream = own[i];
            ctxs[i]->timing = timing[i];
        }
    }
};

// One colour pass's launches: the packings whose pass i has cells, grouped by kernel (sweep_batch_key), each group in
// one launch; packings that run one per launch (soft, counting, polysphere, traced) are launched on their own.
struct PassGroup {
    int key;
    size_t entry0, entries, first0;
    unsigned blocks;
    size_t lds;
};

int runReplicas(rp_ctx *const *ctxs, size_t n, size_t nCycles, bool adapt, rp_counters *out) {
    SharedStream shared;
    for (size_t i = 0; i < n; i++) RP_HIP(hipStreamSynchronize(ctxs[i]->stream));
    for (size_t i = 0; i < n; i++) {
        shared.ctxs.push_back(ctxs[i]);
        shared.own.push_back(ctxs[i]->stream);
        shared.timing.push_back(ctxs[i]->timing);
        ctxs[i]->stream = ctxs[0]->stream;
        ctxs[i]->timing = 0;  // per-launch timing events belong to single-packing runs
    }
    const hipStream_t s = ctxs[0]->stream;
    auto named = [&](size_t i, int r) {
        return r ? fail(r, "replica " + std::to_string(i) + ": " + rp_last_error()) : RP_OK;
    };
    for (size_t i = 0; i < n; i++) {
        int r = validateStart(ctxs[i]);
        if (r) return named(i, r);
    }
    BatchMem mem;
    RP_HIP(hipEventCreateWithFlags(&mem.uploaded, hipEventDisableSystemFence));
    bool uploadedOnce = false;
    std::vector<SweepPlan> plans(n);
    std::vector<BoxTrial> trials(n);
    std::vector<char> acc(n);
    std::vector<PassGroup> groups;
    std::vector<std::pair<int, size_t>> members;  // (key, replica) of one pass
    for (size_t cyc = 0; cyc < nCycles; cyc++) {
        int maxColours = 0;
        for (size_t i = 0; i < n; i++) {
            int r = planSweep(ctxs[i], false, plans[i]);
            if (r) return named(i, r);
            maxColours = std::max(maxColours, plans[i].colours);
        }
        // the staging tables are rewritten once the previous cycle's upload has completed
        if (uploadedOnce) RP_HIP(hipEventSynchronize(mem.uploaded));
        {
            int r = mem.reserve((size_t)maxColours * n, (size_t)maxColours * 2 * n);
            if (r) return r;
        }
        groups.clear();
        size_t entries = 0, firsts = 0;
        for (int pass = 0; pass < maxColours; pass++) {
            members.clear();
            for (size_t i = 0; i < n; i++) {
                const SweepPlan &pl = plans[i];
                if (pass >= pl.colours || !colourHasCells(pl.sp.g, pl.order[pass])) continue;
                members.emplace_back(sweep_batch_key(pl.sp, pl.waves, pl.g32), i);
                ctxs[i]->blocksLaunched += (uint64_t)pl.sp.na[0] * pl.sp.na[1] * pl.sp.na[2];
            }
            std::stable_sort(members.begin(), members.end(),
                             [](const std::pair<int, size_t> &a, const std::pair<int, size_t> &b) {
                                 return a.first < b.first;
                             });
            for (size_t m = 0; m < members.size();) {
                const int key = members[m].first;
                size_t e = m;
                while (e < members.size() && members[e].first == key) e++;
                if (key < 0) {
                    // one launch per packing, in pass order (launched below with the batched groups of this pass)
                    for (size_t q = m; q < e; q++) {
                        PassGroup pg{-1 - (int)members[q].second, (size_t)pass, 0, 0, 0, 0};
                        groups.push_back(pg);
                    }
                } else {
                    PassGroup pg{key, entries, e - m, firsts, 0, 0};
                    mem.hFirst[firsts] = 0;
                    for (size_t q = m; q < e; q++) {
                        const size_t i = members[q].second;
                        SweepParams sp = plans[i].sp;
                        sp.colour = plans[i].order[pass];
                        mem.hTable[entries++] = sp;
                        pg.blocks += (unsigned)(sp.na[0] * sp.na[1] * sp.na[2]);
                        mem.hFirst[++firsts] = (int)pg.blocks;
                        pg.lds = std::max(pg.lds, sweep_lds_bytes(sp.k.kind, sp.cap, plans[i].waves, plans[i].g32));
                    }
                    firsts++;
                    groups.push_back(pg);
                }
                m = e;
            }
            groups.push_back(PassGroup{INT32_MIN, (size_t)pass, 0, 0, 0, 0});  // end of pass
        }
        if (entries) {
            RP_HIP(hipMemcpyAsync(mem.dTable, mem.hTable, sizeof(SweepParams) * entries, hipMemcpyHostToDevice, s));
            RP_HIP(hipMemcpyAsync(mem.dFirst, mem.hFirst, sizeof(int) * firsts, hipMemcpyHostToDevice, s));
        }
        RP_HIP(hipEventRecord(mem.uploaded, s));
        uploadedOnce = true;
        int pass = 0;
        for (const PassGroup &pg : groups) {
            if (pg.key == INT32_MIN) {
                pass++;
                continue;
            }
            if (pg.key < 0) {
                const size_t i = (size_t)(-1 - pg.key);
                SweepParams sp = plans[i].sp;
                sp.colour = plans[i].order[pass];
                RP_HIP(launch_sweep(s, sp, plans[i].waves, Timed{}, plans[i].g32));
                continue;
            }
            const int packings = (int)pg.entries;
            RP_HIP(launch_sweep_batch(s, pg.key, mem.dTable + pg.entry0, mem.dFirst + pg.first0, packings, pg.blocks,
                                      pg.lds));
        }
        bool pending = false;
        for (size_t i = 0; i < n; i++) {
            int r = finishSweep(ctxs[i], plans[i]);
            if (r) return named(i, r);
            bool a = false;
            trials[i] = BoxTrial{};
            if (ctxs[i]->boxKind != RP_BOX_NONE) {
                r = boxMove(ctxs[i], a, &trials[i]);
                if (r) return named(i, r);
                pending = pending || trials[i].pending;
            }
            acc[i] = a;
        }
        if (pending) RP_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < n; i++) {
            bool a = acc[i] != 0;
            int r = boxMoveEnd(ctxs[i], trials[i], a);
            if (r) return named(i, r);
            endCycle(ctxs[i], a);
            r = afterCycle(ctxs[i], adapt);
            if (r) return named(i, r);
        }
    }
    for (size_t i = 0; i < n; i++) {
        int r = reportCounters(ctxs[i], out ? out + i : nullptr);
        if (r) return named(i, r);
    }
    return RP_OK;
}
}  // namespace

extern "C" {

int rp_ramsnap_save(rp_ctx *c, const char *path, const char *aux_text) {
    if (!c || !path) return fail(RP_ERR_PRECONDITION, "null argument");
    std::vector<double> box(9), pos(3 * c->n), rot(9 * c->n);
    int r = rp_download(c, box.data(), pos.data(), rot.data());
    if (r) return r;
    return rp_ramsnap_write(path, box.data(), c->n, pos.data(), rot.data(), aux_text);
}

int rp_ramtrj_record_packing(rp_ramtrj_recorder *r, rp_ctx *c, size_t cycle) {
    if (!r || !c) return fail(RP_ERR_PRECONDITION, "null argument");
    std::vector<double> box(9), pos(3 * c->n), rot(9 * c->n);
    int e = rp_download(c, box.data(), pos.data(), rot.data());
    if (e) return e;
    return rp_ramtrj_record(r, box.data(), c->n, pos.data(), rot.data(), cycle);
}

const char *rp_last_error(void) { return g_lastError.c_str(); }


int rp_create(rp_ctx **out, const rp_shape_desc *shape, const double box[9], size_t n, const double *pos,
              const double *rot, uint64_t seed, int precision) {
    if (!out || !shape || !box || !pos || !rot) return fail(RP_ERR_PRECONDITION, "null argument");
    *out = nullptr;
    if (n == 0) return fail(RP_ERR_PRECONDITION, "empty packing (Packing.cpp:69)");
    if (n > (size_t)0x3fffffff) return fail(RP_ERR_PRECONDITION, "too many particles");
    if (precision != RP_PREC_FP64 && precision != RP_PREC_GUARDED_FP32 && precision != RP_PREC_GUARDED_FP32_VERIFY)
        return fail(RP_ERR_PRECONDITION, "unknown precision");
    if (int v = validateShape(*shape)) return v;
    HostBox hb;
    hb.set(box);
    if (hb.volume() == 0) return fail(RP_ERR_PRECONDITION, "box volume is 0 (Packing.cpp:68)");
    for (size_t i = 0; i < n; i++) {
        V3 rel = mulv(hb.inv, v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
        if (rel.x < 0 || rel.x >= 1 || rel.y < 0 || rel.y >= 1 || rel.z < 0 || rel.z >= 1)
            return fail(RP_ERR_PRECONDITION, "particle " + std::to_string(i) + " lies outside the box (Packing.cpp:71)");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(RP_ERR_DEVICE, "no HIP device");

    rp_ctx *c = new rp_ctx();
    c->device = dev;
    int ldsBytes = 0;
    if (hipDeviceGetAttribute(&ldsBytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && ldsBytes > 0)
        // dynamic LDS a workgroup may use: the device limit less room for the kernels' static LDS tables (stencil
        // cell tables, bead tables, vote slots: under 4 KB)
        c->ldsLimit = (size_t)std::min(ldsBytes, 160 * 1024) - 4096;
    if (const char *w = std::getenv("RAMPACK_SWEEP_WAVES")) c->sweepWaves = std::atoi(w);
    if (const char *p = std::getenv("RAMPACK_COLOURING")) c->colourPeriod = std::atoi(p);
    if (const char *p = std::getenv("RAMPACK_SWEEP_PIPELINE")) c->pipelined = std::atoi(p);
    if (const char *t = std::getenv("RAMPACK_SWEEP_TUNING")) c->sweepTuning = std::atoi(t);
    if (const char *tr = std::getenv("RAMPACK_SWEEP_TRACE")) {
        // path:launch -- the workgroups of the launch-th sweep launch of this context, one line each
        std::string v(tr);
        size_t colon = v.rfind(':');
        if (colon != std::string::npos && colon > 0) {
            c->tracePath = v.substr(0, colon);
            c->traceLaunch = std::strtoull(v.c_str() + colon + 1, nullptr, 10);
        }
    }
    if (const char *ph = std::getenv("RAMPACK_SWEEP_PHASES"); ph && std::atoi(ph) != 0) {
        unsigned long long init[RP_PHASE_NUM] = {};
        if (hipMalloc(&c->dPhases, sizeof(unsigned long long) * RP_PHASE_NUM) != hipSuccess ||
            hipMemcpy(c->dPhases, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess) {
            rp_destroy(c);
            return fail(RP_ERR_DEVICE, "phase timer set-up failed");
        }
    }
    c->shape = *shape;
    c->shape.centres = nullptr;  // the caller's arrays are not kept; the device copy is k.beads
    c->shape.radii = nullptr;
    c->k = makeConsts(*shape);
    c->range = shapeRange(*shape);
    c->totalRange = shapeTotalRange(*shape);
    c->n = n;
    c->box = hb;
    c->precision = precision;
    c->seed = seed;
    c->hostRng = philox_init(seed, 0xA5A5A5A5u, 0x0B0Bu, 0x1u);
    auto cleanup = [&](int code) {
        rp_destroy(c);
        return code;
    };
    if (hipStreamCreate(&c->stream) != hipSuccess) return cleanup(fail(RP_ERR_DEVICE, "hipStreamCreate failed"));
    int r = uploadBeads(*shape, &c->dBeads);
    if (r) return cleanup(r);
    c->k.beads = c->dBeads;
    c->k.nc = shape->kind == RP_SHAPE_POLYSPHERE ? shape->n_centres : 0;
    // polysphere sweep prefilter: particles whose centres are at least the total range apart share no bead
    // contact; the 1e-9 margin keeps every pair whose rounded bead distances could decide otherwise
    if (shape->kind == RP_SHAPE_POLYSPHERE) c->k.range2 = std::pow(c->totalRange * (1 + 1e-9), 2);
    if (shape->kind == RP_SHAPE_POLYSPHERE) c->k.bound = shapeBound(*shape);
    if (shape->kind == RP_SHAPE_POLYSPHERE) {
        // PolysphereGeometry::calculateVolume (PolysphereTraits.cpp:138-146) needs beads that do not overlap by
        // spheresOverlap's test (PolysphereTraits.cpp:211-224): distance^2 (1 + 1e-12) < (r_i + r_j)^2
        double v = 0;
        bool overlapping = false;
        for (int i = 0; i < shape->n_centres; i++) {
            const double r = shape->radii[i];
            v = v + 4 * M_PI / 3 * r * r * r;
            for (int j = i + 1; j < shape->n_centres; j++) {
                double d2 = 0;
                for (int q = 0; q < 3; q++) {
                    double d = shape->centres[3 * i + q] - shape->centres[3 * j + q];
                    d2 += d * d;
                }
                overlapping = overlapping || d2 * (1e-12 + 1) < std::pow(r + shape->radii[j], 2);
            }
        }
        c->beadVolume = overlapping ? std::nan("") : v;
    }
    const size_t items = n * (size_t)std::max(1, c->k.nc);  // rebuild scratch: particles or bead items
    r = allocParticles(c->cur, n);
    if (r) return cleanup(r);
    r = allocParticles(c->alt, n);
    if (r) return cleanup(r);
    if (hipMalloc(&c->tmpSlot, sizeof(int) * 2 * items) != hipSuccess ||
        hipMalloc(&c->tmpOrder, sizeof(int) * items + items + 256) != hipSuccess ||
        hipMalloc(&c->tmpPos, sizeof(double) * 3 * items) != hipSuccess ||
        hipMalloc(&c->dCounters, sizeof(unsigned long long) * CNT_NUM) != hipSuccess ||
        hipHostMalloc(&c->hCounters, sizeof(unsigned long long) * CNT_NUM, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(RP_ERR_DEVICE, "device allocation failed"));

    // upload as SoA in the caller's order
    std::vector<double> col(n);
    std::vector<int> tags(n);
    auto up = [&](double *dst, const double *src, int stride, int off) -> int {
        for (size_t i = 0; i < n; i++) col[i] = src[stride * i + off];
        RP_HIP(hipMemcpy(dst, col.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        return RP_OK;
    };
    if ((r = up(c->cur.p.x, pos, 3, 0)) || (r = up(c->cur.p.y, pos, 3, 1)) || (r = up(c->cur.p.z, pos, 3, 2)))
        return cleanup(r);
    for (int k = 0; k < 9; k++)
        if ((r = up(c->cur.p.o[k], rot, 9, k))) return cleanup(r);
    for (size_t i = 0; i < n; i++) tags[i] = (int)i;
    if (hipMemcpy(c->cur.p.tag, tags.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(RP_ERR_DEVICE, "upload failed"));
    if (launch_fill_u64(c->stream, c->dCounters, CNT_NUM, 0) != hipSuccess)
        return cleanup(fail(RP_ERR_DEVICE, "counter init failed"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(RP_ERR_DEVICE, "sync failed"));
    *out = c;
    return RP_OK;
}

int rp_set_environment(rp_ctx *c, double temperature, double pressure, const rp_move_desc *move,
                       const rp_box_desc *box) {
    if (!c || !move) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!(temperature > 0)) return fail(RP_ERR_PRECONDITION, "temperature must be > 0");
    const int boxKind = box ? box->kind : RP_BOX_NONE;
    if (boxKind != RP_BOX_NONE && boxKind != RP_BOX_DELTA_TRICLINIC && boxKind != RP_BOX_LOG_ISOTROPIC &&
        boxKind != RP_BOX_DELTA_V && boxKind != RP_BOX_LINEAR)
        return fail(RP_ERR_UNSUPPORTED, "unknown box move");
    if (boxKind != RP_BOX_NONE && !(box->step > 0))
        return fail(RP_ERR_PRECONDITION, "box step must be > 0 (TriclinicBoxScaler.cpp:9-11)");
    if (boxKind != RP_BOX_NONE)
        for (int a = 0; a < 3; a++)
            if (box->groups[a] < 0 || box->groups[a] > 3)
                return fail(RP_ERR_PRECONDITION, "box scaling groups are 0, 1, 2 or 3 (not scaled)");
    int r = rp_set_move_samplers(c, move, 1);
    if (r) return r;
    c->temperature = temperature;
    c->pressure = pressure;
    c->boxKind = boxKind;
    if (boxKind != RP_BOX_NONE) {
        c->boxStep = box->step;
        for (int a = 0; a < 3; a++) c->boxGroups[a] = box->groups[a];
        c->boxIndependent = box->independent ? 1 : 0;
        c->boxOneSide = box->one_side ? 1 : 0;
    }
    c->envSet = true;
    return RP_OK;
}

int rp_set_move_samplers(rp_ctx *c, const rp_move_desc *moves, size_t n) {
    if (!c || !moves) return fail(RP_ERR_PRECONDITION, "null argument");
    if (n < 1 || n > RP_MAX_MOVE_SAMPLERS)
        return fail(RP_ERR_PRECONDITION, "1 to " + std::to_string(RP_MAX_MOVE_SAMPLERS) + " move samplers");
    std::vector<rp_ctx::Sampler> list;
    for (size_t i = 0; i < n; i++) {
        const rp_move_desc &m = moves[i];
        rp_ctx::Sampler sm;
        sm.kind = m.kind;
        if (m.kind == RP_MOVE_ROTATION) {
            if (!(m.rot_step > 0)) return fail(RP_ERR_PRECONDITION, "rotation step must be > 0 (RotationSampler.cpp:10)");
            sm.rotStep = m.rot_step;
            sm.transStep = 0;
        } else if (m.kind == RP_MOVE_TRANSLATION || m.kind == RP_MOVE_ROTOTRANSLATION) {
            // TranslationSampler.cpp:12-18, RototranslationSampler.cpp:14-16
            if (!(m.trans_step > 0)) return fail(RP_ERR_PRECONDITION, "translation step must be > 0");
            if (m.max_trans_step < 0 || (m.max_trans_step > 0 && m.trans_step > m.max_trans_step))
                return fail(RP_ERR_PRECONDITION, "invalid maximal translation step (RototranslationSampler.cpp:14-16)");
            sm.transStep = m.trans_step;
            sm.maxTransStep = m.max_trans_step;
            sm.maxTransImplicit = m.max_trans_step == 0;
            if (m.kind == RP_MOVE_TRANSLATION) {
                sm.rotStep = 0;
            } else if (m.rot_step > 0) {
                sm.rotStep = m.rot_step;
            } else {
                // RototranslationSampler::setupForShapeTraits (RototranslationSampler.cpp:95-114)
                double ratio = sm.transStep / c->totalRange;
                sm.rotStep = 0.133183 - 1.18634 * ratio + 264.37 * ratio * ratio;
            }
        } else {
            return fail(RP_ERR_UNSUPPORTED, "unknown move sampler");
        }
        list.push_back(sm);
    }
    c->samplers = list;
    if (c->dCounters) {
        // the per-sampler counters restart with the new list
        int r = readCounters(c);
        if (r) return r;
        for (int i = 0; i < RP_MAX_MOVE_SAMPLERS; i++) {
            c->typeAccBase[i] = c->hCounters[CNT_TYPE_ACC + i];
            c->typeAttBase[i] = c->hCounters[CNT_TYPE_ATT + i];
        }
    }
    return RP_OK;
}

int rp_get_move_samplers(rp_ctx *c, rp_move_desc *out, size_t n, size_t *nMoves) {
    if (!c || (n && !out)) return fail(RP_ERR_PRECONDITION, "null argument");
    if (nMoves) *nMoves = c->samplers.size();
    for (size_t i = 0; i < n && i < c->samplers.size(); i++) {
        const auto &sm = c->samplers[i];
        out[i] = rp_move_desc{};
        out[i].kind = sm.kind;
        out[i].trans_step = sm.transStep;
        out[i].rot_step = sm.rotStep;
        out[i].max_trans_step = sm.maxTransImplicit ? 0 : sm.maxTransStep;
    }
    return RP_OK;
}

// nCycles cycles of a validated packing (the body of rp_run_cycles after its start check).
static int cycleLoop(rp_ctx *c, size_t nCycles, bool adapt) {
    for (size_t i = 0; i < nCycles; i++) {
        int r = runCycle(c);
        if (r) return r;
        r = afterCycle(c, adapt);
        if (r) return r;
        if (c->timing && c->evUsed >= 8192) {  // bounded event pool; otherwise read once at the end
            r = timingHarvest(c);
            if (r) return r;
        }
    }
    return timingHarvest(c);
}

int rp_run_cycles(rp_ctx *c, size_t nCycles, int adapt, rp_counters *out) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    int r = validateStart(c);
    if (r) return r;
    r = cycleLoop(c, nCycles, adapt != 0);
    if (r) return r;
    return reportCounters(c, out);
}

int rp_run_replicas(rp_ctx *const *ctxs, size_t n, size_t nCycles, int adapt, rp_counters *out) {
    if (n && !ctxs) return fail(RP_ERR_PRECONDITION, "null context list");
    for (size_t i = 0; i < n; i++) {
        if (!ctxs[i]) return fail(RP_ERR_PRECONDITION, "replica " + std::to_string(i) + ": null context");
        for (size_t j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i])
                return fail(RP_ERR_PRECONDITION, "replicas " + std::to_string(j) + " and " + std::to_string(i) +
                                                     " are the same context");
    }
    for (size_t i = 0; i < n; i++) {
        if (!ctxs[i]->envSet) return fail(RP_ERR_PRECONDITION, "replica " + std::to_string(i) + ": environment not set");
        if (ctxs[i]->device != ctxs[0]->device)
            return fail(RP_ERR_PRECONDITION, "replica " + std::to_string(i) + " lives on another device");
    }
    if (n == 0) return RP_OK;
    return runReplicas(ctxs, n, nCycles, adapt != 0, out);
}

// Sorted device records (storage order of the current cell list) into host arrays; null arrays are skipped.
static int downloadSorted(rp_ctx *c, double *pos, double *rot, int *tag, int *cell) {
    const size_t n = c->n;
    std::vector<double> col(n);
    auto down = [&](const double *src, double *dst, int stride, int off) -> int {
        RP_HIP(hipMemcpy(col.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) dst[stride * i + off] = col[i];
        return RP_OK;
    };
    int r;
    if (pos && ((r = down(c->cur.p.x, pos, 3, 0)) || (r = down(c->cur.p.y, pos, 3, 1)) ||
                (r = down(c->cur.p.z, pos, 3, 2))))
        return r;
    if (rot)
        for (int k = 0; k < 9; k++)
            if ((r = down(c->cur.p.o[k], rot, 9, k))) return r;
    if (tag) RP_HIP(hipMemcpy(tag, c->cur.p.tag, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (cell) RP_HIP(hipMemcpy(cell, c->cur.p.cell, sizeof(int) * n, hipMemcpyDeviceToHost));
    return RP_OK;
}

int rp_record_sweep(rp_ctx *c, size_t capacity, rp_sweep_record *rec, double *prePos, double *preRot, int *preTag,
                    int *preCell, int *propK, double *propT, double *propRot, uint8_t *accepted, double *postPos,
                    double *postRot) {
    if (!c || !rec) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    if (c->k.soft || c->counting)
        return fail(RP_ERR_PRECONDITION, "decision records cover hard interactions outside overlap counting");
    int r = prepareCycleGrid(c);  // idempotent until the cycle runs: a too-small call leaves the grid for the next
    if (r) return r;
    *rec = rp_sweep_record{};
    std::memcpy(rec->box, c->box.dims, sizeof rec->box);
    for (int d = 0; d < 3; d++) {
        rec->shift[d] = c->grid.shift[d];
        rec->cells[d] = c->grid.n[d];
    }
    rec->period = c->grid.period;
    rec->n_samplers = (int)c->samplers.size();
    rec->n_proposals = cycleProposals(c);
    if (capacity < rec->n_proposals)
        return fail(RP_ERR_PRECONDITION, "capacity " + std::to_string(capacity) + " below the cycle's " +
                                             std::to_string(rec->n_proposals) + " proposals");
    if (!prePos || !preRot || !preTag || !preCell || !propK || !propT || !propRot || !accepted || !postPos ||
        !postRot)
        return fail(RP_ERR_PRECONDITION, "null record array");
    RP_HIP(hipStreamSynchronize(c->stream));
    r = downloadSorted(c, prePos, preRot, preTag, preCell);
    if (r) return r;
    SweepInfo info;
    r = sweepCycle(c, true, &info);
    if (r) return r;
    RP_HIP(hipStreamSynchronize(c->stream));
    if (info.proposals != rec->n_proposals) return fail(RP_ERR_DEVICE, "proposal count changed");
    rec->n_colours = info.colours;
    std::memcpy(rec->colour_order, info.order, sizeof info.order);
    rec->moves_per_cell = info.moves;
    const size_t np = info.proposals, stride = (size_t)c->prop.stride;
    RP_HIP(hipMemcpy(propK, c->prop.k, sizeof(int) * np, hipMemcpyDeviceToHost));
    RP_HIP(hipMemcpy(accepted, c->prop.acc, np, hipMemcpyDeviceToHost));
    std::vector<double> col(np);
    for (int d = 0; d < 3; d++) {
        RP_HIP(hipMemcpy(col.data(), c->prop.t + d * stride, sizeof(double) * np, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < np; i++) propT[3 * i + d] = col[i];
    }
    for (int e = 0; e < 9; e++) {
        RP_HIP(hipMemcpy(col.data(), c->prop.R + e * stride, sizeof(double) * np, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < np; i++) propRot[9 * i + e] = col[i];
    }
    r = downloadSorted(c, postPos, postRot, nullptr, nullptr);
    if (r) return r;
    // a cycle without its box move: the next cycle draws new proposals and a new grid origin
    c->cycle++;
    c->cycles++;
    return RP_OK;
}

int rp_count_total_overlaps(rp_ctx *c, int earlyExit, uint64_t *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    return countExact(c, earlyExit, *out);
}

namespace {

// NematicOrder::calculateEigenvalues (NematicOrder.cpp:31-54, the Cardano method of arXiv:1306.6291) and the
// largest-magnitude pick of NematicOrder::calculate (NematicOrder.cpp:22-27).  A Q tensor proportional to the
// identity (p = 0, where the reference asserts) gives 0.
double nematicP2(const double Q[9]) {
    double A11 = Q[0], A12 = Q[1], A13 = Q[2], A22 = Q[4], A23 = Q[5], A33 = Q[8];
    double b = A11 + A22 + A33;
    double c = A11 * A22 + A11 * A33 + A22 * A33 - A12 * A12 - A13 * A13 - A23 * A23;
    double d = A11 * A23 * A23 + A22 * A13 * A13 + A33 * A12 * A12 - A11 * A22 * A33 - 2 * A12 * A13 * A23;
    double p = b * b - 3 * c;
    if (!(p > 0)) return 0;
    double sp = std::sqrt(p);
    double q = 2 * b * b * b - 9 * b * c - 27 * d;
    double ratio = q / 2 / (sp * sp * sp);
    if (std::fabs(ratio) > 1) ratio = ratio < 0 ? -1 : 1;
    double delta = std::acos(ratio);
    double e[3] = {1. / 3 * (b + 2 * sp * std::cos(delta / 3)), 1. / 3 * (b + 2 * sp * std::cos((delta + 2 * M_PI) / 3)),
                   1. / 3 * (b + 2 * sp * std::cos((delta - 2 * M_PI) / 3))};
    double best = e[0];
    for (int i = 1; i < 3; i++)
        if (std::fabs(best) < std::fabs(e[i])) best = e[i];
    return best;
}

// ShapeGeometry::getVolume: SphereTraits.cpp:31-33, SpherocylinderTraits.cpp:48-50
double shapeVolume(const rp_shape_desc &s) {
    if (s.kind == RP_SHAPE_SPHERE) return 4. / 3 * M_PI * std::pow(s.radius, 3);
    return M_PI * s.radius * s.radius * s.length + 4. / 3 * M_PI * std::pow(s.radius, 3);
}

}  // namespace

int rp_get_observables(rp_ctx *c, rp_observables *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    *out = rp_observables{};
    // Packing::getPackingFraction / getNumberDensity (Packing.cpp:225-232)
    out->volume = c->box.volume();
    out->number_density = (double)c->n / out->volume;
    out->packing_fraction =
        out->number_density * (c->shape.kind == RP_SHAPE_POLYSPHERE ? c->beadVolume : shapeVolume(c->shape));
    if (c->shape.kind != RP_SHAPE_SPHEROCYLINDER) {
        // a sphere has no primary axis (ShapeGeometry::getPrimaryAxis is not provided by SphereTraits), and a
        // polysphere given by its beads alone has none either (the axis of PolysphereGeometry is optional)
        out->nematic_p2 = std::nan("");
        for (double &q : out->q_tensor) q = std::nan("");
        return RP_OK;
    }
    if (!c->obsMem) RP_HIP(hipMalloc(&c->obsMem, sizeof(double) * nematic_scratch_doubles()));
    RP_HIP(launch_nematic_sums(c->stream, c->cur.p, (int)c->n, c->obsMem));
    double sums[6];
    RP_HIP(hipMemcpyAsync(sums, c->obsMem + nematic_scratch_doubles() - 6, sizeof sums, hipMemcpyDeviceToHost,
                          c->stream));
    RP_HIP(hipStreamSynchronize(c->stream));
    // Q /= N; Q = 0.5 (3 Q - I)  (NematicOrder.cpp:18-20)
    const int idx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    double Q[9];
    for (int i = 0; i < 9; i++) {
        double v = sums[idx[i]] / (double)c->n;
        Q[i] = 0.5 * (3. * v - (i % 4 == 0 ? 1.0 : 0.0));
    }
    const int up[6] = {0, 1, 2, 4, 5, 8};
    for (int k = 0; k < 6; k++) out->q_tensor[k] = Q[up[k]];
    out->nematic_p2 = nematicP2(Q);
    return RP_OK;
}

int rp_download(rp_ctx *c, double box[9], double *pos, double *rot) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    RP_HIP(hipStreamSynchronize(c->stream));
    if (box) std::memcpy(box, c->box.dims, sizeof c->box.dims);
    size_t n = c->n;
    std::vector<int> tags(n);
    std::vector<double> col(n);
    RP_HIP(hipMemcpy(tags.data(), c->cur.p.tag, sizeof(int) * n, hipMemcpyDeviceToHost));
    auto down = [&](const double *src, double *dst, int stride, int off) -> int {
        RP_HIP(hipMemcpy(col.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) dst[stride * (size_t)tags[i] + off] = col[i];
        return RP_OK;
    };
    int r;
    if (pos && ((r = down(c->cur.p.x, pos, 3, 0)) || (r = down(c->cur.p.y, pos, 3, 1)) ||
                (r = down(c->cur.p.z, pos, 3, 2))))
        return r;
    if (rot)
        for (int k = 0; k < 9; k++)
            if ((r = down(c->cur.p.o[k], rot, 9, k))) return r;
    return RP_OK;
}

int rp_set_timing(rp_ctx *c, int enable) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    int r = timingHarvest(c);
    if (r) return r;
    if (enable < 0 || enable > RP_TIMING_SAMPLED) return fail(RP_ERR_PRECONDITION, "unknown timing mode");
    c->timing = enable;
    rp_timing keep = c->tstats;
    c->tstats = rp_timing{};
    c->tstats.lds_limit = keep.lds_limit;
    c->tstats.sweep_cap = keep.sweep_cap;
    c->tstats.sweep_waves = keep.sweep_waves;
    c->tstats.sweep_period = keep.sweep_period;
    return RP_OK;
}

int rp_autotune(rp_ctx *c, size_t cycles, rp_autotune_entry *out, size_t *nOut, size_t *chosen) {
    if (!c || !nOut) return fail(RP_ERR_PRECONDITION, "null argument");
    if (cycles == 0) return fail(RP_ERR_PRECONDITION, "cycles_per_shape must be positive");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    struct Shape {
        int waves, schedule;
    };
    std::vector<Shape> cands;
    const bool hard = !c->k.soft && !c->counting;
    if (hard && c->k.kind == RP_SHAPE_SPHEROCYLINDER && c->precision != RP_PREC_FP64) cands = {{8, 2}};
    else if (hard && c->k.kind != RP_SHAPE_POLYSPHERE) cands = {{8, 2}, {4, 2}, {2, 2}, {4, 1}, {2, 1}, {1, 0}};
    else if (hard) cands = {{16, 1}, {8, 1}, {4, 1}, {2, 1}, {1, 0}};
    else cands = {{c->sweepWaves, c->pipelined}};
    int r = timingHarvest(c);
    if (r) return r;
    // one start check, then the candidates' cycles back to back: the same cycles as rp_run_cycles over all of them
    r = validateStart(c);
    if (r) return r;
    const int keepTiming = c->timing;
    const rp_timing keepStats = c->tstats;
    double best = 0;
    size_t bestI = 0;
    for (size_t i = 0; i < cands.size(); i++) {
        c->sweepWaves = cands[i].waves;
        c->pipelined = cands[i].schedule;
        c->timing = 1;  // every launch timed
        c->tstats = rp_timing{};
        r = cycleLoop(c, cycles, false);
        if (r) {
            c->timing = keepTiming;
            c->tstats = keepStats;
            return r;
        }
        const double us = c->tstats.sweep_ms * 1e3 / (double)cycles;
        if (out && i < RP_AUTOTUNE_MAX) out[i] = rp_autotune_entry{cands[i].waves, cands[i].schedule, us};
        if (i == 0 || us < best) {
            best = us;
            bestI = i;
        }
    }
    c->sweepWaves = cands[bestI].waves;
    c->pipelined = cands[bestI].schedule;
    c->timing = keepTiming;
    c->tstats = keepStats;
    *nOut = cands.size();
    if (chosen) *chosen = bestI;
    return RP_OK;
}

int rp_get_timing(rp_ctx *c, rp_timing *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    int r = timingHarvest(c);
    if (r) return r;
    *out = c->tstats;
    return RP_OK;
}

void rp_destroy(rp_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->dPhases) {
        // wall_clock64 ticks at the device's wall-clock rate (kHz attribute)
        int khz = 100000;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        const double perUs = khz / 1000.0;
        unsigned long long ph[RP_PHASE_NUM] = {};
        if (hipMemcpy(ph, c->dPhases, sizeof ph, hipMemcpyDeviceToHost) == hipSuccess && ph[RP_PHASE_BLOCKS]) {
            double b = (double)ph[RP_PHASE_BLOCKS];
            std::fprintf(stderr,
                         "sweep phases per workgroup (us): table %.2f stage %.2f moves %.2f write %.2f; "
                         "%llu workgroups, longest %.1f us; wall clock %d kHz, %d workgroups per CU\n",
                         ph[RP_PHASE_TABLE] / b / perUs, ph[RP_PHASE_STAGE] / b / perUs,
                         ph[RP_PHASE_MOVES] / b / perUs, ph[RP_PHASE_WRITE] / b / perUs,
                         (unsigned long long)ph[RP_PHASE_BLOCKS], ph[RP_PHASE_LAST] / perUs, khz,
                         sweep_occupancy(c->shape.kind, c->tstats.sweep_cap ? (int)c->tstats.sweep_cap : 64,
                                         c->tstats.sweep_waves ? (int)c->tstats.sweep_waves : 4));
            if (ph[RP_PHASE_GROUPS]) {
                const double ng = (double)ph[RP_PHASE_GROUPS];
                std::fprintf(stderr,
                             "grouped sweep per group (us): loop top %.3f, planner busy %.3f, tester busy %.3f "
                             "(segment tests %.3f), barrier to end %.3f; %.2f groups per workgroup, %.1f %% replanned\n",
                             ph[RP_PHASE_TOP] / ng / perUs, ph[RP_PHASE_PLAN] / ng / perUs,
                             ph[RP_PHASE_SCAN] / ng / perUs, ph[RP_PHASE_PASSB] / ng / perUs,
                             ph[RP_PHASE_DECIDE] / ng / perUs, ng / b,
                             100.0 * ph[RP_PHASE_REPLANS] / ng);
            }
        }
        hipFree(c->dPhases);
    }
    for (auto e : c->evStart) hipEventDestroy(e);
    for (auto e : c->evStop) hipEventDestroy(e);
    if (c->cur.mem) hipFree(c->cur.mem);
    if (c->alt.mem) hipFree(c->alt.mem);
    if (c->last.mem) hipFree(c->last.mem);
    if (c->beadA.mem) hipFree(c->beadA.mem);
    if (c->beadB.mem) hipFree(c->beadB.mem);
    if (c->dBeads) hipFree(c->dBeads);
    if (c->beadCellStart) hipFree(c->beadCellStart);
    if (c->beadCellCount) hipFree(c->beadCellCount);
    if (c->cellStart) hipFree(c->cellStart);
    if (c->cellCount) hipFree(c->cellCount);
    if (c->propMem) hipFree(c->propMem);
    if (c->obsMem) hipFree(c->obsMem);
    if (c->cellEnergy) hipFree(c->cellEnergy);
    if (c->sumScratch) hipFree(c->sumScratch);
    if (c->dFixCounters) hipFree(c->dFixCounters);
    if (c->dSlots) hipFree(c->dSlots);
    if (c->scanTmp) hipFree(c->scanTmp);
    if (c->mlpTheta) hipFree(c->mlpTheta);
    for (auto *d : {&c->mlpSet, &c->mlpEval})
        if (d->x) hipFree(d->x);
    if (c->tmpSlot) hipFree(c->tmpSlot);
    if (c->tmpOrder) hipFree(c->tmpOrder);
    if (c->tmpPos) hipFree(c->tmpPos);
    if (c->dCounters) hipFree(c->dCounters);
    if (c->hCounters) hipHostFree(c->hCounters);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

namespace {

// Packing::renormalizeOrientations (Packing.cpp:1178-1209) through Simulation::fixRotationMatrices
// (Simulation.cpp:750-769): every orientation is replaced by fix_rotation(O) unless that creates an overlap; a shape
// without a hard part takes every fix.  The particles are visited cell by cell in the checkerboard colours (the
// reference visits them in index order; only a fix that itself creates an overlap can tell the orders apart).
int fixOrientations(rp_ctx *c, uint64_t &rejected) {
    rejected = 0;
    if (c->k.soft && !c->k.hardCore) {
        RP_HIP(launch_fix_all(c->stream, c->cur.p, (int)c->n));
        RP_HIP(hipStreamSynchronize(c->stream));
        return RP_OK;
    }
    int r = prepareCycleGrid(c);
    if (r) return r;
    r = foldSlots(c);  // the slots are about to count the renormalisation
    if (r) return r;
    if (!c->dFixCounters) RP_HIP(hipMalloc(&c->dFixCounters, sizeof(unsigned long long) * CNT_NUM));
    RP_HIP(launch_fill_u64(c->stream, c->dFixCounters, CNT_NUM, 0));
    const Grid &g = c->grid;
    SweepParams sp{};
    sp.g = g;
    sp.p = c->cur.p;
    sp.k = c->k;
    sp.k.soft = 0;  // a compound interaction's fix is judged by its hard part alone (Packing.cpp:1203-1206)
    for (int d = 0; d < 3; d++) sp.na[d] = std::max(1, g.n[d] / g.period);
    sp.moves = 0;
    sp.mult = 1;
    sp.fixMode = 1;
    sp.pipelined = 0;
    sp.counters = c->dFixCounters;
    r = ensureSlots(c, g.ncell);
    if (r) return r;
    sp.slots = c->dSlots;
    const double lambda = (double)c->n / (double)g.ncell;
    const int side = 2 * g.reach + 1;
    const double expect = side * side * side * lambda;
    const int waves = 4;
    int cap = (int)std::ceil(expect + 8 * std::sqrt(expect) + 64);
    cap = (cap + 63) / 64 * 64;
    const double perEntry =
        (double)(sweep_lds_bytes(c->k.kind, 1024, waves) - sweep_lds_bytes(c->k.kind, 0, waves)) / 1024.0;
    const int maxCap = (int)((double)(c->ldsLimit - sweep_lds_bytes(c->k.kind, 0, waves)) / perEntry) / 64 * 64;
    sp.cap = std::max(64, std::min(cap, maxCap));
    if (c->k.kind == RP_SHAPE_POLYSPHERE) sp.cap = 64;  // read from HBM
    if (g.allPairs) sp.cap = allPairsStaged(c);
    const int colours = g.period * g.period * g.period;
    for (int colour = 0; colour < colours; colour++) {
        const int par[3] = {colour % g.period, (colour / g.period) % g.period, colour / (g.period * g.period)};
        if ((g.n[0] == 1 && par[0]) || (g.n[1] == 1 && par[1]) || (g.n[2] == 1 && par[2])) continue;
        sp.colour = colour;
        RP_HIP(launch_sweep(c->stream, sp, waves));
    }
    RP_HIP(launch_fold_slots(c->stream, c->dSlots, g.ncell, c->dFixCounters));
    unsigned long long h[CNT_NUM];
    RP_HIP(hipMemcpyAsync(h, c->dFixCounters, sizeof h, hipMemcpyDeviceToHost, c->stream));
    RP_HIP(hipStreamSynchronize(c->stream));
    rejected = h[CNT_ATTEMPTED] - h[CNT_ACCEPTED];
    c->fixRejected += rejected;
    c->sortedForCycle = false;  // the next cycle draws its own grid origin
    return RP_OK;
}

}  // namespace

int rp_fix_orientations(rp_ctx *c, uint64_t *rejected) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    uint64_t rej = 0;
    int r = fixOrientations(c, rej);
    if (r) return r;
    if (rejected) *rejected = rej;
    return RP_OK;
}

int rp_set_overlap_counting(rp_ctx *c, int enable) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    if (!enable || c->k.soft) {
        c->counting = false;
        c->numOverlaps = 0;
        return RP_OK;
    }
    uint64_t full = 0;
    int r = countExact(c, 0, full);
    if (r == RP_ERR_UNSUPPORTED) {
        double zero[3] = {0, 0, 0};
        Grid g{};
        c->sortedForCycle = false;
        r = chooseSweepGrid(c, c->box, zero, g);
        if (r) return r;
        r = rebuild(c, g, 0, nullptr);
        if (r) return r;
        r = scanOverlaps(c, g, 0, full);
    }
    if (r) return r;
    RP_HIP(launch_fill_u64(c->stream, c->dCounters + CNT_OVERLAP_DELTA, 1, 0));
    c->counting = true;
    c->numOverlaps = (long long)full;
    return RP_OK;
}

int rp_cached_overlaps(rp_ctx *c, uint64_t *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!c->counting) return fail(RP_ERR_PRECONDITION, "overlap counting is off; no count is cached (Packing.cpp:1059)");
    *out = (uint64_t)c->numOverlaps;
    return RP_OK;
}

int rp_relax_overlaps(rp_ctx *c, size_t max_cycles, size_t *cycles_run, uint64_t *overlaps_left) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    int r = rp_set_overlap_counting(c, 1);
    if (r) return r;
    size_t done = 0;
    while (c->numOverlaps > 0 && done < max_cycles) {  // Simulation::relaxOverlaps (Simulation.cpp:249-256)
        r = rp_run_cycles(c, 1, 1, nullptr);
        if (r) return r;
        done++;
    }
    if (cycles_run) *cycles_run = done;
    if (overlaps_left) *overlaps_left = (uint64_t)c->numOverlaps;
    c->counting = false;
    return RP_OK;
}

int rp_set_orientation_fix_every(rp_ctx *c, uint64_t every) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    c->fixEvery = every;
    return RP_OK;
}

// Packing::tryScaling (Packing.cpp:172-208) on the device-resident state.  The state before the call is kept for
// rp_revert_scaling; the scaled state becomes current whatever the result, as in the reference.
int rp_try_scaling(rp_ctx *c, const double new_box[9], double *dE) {
    if (!c || !new_box || !dE) return fail(RP_ERR_PRECONDITION, "null argument");
    HostBox nb;
    nb.set(new_box);
    if (nb.volume() == 0) return fail(RP_ERR_PRECONDITION, "tryScaling: the new box has zero volume");
    if (!c->last.mem) {
        int r = allocParticles(c->last, c->n);
        if (r) return r;
    }
    static constexpr double INF = std::numeric_limits<double>::infinity();
    // energy of the current state, on a grid the current arrays are sorted by
    double eOld = 0;
    if (c->k.soft) {
        int r = prepareCycleGrid(c);
        if (r) return r;
        r = totalEnergy(c, c->grid, c->cur.p, eOld);
        if (r) return r;
    }
    // the scaled state, sorted by a checkerboard grid of the new box (no origin shift), into `alt`
    Grid g{};
    const double zero[3] = {0, 0, 0};
    int r = chooseSweepGrid(c, nb, zero, g);
    if (r) return r;
    RP_HIP(hipMemcpyAsync(c->last.mem, c->cur.mem, c->cur.bytes, hipMemcpyDeviceToDevice, c->stream));
    r = rebuild(c, g, 1, c->box.inv);
    if (r) return r;
    double result = 0;
    long long count = c->numOverlaps;
    // the hard part first (Packing.cpp:190-204); a compound interaction's energy only counts without an overlap
    const bool hard = !c->k.soft || c->k.hardCore;
    if (hard) {
        if (c->counting) {
            uint64_t after = 0;
            r = c->k.kind == RP_SHAPE_POLYSPHERE ? countBeads(c, c->alt.p, nb, 0, after) : scanOverlaps(c, g, 0, after);
            if (r) return r;
            count = (long long)after;
            if (count < c->numOverlaps) result = -INF;
            else if (count > c->numOverlaps) result = INF;
        } else if (c->k.kind == RP_SHAPE_POLYSPHERE) {  // not convex (PolysphereTraits.h:133): always counted
            uint64_t found = 0;
            r = countBeads(c, c->alt.p, nb, 1, found);
            if (r) return r;
            if (found) result = INF;
        } else if (!isBoxUpscaled(c->box, nb)) {  // spheres and spherocylinders are convex
            uint64_t found = 0;
            r = scanOverlaps(c, g, 1, found);
            if (r) return r;
            if (found) result = INF;
        }
    }
    if (c->k.soft && result == 0) {
        double eNew = 0;
        r = totalEnergy(c, g, c->alt.p, eNew);
        if (r) return r;
        result = eNew - eOld;
    }
    c->lastBox = c->box;
    c->lastNumOverlaps = c->numOverlaps;
    c->scalingSaved = true;
    c->box = nb;
    c->numOverlaps = count;
    swapBuffers(c);
    c->grid = g;
    c->sortedForCycle = true;
    *dE = result;
    return RP_OK;
}

// Packing::revertScaling (Packing.cpp:363-371): the state saved by the last rp_try_scaling becomes current.
int rp_revert_scaling(rp_ctx *c) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    if (!c->scalingSaved) return fail(RP_ERR_PRECONDITION, "revertScaling without a preceding tryScaling");
    RP_HIP(hipMemcpyAsync(c->cur.mem, c->last.mem, c->cur.bytes, hipMemcpyDeviceToDevice, c->stream));
    c->box = c->lastBox;
    c->numOverlaps = c->lastNumOverlaps;
    c->sortedForCycle = false;  // the cell arrays describe the scaled state's grid: the next cycle re-sorts
    RP_HIP(hipStreamSynchronize(c->stream));
    return RP_OK;
}

int rp_total_energy(rp_ctx *c, double *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    *out = 0;
    if (!c->k.soft) return RP_OK;  // Packing::getTotalEnergy returns 0 without a soft part (Packing.cpp:691-692)
    c->sortedForCycle = false;
    Grid g{};
    int disabled = 0;
    int r = makeExactGrid(c, c->box, g, disabled);
    if (r) return r;
    if (disabled) {
        // the reference would sum all pairs; every pair within range is also a half-shell pair of the
        // checkerboard grid, which exists whenever the box is at least two ranges high
        double zero[3] = {0, 0, 0};
        r = chooseSweepGrid(c, c->box, zero, g);
        if (r) return r;
    }
    r = rebuild(c, g, 0, nullptr);
    if (r) return r;
    return totalEnergy(c, g, c->alt.p, *out);
}

int rp_pair_energies(const rp_shape_desc *shape, size_t n, const double *records, double *energies) {
    if (!shape || (n && (!records || !energies))) return fail(RP_ERR_PRECONDITION, "null argument");
    if (softKind(*shape) == RP_SOFT_NONE)
        return fail(RP_ERR_UNSUPPORTED, "no soft interaction for this shape descriptor");
    if (int v = validateShape(*shape)) return v;
    if (n == 0) return RP_OK;
    PairConsts k = makeConsts(*shape);
    // Interaction::calculateEnergyBetween for one pair: the Lennard-Jones formula is not truncated
    if (softKind(*shape) == RP_SOFT_LJ) k.cut2 = std::numeric_limits<double>::infinity();
    double *dRec = nullptr, *dOut = nullptr;
    RP_HIP(hipMalloc(&dRec, sizeof(double) * 27 * n));
    if (hipMalloc(&dOut, sizeof(double) * n) != hipSuccess) {
        hipFree(dRec);
        return fail(RP_ERR_DEVICE, "hipMalloc failed");
    }
    int r = RP_OK;
    hipError_t e = hipMemcpy(dRec, records, sizeof(double) * 27 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_pair_energies(nullptr, k, n, dRec, dOut);
    if (e == hipSuccess) e = hipMemcpy(energies, dOut, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fail(RP_ERR_DEVICE, std::string("pair energies: ") + hipGetErrorString(e));
    hipFree(dRec);
    hipFree(dOut);
    return r;
}

namespace {

// Room for `need` pairs in d (one block: x 12, y 1, score 1 floats and meta 3 ints per pair), keeping the first
// `keep` pairs.
int ensureMlpData(rp_ctx *c, rp_ctx::MlpData &d, size_t need, size_t keep) {
    if (need <= d.cap) return RP_OK;
    const size_t cap = need + need / 4 + 1024;
    void *mem = nullptr;
    RP_HIP(hipMalloc(&mem, cap * (sizeof(float) * (MLP_FEATURES + 2) + sizeof(int) * 3)));
    rp_ctx::MlpData n;
    n.x = static_cast<float *>(mem);
    n.y = n.x + cap * MLP_FEATURES;
    n.score = n.y + cap;
    n.meta = reinterpret_cast<int *>(n.score + cap);
    n.cap = cap;
    if (d.x && keep) {
        RP_HIP(hipMemcpyAsync(n.x, d.x, sizeof(float) * keep * MLP_FEATURES, hipMemcpyDeviceToDevice, c->stream));
        RP_HIP(hipMemcpyAsync(n.y, d.y, sizeof(float) * keep, hipMemcpyDeviceToDevice, c->stream));
        RP_HIP(hipMemcpyAsync(n.meta, d.meta, sizeof(int) * 3 * keep, hipMemcpyDeviceToDevice, c->stream));
    }
    RP_HIP(hipStreamSynchronize(c->stream));
    if (d.x) hipFree(d.x);
    d = n;
    return RP_OK;
}

// Candidate pairs of M trial moves sampled on the current state (k_mlp_pairs) into d from index `at` on; P receives
// the number written.  Moves use the first move sampler's step sizes.
int mlpSample(rp_ctx *c, int M, rp_ctx::MlpData &d, size_t at, size_t &P) {
    int r = prepareCycleGrid(c);
    if (r) return r;
    SweepParams sp{};
    sp.g = c->grid;
    sp.p = c->cur.p;
    sp.k = c->k;
    sp.mlpInvRange = 1 / c->totalRange;
    const auto &sm = c->samplers.front();
    const double maxT = sm.maxTransImplicit ? minHeight(c->box) / 2 : sm.maxTransStep;
    const double ts = sm.kind == RP_MOVE_ROTATION ? 0.0 : std::min(maxT, sm.transStep);
    const double rs = sm.kind == RP_MOVE_TRANSLATION ? 0.0 : std::min(sm.rotStep, M_PI);
    void *tmp = nullptr;
    RP_HIP(hipMalloc(&tmp, (sizeof(int) + sizeof(long long)) * (size_t)M));
    int *dCounts = static_cast<int *>(tmp);
    long long *dOff = reinterpret_cast<long long *>(dCounts + M + (M & 1));
    std::vector<int> counts((size_t)M);
    std::vector<long long> off((size_t)M);
    const uint32_t stream = c->mlpStream++;
    hipError_t e = launch_mlp_pairs(c->stream, sp, M, c->seed, stream, ts, rs, 0, dCounts, nullptr, nullptr, nullptr,
                                    nullptr);
    if (e == hipSuccess) e = hipMemcpyAsync(counts.data(), dCounts, sizeof(int) * (size_t)M, hipMemcpyDeviceToHost,
                                            c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    size_t total = 0;
    for (int q = 0; q < M; q++) {
        off[(size_t)q] = (long long)(at + total);
        total += (size_t)counts[(size_t)q];
    }
    if (e == hipSuccess && ensureMlpData(c, d, at + total, at) != RP_OK) e = hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipMemcpy(dOff, off.data(), sizeof(long long) * (size_t)M, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_mlp_pairs(c->stream, sp, M, c->seed, stream, ts, rs, 1, nullptr, dOff, d.x, d.y, d.meta);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) return fail(RP_ERR_DEVICE, std::string("candidate pairs: ") + hipGetErrorString(e));
    P = total;
    return RP_OK;
}

// Mean loss of the network over the first n pairs of d, in batches of B (the gradient pass's loss sums).
int mlpDatasetLoss(rp_ctx *c, const rp_ctx::MlpData &d, size_t n, int B, MlpWork &w, float *grad, const int *dRows,
                   float *bx, float *by, double &loss) {
    double sum = 0;
    size_t done = 0;
    std::vector<float> parts((size_t)mlp_loss_blocks(B));
    std::vector<int> rows((size_t)B);
    for (size_t b0 = 0; b0 + (size_t)B <= n; b0 += (size_t)B) {
        for (int i = 0; i < B; i++) rows[(size_t)i] = (int)(b0 + (size_t)i);
        RP_HIP(hipMemcpyAsync(const_cast<int *>(dRows), rows.data(), sizeof(int) * (size_t)B, hipMemcpyHostToDevice,
                              c->stream));
        RP_HIP(mlp_gather(c->stream, d.x, d.y, dRows, B, bx, by));
        RP_HIP(mlp_forward_backward(c->stream, c->mlpTheta, B, bx, by, w, grad));
        RP_HIP(hipMemcpyAsync(parts.data(), w.lossPart, sizeof(float) * parts.size(), hipMemcpyDeviceToHost,
                              c->stream));
        RP_HIP(hipStreamSynchronize(c->stream));
        for (float v : parts) sum += v;
        done += (size_t)B;
    }
    loss = done ? sum / (double)done : 0;
    return RP_OK;
}

}  // namespace

int rp_mlp_train(rp_ctx *c, size_t snapshots, size_t cyclesBetween, size_t movesPerSnapshot, size_t epochs,
                 size_t batch, double learningRate, rp_mlp_stats *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    if (c->k.kind != RP_SHAPE_POLYSPHERE)
        return fail(RP_ERR_UNSUPPORTED, "the learned ordering is built for polysphere sweeps");
    if (snapshots == 0 || movesPerSnapshot == 0 || movesPerSnapshot > (1u << 24) || batch < 16 ||
        batch > (1u << 16) || !(learningRate > 0))
        return fail(RP_ERR_PRECONDITION, "bad training arguments");
    *out = rp_mlp_stats{};
    // pairs from the run's own trajectory: cyclesBetween cycles, then movesPerSnapshot sampled trial moves
    size_t n = 0;
    for (size_t s = 0; s < snapshots; s++) {
        for (size_t i = 0; i < cyclesBetween; i++) {
            int r = runCycle(c);
            if (r) return r;
        }
        size_t P = 0;
        int r = mlpSample(c, (int)movesPerSnapshot, c->mlpSet, n, P);
        if (r) return r;
        n += P;
    }
    const int B = (int)batch;
    if (n < (size_t)B) return fail(RP_ERR_PRECONDITION, "fewer training pairs than one batch");
    // parameters, Adam moments, gradient
    if (!c->mlpTheta) RP_HIP(hipMalloc(&c->mlpTheta, sizeof(float) * 4 * MLP_PARAMS));
    float *m = c->mlpTheta + MLP_PARAMS, *v = m + MLP_PARAMS, *grad = v + MLP_PARAMS;
    std::vector<float> y(n);
    RP_HIP(hipMemcpy(y.data(), c->mlpSet.y, sizeof(float) * n, hipMemcpyDeviceToHost));
    double pos = 0;
    for (float t : y) pos += t;
    const double frac = pos / (double)n;
    // Glorot-uniform weights, zero biases, the output bias at the log-odds of the overlap fraction
    {
        std::vector<float> th(MLP_PARAMS, 0.f);
        Philox rng = philox_init(c->seed ^ 0x6d6c70696e6974ull, 0, 0, 0);
        auto fill = [&](int offset, int fanOut, int fanIn) {
            const double a = std::sqrt(6.0 / (fanIn + fanOut));
            for (int i = 0; i < fanOut * fanIn; i++) th[(size_t)(offset + i)] = (float)philox_uniform(rng, -a, a);
        };
        fill(MLP_OFF_W1, MLP_H, MLP_IN);
        fill(MLP_OFF_W2, MLP_H, MLP_H);
        fill(MLP_OFF_W3, MLP_H, MLP_H);
        fill(MLP_OFF_W4, 1, MLP_H);
        const double f = std::min(std::max(frac, 1e-6), 1 - 1e-6);
        th[MLP_OFF_B4] = (float)std::log(f / (1 - f));
        RP_HIP(hipMemcpy(c->mlpTheta, th.data(), sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice));
        RP_HIP(hipMemset(m, 0, sizeof(float) * 2 * MLP_PARAMS));
    }
    // batch work arrays: rows, gathered batch, activations
    const size_t lossBlocks = (size_t)mlp_loss_blocks(B);
    void *work = nullptr;
    RP_HIP(hipMalloc(&work, sizeof(int) * (size_t)B + sizeof(float) * ((size_t)B * (MLP_IN + 1) +
                                                                       9 * (size_t)B * MLP_H + B + lossBlocks)));
    int *dRows = static_cast<int *>(work);
    float *bx = reinterpret_cast<float *>(dRows + B), *by = bx + (size_t)B * MLP_IN;
    MlpWork w;
    float *q = by + B;
    float **arr[9] = {&w.pre1, &w.a1, &w.pre2, &w.a2, &w.pre3, &w.a3, &w.d1, &w.d2, &w.d3};
    for (auto p : arr) {
        *p = q;
        q += (size_t)B * MLP_H;
    }
    w.dz = q;
    w.lossPart = q + B;
    struct FreeWork {
        void *p;
        ~FreeWork() { hipFree(p); }
    } freeWork{work};
    double loss0 = 0, loss1 = 0;
    int r = mlpDatasetLoss(c, c->mlpSet, n, B, w, grad, dRows, bx, by, loss0);
    if (r) return r;
    // epochs over a fresh permutation each, full batches only
    std::vector<int> perm(n);
    Philox prng = philox_init(c->seed ^ 0x6d6c7065726dull, 1, 0, 0);
    int step = 0;
    for (size_t ep = 0; ep < epochs; ep++) {
        for (size_t i = 0; i < n; i++) perm[i] = (int)i;
        for (size_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[philox_below(prng, (uint32_t)(i + 1))]);
        for (size_t b0 = 0; b0 + (size_t)B <= n; b0 += (size_t)B) {
            RP_HIP(hipMemcpyAsync(dRows, perm.data() + b0, sizeof(int) * (size_t)B, hipMemcpyHostToDevice,
                                  c->stream));
            RP_HIP(mlp_gather(c->stream, c->mlpSet.x, c->mlpSet.y, dRows, B, bx, by));
            RP_HIP(mlp_forward_backward(c->stream, c->mlpTheta, B, bx, by, w, grad));
            RP_HIP(mlp_adam(c->stream, c->mlpTheta, grad, m, v, (float)learningRate, ++step));
            // the host buffer of the next batch is rewritten only after this copy has run
            RP_HIP(hipStreamSynchronize(c->stream));
        }
    }
    r = mlpDatasetLoss(c, c->mlpSet, n, B, w, grad, dRows, bx, by, loss1);
    if (r) return r;
    c->mlpTrained = true;
    out->pairs = n;
    out->overlap_fraction = frac;
    out->loss_initial = loss0;
    out->loss_final = loss1;
    out->steps = (uint64_t)step;
    return RP_OK;
}

int rp_mlp_order_stats(rp_ctx *c, size_t moves, rp_mlp_stats *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    if (c->k.kind != RP_SHAPE_POLYSPHERE)
        return fail(RP_ERR_UNSUPPORTED, "the learned ordering is built for polysphere sweeps");
    if (!c->mlpTrained) return fail(RP_ERR_PRECONDITION, "the ordering network is not trained (rp_mlp_train)");
    if (moves == 0 || moves > (1u << 24)) return fail(RP_ERR_PRECONDITION, "moves must be 1 .. 2^24");
    *out = rp_mlp_stats{};
    size_t P = 0;
    int r = mlpSample(c, (int)moves, c->mlpEval, 0, P);
    if (r) return r;
    RP_HIP(mlp_infer(c->stream, c->mlpTheta, (int)P, c->mlpEval.x, c->mlpEval.score));
    std::vector<float> y(P), score(P);
    std::vector<int> meta(3 * P);
    RP_HIP(hipMemcpyAsync(y.data(), c->mlpEval.y, sizeof(float) * P, hipMemcpyDeviceToHost, c->stream));
    RP_HIP(hipMemcpyAsync(score.data(), c->mlpEval.score, sizeof(float) * P, hipMemcpyDeviceToHost, c->stream));
    RP_HIP(hipMemcpyAsync(meta.data(), c->mlpEval.meta, sizeof(int) * 3 * P, hipMemcpyDeviceToHost, c->stream));
    RP_HIP(hipStreamSynchronize(c->stream));
    // per move: the pairs are contiguous and in natural order; exact tests with early exit = 1 + the position of the
    // first overlapping candidate in the order, or every candidate when none overlaps
    double natural = 0, shell = 0, learned = 0, cand = 0, pos = 0;
    uint64_t rejected = 0;
    std::vector<size_t> idx;
    for (size_t a = 0; a < P;) {
        size_t b = a;
        while (b < P && meta[3 * b] == meta[3 * a]) b++;
        const size_t cnt = b - a;
        idx.resize(cnt);
        auto firstHit = [&](void) {
            for (size_t i = 0; i < cnt; i++)
                if (y[idx[i]] > 0.5f) return (double)(i + 1);
            return (double)cnt;
        };
        for (size_t i = 0; i < cnt; i++) idx[i] = a + i;
        const double nat = firstHit();
        std::stable_sort(idx.begin(), idx.end(), [&](size_t u, size_t v) { return meta[3 * u + 2] < meta[3 * v + 2]; });
        const double sh = firstHit();
        for (size_t i = 0; i < cnt; i++) idx[i] = a + i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t u, size_t v) { return score[u] > score[v]; });
        const double le = firstHit();
        bool any = false;
        for (size_t i = a; i < b; i++) {
            any = any || y[i] > 0.5f;
            pos += y[i];
        }
        rejected += any ? 1 : 0;
        natural += nat;
        shell += sh;
        learned += le;
        cand += (double)cnt;
        a = b;
    }
    // moves with no close candidate (or leaving their cell) need no exact test in any order
    out->moves = moves;
    out->moves_rejected = rejected;
    out->pairs = P;
    out->overlap_fraction = P ? pos / (double)P : 0;
    out->candidates_per_move = cand / (double)moves;
    out->tests_natural = natural / (double)moves;
    out->tests_nearest_first = shell / (double)moves;
    out->tests_learned = learned / (double)moves;
    return RP_OK;
}

int rp_mlp_get_weights(rp_ctx *c, float *theta) {
    if (!c || !theta) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!c->mlpTrained) return fail(RP_ERR_PRECONDITION, "the ordering network is not trained (rp_mlp_train)");
    RP_HIP(hipMemcpy(theta, c->mlpTheta, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost));
    return RP_OK;
}

int rp_mlp_set_weights(rp_ctx *c, const float *theta) {
    if (!c || !theta) return fail(RP_ERR_PRECONDITION, "null argument");
    if (!c->mlpTheta) RP_HIP(hipMalloc(&c->mlpTheta, sizeof(float) * 4 * MLP_PARAMS));
    RP_HIP(hipMemcpy(c->mlpTheta, theta, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice));
    c->mlpTrained = true;
    return RP_OK;
}

int rp_mlp_pairs(rp_ctx *c, size_t moves, size_t cap, float *x, float *y, int *meta, size_t *n) {
    if (!c || !n) return fail(RP_ERR_PRECONDITION, "null argument");
    if (c->k.kind != RP_SHAPE_POLYSPHERE)
        return fail(RP_ERR_UNSUPPORTED, "the learned ordering is built for polysphere sweeps");
    if (!c->envSet) return fail(RP_ERR_PRECONDITION, "environment not set");
    if (moves == 0 || moves > (1u << 24)) return fail(RP_ERR_PRECONDITION, "moves must be 1 .. 2^24");
    size_t P = 0;
    int r = mlpSample(c, (int)moves, c->mlpEval, 0, P);
    if (r) return r;
    *n = P;
    if (P > cap) return fail(RP_ERR_PRECONDITION, "capacity below the number of pairs");
    if (P == 0) return RP_OK;
    if (!x || !y || !meta) return fail(RP_ERR_PRECONDITION, "null pair array");
    RP_HIP(hipMemcpy(x, c->mlpEval.x, sizeof(float) * MLP_FEATURES * P, hipMemcpyDeviceToHost));
    RP_HIP(hipMemcpy(y, c->mlpEval.y, sizeof(float) * P, hipMemcpyDeviceToHost));
    RP_HIP(hipMemcpy(meta, c->mlpEval.meta, sizeof(int) * 3 * P, hipMemcpyDeviceToHost));
    return RP_OK;
}

int rp_mlp_gradient(const float *theta, size_t n, const float *x, const float *y, double *loss, float *grad) {
    if (!theta || !loss || !grad || (n && (!x || !y))) return fail(RP_ERR_PRECONDITION, "null argument");
    if (n == 0 || n > (1u << 20)) return fail(RP_ERR_PRECONDITION, "batch size must be 1 .. 2^20");
    const int B = (int)n;
    const size_t lossBlocks = (size_t)mlp_loss_blocks(B);
    // theta, grad, x, y, 9 activation arrays of B x 64, dz, loss partials
    const size_t floats = 2 * (size_t)MLP_PARAMS + (size_t)B * (MLP_IN + 1) + 9 * (size_t)B * MLP_H + B + lossBlocks;
    float *d = nullptr;
    RP_HIP(hipMalloc(&d, sizeof(float) * floats));
    float *dTheta = d, *dGrad = dTheta + MLP_PARAMS, *dX = dGrad + MLP_PARAMS, *dY = dX + (size_t)B * MLP_IN;
    MlpWork w;
    float *q = dY + B;
    float **arr[9] = {&w.pre1, &w.a1, &w.pre2, &w.a2, &w.pre3, &w.a3, &w.d1, &w.d2, &w.d3};
    for (auto p : arr) {
        *p = q;
        q += (size_t)B * MLP_H;
    }
    w.dz = q;
    w.lossPart = q + B;
    std::vector<float> parts(lossBlocks);
    hipError_t e = hipMemcpy(dTheta, theta, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dX, x, sizeof(float) * (size_t)B * MLP_IN, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dY, y, sizeof(float) * (size_t)B, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mlp_forward_backward(nullptr, dTheta, B, dX, dY, w, dGrad);
    if (e == hipSuccess) e = hipMemcpy(grad, dGrad, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(parts.data(), w.lossPart, sizeof(float) * lossBlocks, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(RP_ERR_DEVICE, std::string("mlp gradient: ") + hipGetErrorString(e));
    double sum = 0;
    for (float v : parts) sum += v;
    *loss = sum / (double)B;
    return RP_OK;
}

int rp_mlp_infer(const float *theta, size_t n, const float *x, float *score) {
    if (!theta || (n && (!x || !score))) return fail(RP_ERR_PRECONDITION, "null argument");
    if (n == 0) return RP_OK;
    if (n > (size_t)0x7fffffff / MLP_IN) return fail(RP_ERR_PRECONDITION, "too many pairs");
    float *d = nullptr;
    RP_HIP(hipMalloc(&d, sizeof(float) * (MLP_PARAMS + n * (MLP_IN + 1))));
    float *dTheta = d, *dX = d + MLP_PARAMS, *dS = dX + n * MLP_IN;
    hipError_t e = hipMemcpy(dTheta, theta, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dX, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mlp_infer(nullptr, dTheta, (int)n, dX, dS);
    if (e == hipSuccess) e = hipMemcpy(score, dS, sizeof(float) * n, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(RP_ERR_DEVICE, std::string("mlp inference: ") + hipGetErrorString(e));
    return RP_OK;
}

int rp_set_precision(rp_ctx *c, int precision) {
    if (!c) return fail(RP_ERR_PRECONDITION, "null context");
    if (precision != RP_PREC_FP64 && precision != RP_PREC_GUARDED_FP32 && precision != RP_PREC_GUARDED_FP32_VERIFY)
        return fail(RP_ERR_PRECONDITION, "unknown precision");
    c->precision = precision;
    return RP_OK;
}

int rp_get_guard_stats(rp_ctx *c, rp_guard_stats *out) {
    if (!c || !out) return fail(RP_ERR_PRECONDITION, "null argument");
    int r = readCounters(c);
    if (r) return r;
    *out = rp_guard_stats{};
    out->pairs = c->hCounters[CNT_G32_PAIRS];
    out->fp64_fallback = c->hCounters[CNT_G32_GUARD];
    out->mismatches = c->hCounters[CNT_G32_MISMATCH];
    const uint32_t bits = (uint32_t)c->hCounters[CNT_G32_MAXERR];
    float err;
    std::memcpy(&err, &bits, sizeof err);
    out->max_err_centre = err;
    return RP_OK;
}

int rp_box_info(const double box[9], double inv[9], double heights[3], double *volume) {
    if (!box || !inv || !heights || !volume) return fail(RP_ERR_PRECONDITION, "null argument");
    HostBox b;
    b.set(box);
    std::memcpy(inv, b.inv, sizeof b.inv);
    b.heights(heights);
    *volume = b.volume();
    return RP_OK;
}

int rp_apply_moves(const double box[9], size_t n, const double *pos, const double *rot, const double *t,
                   const double *axis, const double *angle, double *outPos, double *outR, double *outRot,
                   double *outSinCos) {
    if (!box || (n && (!pos || !rot || !t || !axis || !angle || !outPos || !outR || !outRot || !outSinCos)))
        return fail(RP_ERR_PRECONDITION, "null argument");
    HostBox b;
    b.set(box);
    if (b.volume() == 0) return fail(RP_ERR_PRECONDITION, "zero-volume box");
    if (n == 0) return RP_OK;
    // one block: inputs pos 3, rot 9, t 3, axis 3, angle 1; outputs pos 3, R 9, rot 9, sin/cos 2 (42 doubles a move)
    double *d = nullptr;
    RP_HIP(hipMalloc(&d, sizeof(double) * 42 * n));
    double *dPos = d, *dRot = dPos + 3 * n, *dT = dRot + 9 * n, *dAxis = dT + 3 * n, *dAngle = dAxis + 3 * n;
    double *oPos = dAngle + n, *oR = oPos + 3 * n, *oRot = oR + 9 * n, *oSc = oRot + 9 * n;
    hipError_t e = hipSuccess;
    const struct { double *dst; const double *src; size_t len; } in[] = {
        {dPos, pos, 3 * n}, {dRot, rot, 9 * n}, {dT, t, 3 * n}, {dAxis, axis, 3 * n}, {dAngle, angle, n}};
    for (const auto &c : in)
        if (e == hipSuccess) e = hipMemcpy(c.dst, c.src, sizeof(double) * c.len, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_apply_moves(nullptr, b.dims, b.inv, n, dPos, dRot, dT, dAxis, dAngle, oPos, oR, oRot, oSc);
    const struct { double *dst; const double *src; size_t len; } out[] = {
        {outPos, oPos, 3 * n}, {outR, oR, 9 * n}, {outRot, oRot, 9 * n}, {outSinCos, oSc, 2 * n}};
    for (const auto &c : out)
        if (e == hipSuccess) e = hipMemcpy(c.dst, c.src, sizeof(double) * c.len, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(RP_ERR_DEVICE, std::string("apply moves: ") + hipGetErrorString(e));
    return RP_OK;
}

int rp_pair_decisions(const rp_shape_desc *shape, size_t n, const double *records, uint8_t *decisions) {
    if (!shape || (n && (!records || !decisions))) return fail(RP_ERR_PRECONDITION, "null argument");
    if (int v = validateShape(*shape)) return v;
    if (!hasHardPart(*shape)) return fail(RP_ERR_UNSUPPORTED, "a soft interaction has no hard part");
    if (n == 0) return RP_OK;
    PairConsts k = makeConsts(*shape);
    double *dBeads = nullptr;
    if (int r = uploadBeads(*shape, &dBeads)) return r;
    k.beads = dBeads;
    k.nc = shape->kind == RP_SHAPE_POLYSPHERE ? shape->n_centres : 0;
    struct FreeBeads {
        double *p;
        ~FreeBeads() {
            if (p) hipFree(p);
        }
    } freeBeads{dBeads};
    double *dRec = nullptr;
    uint8_t *dOut = nullptr;
    RP_HIP(hipMalloc(&dRec, sizeof(double) * 27 * n));
    if (hipMalloc(&dOut, n) != hipSuccess) {
        hipFree(dRec);
        return fail(RP_ERR_DEVICE, "hipMalloc failed");
    }
    int r = RP_OK;
    hipError_t e = hipMemcpy(dRec, records, sizeof(double) * 27 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_pair_decisions(nullptr, k, n, dRec, dOut);
    if (e == hipSuccess) e = hipMemcpy(decisions, dOut, n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fail(RP_ERR_DEVICE, std::string("pair decisions: ") + hipGetErrorString(e));
    hipFree(dRec);
    hipFree(dOut);
    return r;
}

int rp_pair_decisions_guarded(const rp_shape_desc *shape, size_t n, const double *records, uint8_t *decisions,
                              uint8_t *codes, rp_guard_stats *stats) {
    if (!shape || (n && (!records || !decisions))) return fail(RP_ERR_PRECONDITION, "null argument");
    if (shape->kind != RP_SHAPE_SPHERE && shape->kind != RP_SHAPE_SPHEROCYLINDER)
        return fail(RP_ERR_UNSUPPORTED, "unsupported shape kind");
    if (!hasHardPart(*shape)) return fail(RP_ERR_UNSUPPORTED, "a soft interaction has no hard part");
    if (stats) *stats = rp_guard_stats{};
    if (n == 0) return RP_OK;
    PairConsts k = makeConsts(*shape);
    double *dRec = nullptr;
    uint8_t *dOut = nullptr;
    float *dErr = nullptr;
    const size_t outBytes = 2 * n, errBytes = sizeof(float) * 2 * n;
    RP_HIP(hipMalloc(&dRec, sizeof(double) * 27 * n));
    if (hipMalloc(&dOut, outBytes) != hipSuccess || hipMalloc(&dErr, errBytes) != hipSuccess) {
        hipFree(dRec);
        if (dOut) hipFree(dOut);
        return fail(RP_ERR_DEVICE, "hipMalloc failed");
    }
    std::vector<uint8_t> hOut(outBytes);
    std::vector<float> hErr(2 * n);
    int r = RP_OK;
    hipError_t e = hipMemcpy(dRec, records, sizeof(double) * 27 * n, hipMemcpyHostToDevice);
    // without codes or statistics the kernel runs FP64 only on the undecided pairs
    const bool verify = codes || stats;
    if (e == hipSuccess)
        e = launch_pair_decisions_guarded(nullptr, k, n, dRec, dOut, verify ? dOut + n : nullptr, dErr, dErr + n);
    if (e == hipSuccess) e = hipMemcpy(hOut.data(), dOut, verify ? outBytes : n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && verify) e = hipMemcpy(hErr.data(), dErr, errBytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fail(RP_ERR_DEVICE, std::string("guarded pair decisions: ") + hipGetErrorString(e));
    hipFree(dRec);
    hipFree(dOut);
    hipFree(dErr);
    if (r) return r;
    std::memcpy(decisions, hOut.data(), n);
    if (codes) std::memcpy(codes, hOut.data() + n, n);
    if (stats) {
        stats->pairs = n;
        for (size_t i = 0; i < n; i++) {
            const uint8_t c = hOut[n + i];
            if ((c & 3) == 2) stats->fp64_fallback++;
            if (c & 4) stats->segment_stage++;
            if ((c & 4) && (c & 3) == 2) stats->segment_fallback++;
            if (c & 8) stats->mismatches++;
            stats->max_err_centre = std::max(stats->max_err_centre, (double)hErr[i]);
            if (hErr[n + i] >= 0) stats->max_err_segment = std::max(stats->max_err_segment, (double)hErr[n + i]);
        }
    }
    return RP_OK;
}

int rp_pair_decisions_guarded_noalloc(const rp_shape_desc *shape, size_t n, const double *d_records,
                                      uint8_t *d_decisions, uint8_t *d_codes, rp_guard_stats *stats) {
    if (!shape || (n && (!d_records || !d_decisions))) return fail(RP_ERR_PRECONDITION, "null argument");
    if (shape->kind != RP_SHAPE_SPHERE && shape->kind != RP_SHAPE_SPHEROCYLINDER)
        return fail(RP_ERR_UNSUPPORTED, "unsupported shape kind");
    if (!hasHardPart(*shape)) return fail(RP_ERR_UNSUPPORTED, "a soft interaction has no hard part");
    if (n > RP_GUARD_NOALLOC_MAX_PAIRS)
        return fail(RP_ERR_PRECONDITION, std::to_string(n) + " pairs exceed the fixed scratch area (" +
                                             std::to_string((size_t)RP_GUARD_NOALLOC_MAX_PAIRS) + " pairs)");
    if (stats) *stats = rp_guard_stats{};
    if (n == 0) return RP_OK;
    GuardScratch sc{};
    RP_HIP(guard_scratch(sc));
    const PairConsts k = makeConsts(*shape);
    const bool verify = d_codes || stats;
    uint8_t *codes = d_codes ? d_codes : sc.code;
    RP_HIP(launch_pair_decisions_guarded(nullptr, k, n, d_records, d_decisions, verify ? codes : nullptr, sc.errC,
                                         sc.errS));
    if (stats) {
        // the area keeps the last call's contents: the counters are cleared (a fill kernel) before every reduction
        RP_HIP(launch_fill_u64(nullptr, sc.counters, 8, 0));
        RP_HIP(launch_guard_stats(nullptr, n, codes, sc.errC, sc.errS, sc.counters));
        unsigned long long h[8] = {};
        RP_HIP(hipMemcpy(h, sc.counters, sizeof h, hipMemcpyDeviceToHost));
        float ec = 0, es = 0;
        const uint32_t bc = (uint32_t)h[4], bs = (uint32_t)h[5];
        std::memcpy(&ec, &bc, sizeof ec);
        std::memcpy(&es, &bs, sizeof es);
        stats->pairs = n;
        stats->fp64_fallback = h[0];
        stats->segment_stage = h[1];
        stats->segment_fallback = h[2];
        stats->mismatches = h[3];
        stats->max_err_centre = ec;
        stats->max_err_segment = es;
    } else {
        RP_HIP(hipStreamSynchronize(nullptr));
    }
    return RP_OK;
}

}  // extern "C"

// Error reporting for the library's other host sources (C++ linkage).
namespace rp {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace rp
