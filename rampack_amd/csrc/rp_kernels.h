// rp_kernels.h -- device data layout and kernel launch wrappers (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "rampack_amd.h"  // RP_MAX_CENTRES
#include "rp_math.h"

namespace rp {

// Particle state in HBM, structure of arrays, sorted by cell (one sort per rebuild).  x,y,z are the FP64
// accumulated positions, o[0..8] the row-major FP64 orientation matrix, tag the caller's particle index,
// wbits the three wrap bits of the position relative to the current grid origin, cell the cell index.
struct Particles {
    double *x, *y, *z;
    double *o[9];
    int *tag;
    uint8_t *wbits;
    int *cell;
};

// One cell grid over the (triclinic) box in relative coordinates.  exact = 1 reproduces the reference's
// NeighbourGrid binning (NeighbourGrid.cpp:13-34, no origin shift, reach 1); exact = 0 is the checkerboard grid,
// with an origin shift in [0, 1/n) per axis, cells at least range / reach high and a multiple of period = reach + 1
// cells per axis, so that same-colour cells are separated by reach cells.
struct Grid {
    double dims[9], inv[9];
    double T[81];           // translations dims * (i-1, j-1, k-1), index 9i + 3j + k (NeighbourGrid.cpp:203-217)
    double relCell[3];      // 1 / n
    double shift[3];
    int n[3];
    int ncell;
    int exact;
    int reach;              // stencil half-width in cells: cells are at least range / reach high
    int period;             // checkerboard colour period per axis (reach + 1); colours = period^3
    int allPairs;           // 1: no grid (the reference's box too small for its neighbour grid, Packing.cpp:862-873):
                            // one cell holds every particle and pairs meet through the nearest image only
    int *cellStart;         // ncell + 1
    int *cellCount;         // ncell
};

// Trial-move proposals of one cycle, one per (cell, move), structure of arrays with `stride` entries per field.
struct Proposals {
    int *k;         // particle index within the cell
    double *t;      // translation, 3 x stride
    double *R;      // rotation matrix, row-major, 9 x stride
    double *u;      // Metropolis uniform in [0, 1) for soft interactions; nullptr for hard ones
    uint8_t *type;  // with several move samplers: the sampler of each proposal (255: no move); else nullptr
    uint8_t *acc;   // with several move samplers: 1 once the sweep accepts the proposal; else nullptr
    int stride;
};

// The move samplers of a cycle (RototranslationSampler, TranslationSampler, RotationSampler): each proposal picks
// one uniformly (every sampler requests N moves, Simulation.cpp:456-465, 683-691) and uses its step sizes.
struct MoveSet {
    int n;
    int kind[RP_MAX_MOVE_SAMPLERS];   // RP_MOVE_*
    double ts[RP_MAX_MOVE_SAMPLERS];  // translation half-width in use (already capped by the maximum step)
    double rs[RP_MAX_MOVE_SAMPLERS];  // rotation half-angle in use
};

struct SweepParams {
    Grid g;
    Particles p;
    PairConsts k;
    Proposals pr;
    int moves;              // trial moves per active cell and colour pass; 0: mult per particle the cell holds
    int mult;               // moves per particle in that case: the number of move samplers
    int colour;
    int na[3];              // active cells per axis (n / period)
    int cap;                // stencil entries that fit in LDS; without a grid: the particles the all-pairs sweep
                            // stages in LDS (0: it reads them in place)
    int pipelined;          // 1: wavefront 0 plans the next move while the others test; 2: grouped moves
                            // (both need >= 2 wavefronts)
    double temperature;     // soft interactions: the move is accepted if u <= exp(-dE / temperature)
    int counting;           // overlap-counting mode (Packing::toggleOverlapCounting): a move is accepted if it does not
                            // raise the moved particle's overlap count; runs in the soft sweep with 0/1 pair values
    int fixMode;            // orientation renormalisation pass: one trial per own particle in index order, no
                            // translation, orientation fix_rotation(O) instead of R O (Packing::renormalizeOrientations)
    unsigned long long *counters;
    unsigned long long *slots;   // per grid cell: its workgroup's move counts (SLOT_WORDS words), see launch_fold_slots
    unsigned long long *phases;  // optional per-phase wall-clock sums (RP_PHASE_*), nullptr when not profiling
    int tuning;                  // grouped sweep wave priorities: bit 0 raises the planner, bit 1 the testers
                                 // that scan one slot more than the others
    unsigned long long *trace;   // optional per-workgroup record of one launch: start and end wall clock,
                                 // own | stencil << 16 | cell << 32, compute-unit id; nullptr when not tracing
    // Guarded FP32 stencil scan of the grouped sweep (k_sweep<..., G32 = 1 or 2>): candidate positions relative to the
    // active cell's centre in FP32, squared centre distances in FP32, and the FP64 value only for pairs the FP32 value
    // cannot place on one side of a threshold.  g32Lim[s] = {contact2 lower, contact2 upper, range2 lower,
    // range2 upper} bounds of the FP32 guard bands, s = 0 for cells whose stencil was staged translated (no wrapped
    // particle), s = 1 otherwise (positions and image translations up to a box length from the centre).
    float g32Lim[2][4];
    double mlpInvRange;  // candidate pairs of the learned ordering: 1 / total interaction range (feature scale)
};

constexpr int MLP_FEATURES = 12;

// Phases of a sweep workgroup timed with the device wall clock when SweepParams::phases is set.
enum SweepPhase : int {
    RP_PHASE_TABLE = 0,    // stencil cell table and prefix sums
    RP_PHASE_STAGE = 1,    // stencil and own orientations into LDS
    RP_PHASE_MOVES = 2,    // the trial moves
    RP_PHASE_WRITE = 3,    // write-back and counters
    RP_PHASE_BLOCKS = 4,   // workgroups timed
    RP_PHASE_LAST = 6,     // longest workgroup (max)
    RP_PHASE_PLAN = 8,     // grouped sweep: planner wavefront busy (planning the next group)
    RP_PHASE_SCAN = 9,     // grouped sweep: first tester wavefront busy (stencil scan and segment tests)
    RP_PHASE_DECIDE = 10,  // grouped sweep: group barrier to the end of the group (decisions, write-back, replans)
    RP_PHASE_GROUPS = 11,  // grouped sweep: groups
    RP_PHASE_REPLANS = 12, // grouped sweep: groups planned twice
    RP_PHASE_TOP = 13,     // grouped sweep: start of a group to the start of planning / scanning (plan words)
    RP_PHASE_PASSB = 14,   // grouped sweep: first tester wavefront in the segment tests of queued pairs
    RP_PHASE_NUM = 16
};

enum Counter : int {
    CNT_ACCEPTED = 0,
    CNT_ATTEMPTED = 1,
    CNT_LEFT_CELL = 2,
    CNT_UNSTAGED_BLOCKS = 3,
    CNT_OVERLAPS = 4,
    CNT_MAX_STENCIL = 5,
    CNT_TESTS = 6,
    CNT_OVERLAP_DELTA = 7,  // overlap-counting mode: sum of the accepted moves' overlap changes (two's complement)
    CNT_TYPE_ATT = 8,       // per move sampler (RP_MAX_MOVE_SAMPLERS entries): attempted moves
    CNT_TYPE_ACC = 12,      // per move sampler: accepted moves
    CNT_STAGE_OVERFLOW = 16,  // cells of an overlap count or energy sum too full to stage (cleared before each launch)
    CNT_G32_PAIRS = 17,       // guarded FP32 sweep: (candidate, move) pairs classified from FP32 distances
    CNT_G32_GUARD = 18,       // of those, pairs inside a guard band, decided by the FP64 distance
    CNT_G32_MAXERR = 19,      // verify mode: largest |FP32 - FP64| squared centre distance (float bits)
    CNT_G32_MISMATCH = 20,    // verify mode: pairs whose guarded class differs from the FP64 class (0)
    CNT_SCAN_HINT = 21,       // a cell in which the last overlap count found an overlap
    CNT_NUM = 24
};

// Move counts of the sweep's workgroups: the workgroup of cell c stores its counts in slot c (SLOT_WORDS words) with
// plain stores -- every cell is active once per cycle, so each slot is written at most once between folds -- and
// launch_fold_slots, once per cycle, adds the slots of the cycle's grid into the counters and clears them.  A plain
// store completes in the L2 of the workgroup's XCD; a device-scope atomic by every workgroup at the end of each
// colour pass held the end of the launch back by about 2 us.
enum SlotWord : int {
    SLOT_ACCEPTED = 0,
    SLOT_ATTEMPTED = 1,
    SLOT_LEFT_CELL = 2,
    SLOT_TESTS = 3,
    SLOT_G32_PAIRS = 4,
    SLOT_WORDS = 5
};
struct FoldSlots {
    unsigned long long *slots;     // nullptr: nothing to fold
    int cells;
    unsigned long long *counters;
};
hipError_t launch_fold_slots(hipStream_t s, unsigned long long *slots, int ncell, unsigned long long *counters);

// Launch wrappers (rp_kernels.hip).  All enqueue on `s` and return the launch status.
hipError_t launch_pair_decisions(hipStream_t s, PairConsts k, size_t n, const double *rec, uint8_t *out);
hipError_t launch_pair_decisions_guarded(hipStream_t s, PairConsts k, size_t n, const double *rec, uint8_t *out,
                                         uint8_t *code, float *errC, float *errS);
// Guard statistics of n decisions of launch_pair_decisions_guarded (with codes) into counters[0..5]: fallbacks,
// segment stages, segment fallbacks, mismatches, largest centre and segment errors (float bits); counters start at 0.
hipError_t launch_guard_stats(hipStream_t s, size_t n, const uint8_t *code, const float *errC, const float *errS,
                              unsigned long long *counters);
// The current device's fixed scratch area of the allocation-free guarded decisions (RP_GUARD_NOALLOC_MAX_PAIRS
// pairs): counters (8), centre errors, segment errors and codes.  The device address is looked up once per device.
constexpr size_t GUARD_SCRATCH_PAIRS = RP_GUARD_NOALLOC_MAX_PAIRS;
constexpr size_t GUARD_SCRATCH_BYTES = 256 + GUARD_SCRATCH_PAIRS * (2 * sizeof(float) + 1);
struct GuardScratch {
    unsigned long long *counters;
    float *errC, *errS;
    uint8_t *code;
};
hipError_t guard_scratch(GuardScratch &out);
// Start / stop events carried by a launch's own dispatch (hipExtLaunchKernel): the kernel's begin and end times
// are recorded with no marker packets between kernels.  Null events: an untimed launch.
struct Timed {
    hipEvent_t start = nullptr, stop = nullptr;
};
// Absolute bead centres (Packing.cpp:1036-1047) of the n particles of src into items of dst (n * k.nc items: x, y,
// z and tag = item index); g supplies the box (dims, inv) for the periodic correction.
hipError_t launch_centres(hipStream_t s, const Particles &src, int n, const PairConsts &k, const Grid &g,
                          const Particles &dst);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, size_t n, unsigned long long v);
// counters[CNT_OVERLAPS] = counters[CNT_STAGE_OVERFLOW] = 0 in one launch.
hipError_t launch_clear_count(hipStream_t s, unsigned long long *counters);
hipError_t launch_fill_i32(hipStream_t s, int *p, size_t n, int v);
// Cell-list rebuild: src -> dst sorted by cell of `g`.  When rescale != 0 the positions are first mapped
// newDims * (oldInv * pos) (Packing.cpp:183-184) with newDims = g.dims.
// scanTmp: rebuild_scan_ints(g.ncell) ints (none for small grids).
hipError_t launch_rebuild(hipStream_t s, const Grid &g, int n, Particles src, Particles dst, int *tmpSlot,
                          int *tmpOrder, double *tmpPos, int rescale, const double *oldInv, Timed t, int *scanTmp);
size_t rebuild_scan_ints(int ncell);
// Half-shell overlap count over the grid (Packing.cpp:431-559).  Adds to counters[CNT_OVERLAPS]; with
// earlyExit the workgroups stop once any overlap is found.  cap: LDS entries per workgroup (>= largest cell).
// startCell: the first cell counted (the cells follow in index order, wrapping around).
hipError_t launch_count_overlaps(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, int earlyExit,
                                 unsigned long long *counters, int cap, Timed t = {}, int startCell = 0);
size_t count_lds_bytes(int kind, int cap);
// Every pair i < j of the n particles (a grid with allPairs set) through the nearest image, the particle with the lower
// caller index the first operand (Packing::countTotalOverlaps without a grid, Packing.cpp:547-556); adds to
// counters[CNT_OVERLAPS], with earlyExit stopping once any is found.
hipError_t launch_count_allpairs(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, int earlyExit,
                                 unsigned long long *counters);
// Dynamic LDS of the all-pairs sweep that stages n particles and its proposals.
size_t allpairs_lds_bytes(int n);
// Proposals of one cycle: `moves` per cell when moves > 0, else ms.n per particle of its cell (cellOf = cell of every
// sorted record, n records).
hipError_t launch_proposals(hipStream_t s, const Grid &g, const Proposals &pr, int moves, const int *cellOf, int n,
                            const MoveSet &ms, uint64_t seed, uint32_t cycle, Timed t, FoldSlots fold);
// The sweep's move arithmetic on n explicit records (device arrays): pos + t wrapped into the box, the rotation matrix
// of (axis, angle), R O, and sin / cos of the angle.
hipError_t launch_apply_moves(hipStream_t s, const double *dims, const double *inv, size_t n, const double *pos,
                              const double *rot, const double *t, const double *axis, const double *angle,
                              double *outPos, double *outR, double *outRot, double *outSc);
// Adds each proposal's attempt and acceptance (pr.type, pr.acc; `total` proposals) to the per-sampler counters.
hipError_t launch_type_counts(hipStream_t s, const Proposals &pr, size_t total, unsigned long long *counters);
// waves: wavefronts per active cell (1, 2 or 4)
// Soft pair energies on records pos1[3] O1[9] pos2[3] O2[9] T[3].
hipError_t launch_pair_energies(hipStream_t s, PairConsts k, size_t n, const double *rec, double *out);

// Total soft energy over the half-shell pairs of grid g (Packing::getTotalEnergy, Packing.cpp:690-709):
// cellEnergy (g.ncell doubles) receives one partial per cell, scratch (fixed_sum_scratch_doubles()) the fixed-order
// sum, whose last element is the total.  counters[CNT_STAGE_OVERFLOW] counts cells too full to stage.
hipError_t launch_total_energy(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, double *cellEnergy,
                               double *scratch, unsigned long long *counters, int cap);
size_t fixed_sum_scratch_doubles();

// Sums of a_i a_j over the primary axes (orientation column 2), Q_11 Q_12 Q_13 Q_22 Q_23 Q_33 order, written to
// scratch[6 * 256 .. 6 * 256 + 6); scratch holds nematic_scratch_doubles() doubles.
size_t nematic_scratch_doubles();
hipError_t launch_nematic_sums(hipStream_t s, const Particles &p, int n, double *scratch);

// Applies fix_rotation to every orientation (no overlap test: soft shapes have no hard part).
hipError_t launch_fix_all(hipStream_t s, const Particles &p, int n);

// Candidate pairs of the learned ordering for M sampled trial moves on the state of sp (polyspheres; sp.g the sweep
// grid the records are sorted by): mode 0 writes counts[q], mode 1 the pairs of move q from offsets[q] on
// (x: 12 features, y: exact decision, meta: move, natural rank, nearest-first key).
hipError_t launch_mlp_pairs(hipStream_t s, const SweepParams &sp, int M, uint64_t seed, uint32_t stream, double ts,
                            double rs, int mode, int *counts, const long long *offsets, float *x, float *y,
                            int *meta);

// Workgroups of the sweep kernel one compute unit can hold at this LDS size (occupancy API).
int sweep_occupancy(int kind, int cap, int waves);
// g32: 0 FP64 stencil scan; 1 guarded FP32 scan; 2 guarded FP32 scan that also evaluates every pair in FP64 and
// records the largest FP32 error and any class mismatch (counters CNT_G32_*).  g32 > 0 needs eight wavefronts, the
// grouped launch and a hard sphere or spherocylinder.
hipError_t launch_sweep(hipStream_t s, const SweepParams &sp, int waves, Timed t = {}, int g32 = 0);
size_t sweep_lds_bytes(int kind, int cap, int waves, int g32 = 0);
// The kernel launch_sweep(sp, waves, g32) would run, as a key shared by the packings one batched launch may hold, or
// -1 for the sweeps that run one packing per launch (soft, overlap counting, orientation fix, polyspheres, traced).
int sweep_batch_key(const SweepParams &sp, int waves, int g32);
// One colour pass of `packings` packings of one key in one launch: table (device) holds each packing's parameters,
// first (device, packings + 1 entries) the first workgroup of each packing; blocks = first[packings]; lds the largest
// sweep_lds_bytes of the packings.
hipError_t launch_sweep_batch(hipStream_t s, int key, const SweepParams *table, const int *first, int packings,
                              unsigned blocks, size_t lds);

}  // namespace rp
