/* rampack_amd.h -- C ABI of the MI355X-native rampack Monte Carlo sweep.
 *
 * The boundary sits at sweep granularity: the reference calls its physics once per pair through virtual
 * functions (Interaction::overlapBetween, src/core/Interaction.h:81-90; Packing::countParticleOverlaps,
 * src/core/Packing.cpp:373-429), which cannot cross a device boundary per call.  The plug-ins in front of the
 * path (ShapeTraits / Interaction, MoveSampler, TriclinicBoxScaler) are exported as plain descriptors, and one
 * call runs whole cycles on the GPU.
 *
 * Conventions (src/utils/Exceptions.h:19-80, src/core/Packing.h:305-339):
 *   0 on success; RP_ERR_PRECONDITION where the reference throws PreconditionException; RP_ERR_VALIDATION where
 *   it throws ValidationException (e.g. overlaps at the start of integration, src/core/Simulation.cpp:131-132).
 *   Overlaps are not errors.  rp_last_error() describes the last failure of the calling thread.
 *   Host arrays are owned by the caller and copied in and out.  One context is used by one host thread at a time.
 *   Positions are n x 3 doubles, orientations n x 9 doubles (row-major 3x3, Shape.h:20-23), boxes 9 doubles
 *   (row-major matrix whose columns are the box sides, TriclinicBox.h:49-58).
 */
#ifndef RAMPACK_AMD_H
#define RAMPACK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP_OK 0
#define RP_ERR_PRECONDITION (-1)
#define RP_ERR_VALIDATION (-2)
#define RP_ERR_DEVICE (-3)
#define RP_ERR_UNSUPPORTED (-4)

/* Shape kinds: SphereTraits (src/core/shapes/SphereTraits.h:17-70), SpherocylinderTraits
 * (src/core/shapes/SpherocylinderTraits.h:16-86), PolysphereTraits (src/core/shapes/PolysphereTraits.h, hard beads
 * at fixed body-frame positions, PolysphereTraits.cpp:97-106). */
#define RP_SHAPE_SPHERE 0
#define RP_SHAPE_SPHEROCYLINDER 1
#define RP_SHAPE_POLYSPHERE 2
#define RP_MAX_CENTRES 16       /* beads per polysphere */

/* Soft interactions, which replace the hard core:
 * RP_SOFT_WCA -- a sphere with RepulsiveLennardJonesInteraction (src/core/interactions/
 *   RepulsiveLennardJonesInteraction.cpp:8-16 through CentralInteraction.h:58-67): WCA of the centre distance,
 *   range sigma 2^(1/6);
 * RP_SOFT_WCA_SEGMENT -- a spherocylinder whose energy is the WCA of the squared distance between the two axis
 *   segments (no counterpart in the reference), range L + sigma 2^(1/6). */
#define RP_SOFT_NONE 0
#define RP_SOFT_WCA 1
#define RP_SOFT_WCA_SEGMENT 2
/* RP_SOFT_LJ -- a sphere with LennardJonesInteraction (src/core/interactions/LennardJonesInteraction.cpp:8-13):
 *   4 eps ((s/r)^12 - (s/r)^6), range 3 sigma (LennardJonesInteraction.h:25).  The reference adds it for every pair
 *   its neighbour grid visits, so its truncation depends on the grid; sweeps and total energies here truncate at the
 *   range (pairs at and beyond 3 sigma contribute 0), rp_pair_energies does not truncate.
 * RP_SOFT_SIC -- a sphere with SquareInverseCoreInteraction (SquareInverseCoreInteraction.cpp:13-15):
 *   eps max(0, sigma^2 / r^2 - 1), range sigma; eps may be negative. */
#define RP_SOFT_LJ 3
#define RP_SOFT_SIC 4
/* RP_SOFT_HARD_CORE -- or-ed with a soft kind: CompoundInteraction of the shape's hard interaction and the soft one
 *   (src/core/interactions/CompoundInteraction.cpp:9-60): a pair overlaps when the hard shapes do, its energy is the
 *   soft part's, the range is the larger of the two, and the compound is convex (the soft part has no hard core).
 *   E.g. RP_SOFT_WCA_SEGMENT | RP_SOFT_HARD_CORE: hard spherocylinders that also repel by WCA on the segment distance. */
#define RP_SOFT_HARD_CORE 0x100

/* Precision of the overlap tests.  RP_PREC_FP64 evaluates every test in the reference's FP64 operation order.
 * RP_PREC_GUARDED_FP32 (no counterpart in the reference): the stencil scan of the default hard spherocylinder sweep
 * (eight wavefronts per cell, grouped moves; other launches and shapes keep FP64) takes squared centre distances in
 * FP32 from positions
 * relative to the active cell's centre, with rigorous bounds on their error; a pair whose FP32 value lies within that
 * bound of contact or of the interaction range is decided by the FP64 value, so every decision equals RP_PREC_FP64's.
 * Box, inverse, positions and the segment-distance stage stay FP64.  RP_PREC_GUARDED_FP32_VERIFY also evaluates every
 * pair in FP64 and records the largest FP32 error and any class that differs (rp_get_guard_stats). */
#define RP_PREC_FP64 0
#define RP_PREC_GUARDED_FP32 1
#define RP_PREC_GUARDED_FP32_VERIFY 2

/* Box moves: TriclinicDeltaScaler with scale_together (src/core/volume_scalers/TriclinicDeltaScaler.cpp:7-25),
 * log isotropic through TriclinicAdapter + AnisotropicVolumeScaler + LogScalingFactorSampler
 * (src/core/volume_scalers/TriclinicAdapter.h:27-33, AnisotropicVolumeScaler.cpp:18-33,
 * LogScalingFactorSampler.cpp:8-19). */
#define RP_BOX_NONE 0
#define RP_BOX_DELTA_TRICLINIC 1
#define RP_BOX_LOG_ISOTROPIC 2  /* log scaling; anisotropic with rp_box_desc.groups */
#define RP_BOX_DELTA_V 3        /* TriclinicAdapter + DeltaVolumeScaler (src/core/volume_scalers/DeltaVolumeScaler.cpp:10-19) */
#define RP_BOX_LINEAR 4         /* TriclinicAdapter + AnisotropicVolumeScaler + LinearScalingFactorSampler
                                 * (src/core/volume_scalers/LinearScalingFactorSampler.cpp:8-22) */

typedef struct rp_ctx rp_ctx;

/* ShapeTraits / Interaction descriptor (replaces the virtual Interaction, src/core/Interaction.h:18-150). */
typedef struct rp_shape_desc {
    int kind;       /* RP_SHAPE_* */
    double radius;  /* sphere / spherocylinder radius */
    double length;  /* spherocylinder distance between cap centres */
    int soft;       /* RP_SOFT_*; RP_SOFT_NONE: hard core */
    double epsilon; /* WCA energy scale (> 0) */
    double sigma;   /* WCA length scale (> 0) */
    int n_centres;          /* polysphere: number of beads, 1..RP_MAX_CENTRES */
    const double *centres;  /* polysphere: n_centres body-frame bead positions (x, y, z), host memory */
    const double *radii;    /* polysphere: n_centres bead radii (> 0), host memory */
} rp_shape_desc;

/* Move samplers: RototranslationSampler (src/core/move_samplers/RototranslationSampler.cpp:22-49),
 * TranslationSampler (TranslationSampler.cpp:20-37), RotationSampler (RotationSampler.cpp:12-33). */
#define RP_MOVE_ROTOTRANSLATION 0
#define RP_MOVE_TRANSLATION 1
#define RP_MOVE_ROTATION 2
#define RP_MAX_MOVE_SAMPLERS 4

/* Move sampler descriptor (RototranslationSampler.h:17-60, TranslationSampler.h, RotationSampler.h).  kind 0 (the
 * zero default) is the rototranslation sampler. */
typedef struct rp_move_desc {
    double trans_step;     /* translation step (> 0) */
    double rot_step;       /* rotation step; <= 0 selects the automatic rule of RototranslationSampler.cpp:95-114 */
    double max_trans_step; /* 0: implicit, half the smallest box height (RototranslationSampler.cpp:78-85) */
    int kind;              /* RP_MOVE_*; the translation sampler ignores rot_step, the rotation sampler trans_step */
} rp_move_desc;

/* TriclinicBoxScaler descriptor (src/core/TriclinicBoxScaler.h:15-55).  The fields after `step` default to zero:
 * an isotropic log / linear scaler and a delta_triclinic scaler that moves all nine entries. */
typedef struct rp_box_desc {
    int kind;      /* RP_BOX_* */
    double step;   /* initial step size (> 0) */
    int groups[3]; /* log / linear: scaling group (0-2) of the box sides x, y, z -- AnisotropicVolumeScaler's list
                    * of ScalingDirections (AnisotropicVolumeScaler.cpp:13-31); sides with one group share a factor,
                    * all zero is isotropic; 3: the side is not scaled */
    int independent; /* log / linear with several groups: 1 = one group per move, drawn uniformly; 0 = a factor for
                      * every group on each move (AnisotropicVolumeScaler.cpp:16-27) */
    int one_side;    /* delta_triclinic: 1 = one box side per move (scaleTogether = false,
                      * TriclinicDeltaScaler.cpp:17-23); 0 = all nine entries */
} rp_box_desc;

/* Counters and step sizes after a call of rp_run_cycles (cumulative over the context's lifetime). */
typedef struct rp_counters {
    uint64_t cycles;
    uint64_t moves_attempted;
    uint64_t moves_accepted;
    uint64_t moves_left_cell;    /* rejected because the particle left its active cell */
    uint64_t box_attempted;
    uint64_t box_accepted;
    uint64_t box_scans;          /* box moves that ran the device-wide overlap scan */
    uint64_t box_too_small;      /* box moves rejected because a box height fell below two interaction ranges */
    uint64_t unstaged_blocks;    /* active cells whose stencil did not fit LDS and was read from HBM instead */
    uint64_t pair_tests;         /* candidate slots examined by the sweep */
    double trans_step, rot_step, box_step;
} rp_counters;

/* Event-timed sweep kernel statistics (filled when timing is on, see rp_set_timing). */
typedef struct rp_timing {
    uint64_t sweep_launches;
    double sweep_ms;             /* summed duration of the sweep kernel launches */
    uint64_t rebuild_launches;
    double rebuild_ms;           /* summed duration of the cell-list rebuilds */
    uint64_t scan_launches;
    double scan_ms;
    uint64_t proposal_launches;
    double proposal_ms;          /* summed duration of the per-cycle proposal launches */
    uint64_t lds_limit;          /* LDS bytes a workgroup may use on this device */
    uint64_t sweep_cap;          /* stencil entries staged per sweep workgroup in the last cycle */
    uint64_t sweep_waves;        /* wavefronts per active cell in the last cycle */
    uint64_t sweep_period;       /* checkerboard colour period (colours = period^3) in the last cycle */
} rp_timing;

/* Observables of the current state (src/core/observables/PackingFraction.h:19-23, NematicOrder.cpp:9-60). */
typedef struct rp_observables {
    double packing_fraction;  /* N v0 / V (Packing::getPackingFraction, Packing.cpp:225-232) */
    double number_density;    /* N / V */
    double volume;            /* |(s1 x s2) . s3| (TriclinicBox::getVolume, TriclinicBox.cpp:35-38) */
    double nematic_p2;        /* largest-magnitude eigenvalue of Q; NaN for spheres (no primary axis) */
    double q_tensor[6];       /* Q_11 Q_12 Q_13 Q_22 Q_23 Q_33 (NematicOrder::getIntervalValues order) */
} rp_observables;

/* Packing constructor + Simulation constructor (src/core/Packing.cpp:46-82, src/core/Simulation.cpp:54-76):
 * uploads n shapes into a triclinic periodic box.  Every position must lie inside the box (Packing.cpp:1226-1232).
 * seed keys the counter-based random streams. */
int rp_create(rp_ctx **out, const rp_shape_desc *shape, const double box[9], size_t n, const double *pos,
              const double *rot, uint64_t seed, int precision);

/* Simulation::Environment (src/core/Simulation.h:113-165): temperature, pressure, the move sampler and the box
 * scaler (NULL or RP_BOX_NONE: NVT). */
int rp_set_environment(rp_ctx *ctx, double temperature, double pressure, const rp_move_desc *move,
                       const rp_box_desc *box);

/* The environment's list of move samplers (Simulation.h:113-165 holds a vector of MoveSampler): each cycle makes
 * n_moves * N trial moves, each with a sampler drawn uniformly (every sampler requests N moves,
 * Simulation.cpp:456-465, 683-691); each sampler adapts its steps on its own acceptance counter
 * (Simulation.cpp:561-607).  Replaces the sampler that rp_set_environment set; 1 <= n_moves <= RP_MAX_MOVE_SAMPLERS.
 * rp_get_move_samplers copies the samplers with their current step sizes (n entries at most). */
int rp_set_move_samplers(rp_ctx *ctx, const rp_move_desc *moves, size_t n_moves);
int rp_get_move_samplers(rp_ctx *ctx, rp_move_desc *out, size_t n, size_t *n_moves);

/* Simulation::performCycle x n_cycles (src/core/Simulation.cpp:325-363): N trial rototranslations as checkerboard
 * colour passes, then one box move; adapt_steps != 0 applies the thermalisation step adaptation of
 * Simulation.cpp:533-607.  Fails with RP_ERR_VALIDATION if the packing has overlaps (Simulation.cpp:131-132). */
int rp_run_cycles(rp_ctx *ctx, size_t n_cycles, int adapt_steps, rp_counters *out);

/* rp_run_cycles on n independent packings together: the replicas of an equation-of-state scan (the reference runs
 * one `rampack casino` per state point; Simulation::integrate, src/core/Simulation.cpp:92-176, per replica).  One
 * host thread and one stream (the first context's) carry every replica's work for the call: each colour pass of all
 * replicas runs as one launch per sweep kernel (hard spheres and spherocylinders; the other sweeps one launch per
 * replica), and one synchronisation per cycle decides every replica's box move.  Each replica ends bit for bit where
 * rp_run_cycles alone takes it.  out (nullable) receives n counter records.  A context may appear only once, all
 * must have an environment and live on one device (RP_ERR_PRECONDITION otherwise; messages name the replica). */
int rp_run_replicas(rp_ctx *const *ctxs, size_t n, size_t n_cycles, int adapt_steps, rp_counters *out);

/* Decision record of one cycle's particle moves (no counterpart in the reference: a test hook through which a CPU
 * restatement of Simulation::tryMove + Packing::tryMove / countParticleOverlaps / acceptMove (src/core/
 * Simulation.cpp:447-491, src/core/Packing.cpp:136-164, 373-429, 278-302) replays every decision of the device sweep).
 * rp_record_sweep runs the particle moves of one cycle (every checkerboard colour pass, no box move) and returns
 * the state before and after them in the sweep's storage order (records sorted by checkerboard cell), each
 * proposal (particle index within its cell, translation, rotation matrix) and its acceptance flag.  A call whose
 * capacity is below rec->n_proposals fills rec, runs nothing and returns RP_ERR_PRECONDITION; the next call runs
 * the same grid and proposals.  Hard interactions outside overlap-counting mode only. */
typedef struct rp_sweep_record {
    double box[9];
    double shift[3];        /* checkerboard origin shift per axis, relative coordinates in [0, 1/cells) */
    int cells[3];           /* checkerboard cells per axis; cell index = cx + cells[0] (cy + cells[1] cz) */
    int period;             /* colour period: a colour's cells have coordinates = (px, py, pz) mod period */
    int n_colours;
    int colour_order[27];   /* colour passes in launch order, colour = px + period (py + period pz) */
    int moves_per_cell;     /* > 0: that many proposals per cell, cell c's at [c * moves_per_cell, ...); 0: n_samplers
                             * per record of the cell, cell c's at [n_samplers * (first record of c), ...) */
    int n_samplers;
    size_t n_proposals;
} rp_sweep_record;
int rp_record_sweep(rp_ctx *ctx, size_t capacity, rp_sweep_record *rec, double *pre_pos, double *pre_rot,
                    int *pre_tag, int *pre_cell, int *prop_k, double *prop_t, double *prop_rot, uint8_t *accepted,
                    double *post_pos, double *post_rot);

/* Packing::countTotalOverlaps (src/core/Packing.cpp:519-559) on the reference's own neighbour grid
 * (Packing.cpp:858-899); early_exit returns 0 or 1. */
int rp_count_total_overlaps(rp_ctx *ctx, int early_exit, uint64_t *out);

/* Packing::getTotalEnergy (src/core/Packing.cpp:690-709): the soft energy summed over every pair once on the
 * reference's neighbour grid; 0 for hard shapes.  The device sums per cell and then in a fixed order, so a state
 * always gives the same bits; the reference's OpenMP reduction order differs, so values agree to the rounding of the
 * sum. */
int rp_total_energy(rp_ctx *ctx, double *out);

/* Packing::tryScaling(newBox, interaction) (src/core/Packing.cpp:172-208): positions become
 * newBox * (oldBox^-1 * pos) and the scaled state is current on return.  *dE receives +INFINITY if the scaled
 * packing has an overlap (the early-exit count, skipped when the box only grew), in overlap-counting mode
 * -INFINITY / +INFINITY when the full count fell / grew, otherwise getTotalEnergy(new) - getTotalEnergy(old)
 * (0 for hard shapes).  The state before the call is kept for rp_revert_scaling.  RP_ERR_PRECONDITION for a
 * zero-volume box (Expects at Packing.cpp:173); RP_ERR_UNSUPPORTED when a box height is below two interaction
 * ranges (the checkerboard grid needs two cells per axis).
 * rp_revert_scaling: Packing::revertScaling (Packing.cpp:363-371), the state saved by the last rp_try_scaling
 * (box, particles, cached overlap count) becomes current; RP_ERR_PRECONDITION if no scaling was tried. */
int rp_try_scaling(rp_ctx *ctx, const double new_box[9], double *dE);
int rp_revert_scaling(rp_ctx *ctx);

/* Packing::renormalizeOrientations (src/core/Packing.cpp:1178-1216) as Simulation::fixRotationMatrices runs it
 * (Simulation.cpp:750-769): every orientation becomes three iterations of R <- 1.5 R - 0.5 R R^T R unless that
 * creates an overlap; *rejected receives the number of particles left unchanged.  rp_run_cycles runs it after every
 * `every`-th cycle (the reference's rotation_matrix_fix_every, default here 10000; 0 disables). */
int rp_fix_orientations(rp_ctx *ctx, uint64_t *rejected);
int rp_set_orientation_fix_every(rp_ctx *ctx, uint64_t every);

/* Overlap-counting mode (Packing::toggleOverlapCounting, src/core/Packing.cpp:1049-1060; moves as
 * Packing::calculateMoveOverlapEnergy, Packing.cpp:304-323; box moves as Packing::tryScaling's counting branch,
 * Packing.cpp:190-198): a particle move is accepted if it does not raise the particle's overlap count, and the
 * overlap count is kept up to date; rp_run_cycles does not reject an overlapping packing while it is on.
 * rp_cached_overlaps reads the count (an error when counting is off, as the reference throws). */
int rp_set_overlap_counting(rp_ctx *ctx, int enable);
int rp_cached_overlaps(rp_ctx *ctx, uint64_t *out);

/* Simulation::relaxOverlaps (src/core/Simulation.cpp:222-284): cycles with step adaptation in overlap-counting
 * mode until no overlap is left or max_cycles have run; counting is switched off afterwards. */
int rp_relax_overlaps(rp_ctx *ctx, size_t max_cycles, size_t *cycles_run, uint64_t *overlaps_left);

/* RAMSNAP snapshots, the reference's text snapshot format (RamsnapWriter::write, src/core/io/RamsnapWriter.cpp:8-48;
 * RamsnapReader::read, src/core/io/RamsnapReader.cpp:15-92).  aux_text holds "key value" lines (the key is the
 * first whitespace-free token, the value the rest of the line); entries are written sorted by key.  A broken file
 * is RP_ERR_VALIDATION (RamsnapException is a ValidationException). */
int rp_ramsnap_write(const char *path, const double box[9], size_t n, const double *pos, const double *rot,
                     const char *aux_text);
int rp_ramsnap_save(rp_ctx *ctx, const char *path, const char *aux_text);  /* the device-resident state */
/* Number of particles and bytes of aux text (with the terminating NUL) of a RAMSNAP file. */
int rp_ramsnap_info(const char *path, size_t *n, size_t *aux_bytes);
int rp_ramsnap_read(const char *path, double box[9], size_t n, double *pos, double *rot, char *aux_text,
                    size_t aux_cap);

/* RAMTRJ trajectories, the reference's binary trajectory format (RamtrjIO, src/core/io/RamtrjIO.cpp:16-105): a
 * 33-byte header, then per snapshot the box (9 doubles) and per particle its position and the extrinsic XYZ
 * Tait-Bryan angles of its orientation (EulerAngles::forMatrix, src/geometry/EulerAngles.cpp:12-45; read back with
 * Matrix<3,3>::rotation, src/geometry/Matrix.tpp:408-433).  Files are byte-identical to the reference's.
 *
 * Recorder -- RamtrjRecorder (src/core/io/RamtrjRecorder.cpp:8-66) opened as SimulationRecorderFactory does
 * (src/frontend/SimulationRecorderFactory.cpp:13-31): append 0 truncates the file, append 1 continues a recording
 * with the same particle count and cycle step.  Snapshot k (1, 2, ...) must be recorded at cycle k * cycle_step;
 * an orientation that is not a rotation matrix to 1e-12 is a precondition error (EulerAngles.cpp:47-53).  Closing
 * writes the snapshot count into the header. */
typedef struct rp_ramtrj_recorder rp_ramtrj_recorder;
int rp_ramtrj_recorder_open(rp_ramtrj_recorder **out, const char *path, size_t num_particles, size_t cycle_step,
                            int append);
int rp_ramtrj_record(rp_ramtrj_recorder *rec, const double box[9], size_t n, const double *pos, const double *rot,
                     size_t cycle);
int rp_ramtrj_record_packing(rp_ramtrj_recorder *rec, rp_ctx *ctx, size_t cycle);  /* the device-resident state */
int rp_ramtrj_recorder_last_cycle(rp_ramtrj_recorder *rec, size_t *cycle);      /* getLastCycleNumber */
int rp_ramtrj_recorder_close(rp_ramtrj_recorder *rec);                           /* also frees rec */

/* Player -- RamtrjPlayer (src/core/io/RamtrjPlayer.cpp:11-128) opened as IO::loadRamtrjPlayer does
 * (src/frontend/IO.cpp:32-64).  expected_particles 0 reads strictly: a file whose length disagrees with its header
 * is RP_ERR_VALIDATION.  Otherwise a truncated file (a recorder that never closed) is repaired in memory as
 * RamtrjPlayer::AutoFix does (RamtrjPlayer.cpp:146-189) and *fix, if given, reports it.  next_snapshot reads the
 * following snapshot; jump_to_snapshot the one recorded at `cycle` (a positive multiple of the cycle step). */
typedef struct rp_ramtrj_player rp_ramtrj_player;
typedef struct rp_ramtrj_header {
    uint8_t version_major, version_minor;
    uint64_t num_particles, num_snapshots, cycle_step;
    uint64_t current_snapshot; /* snapshots read; current_snapshot * cycle_step = getCurrentSnapshotCycles */
} rp_ramtrj_header;
typedef struct rp_ramtrj_fix {
    int fixing_needed, fixing_successful;
    uint64_t header_snapshots, inferred_snapshots, bytes_remainder, bytes_per_snapshot;
} rp_ramtrj_fix;
int rp_ramtrj_player_open(rp_ramtrj_player **out, const char *path, size_t expected_particles, rp_ramtrj_fix *fix);
int rp_ramtrj_player_header(rp_ramtrj_player *player, rp_ramtrj_header *out);
int rp_ramtrj_next_snapshot(rp_ramtrj_player *player, double box[9], size_t n, double *pos, double *rot);
int rp_ramtrj_jump_to_snapshot(rp_ramtrj_player *player, size_t cycle, double box[9], size_t n, double *pos,
                               double *rot);
void rp_ramtrj_player_close(rp_ramtrj_player *player);

/* Current state, in the caller's particle order (Packing::operator[], Packing::getBox). */
int rp_download(rp_ctx *ctx, double box[9], double *pos, double *rot);

/* PackingFraction::calculate and NematicOrder::calculate on the device-resident state: the Q-tensor sum runs on the
 * GPU (a fixed reduction tree, so the same state always gives the same bits; the reference sums sequentially, so
 * results agree to the rounding of an N-term sum), the 3x3 eigenvalue step on the host. */
int rp_get_observables(rp_ctx *ctx, rp_observables *out);

/* Event timing of the cycle's kernels (no reference counterpart; the reference's CasinoMode.cpp:400-442 timers
 * measure whole phases on the host).  enable: 0 off; 1 every sweep, rebuild, overlap-scan and proposal launch;
 * RP_TIMING_SAMPLED one sweep launch per cycle (the colour pass whose index is the cycle number modulo the colour
 * count), which leaves the other launches untouched.  rp_get_timing reads the accumulated statistics. */
#define RP_TIMING_SAMPLED 2
int rp_set_timing(rp_ctx *ctx, int enable);
int rp_get_timing(rp_ctx *ctx, rp_timing *out);

/* Launch-shape autotuner (no counterpart in the reference, whose domain decomposition has no launch shape): runs
 * cycles_per_shape cycles (no step adaptation) with each candidate sweep launch -- wavefronts per active cell and
 * the move schedule (0 plain, 1 planner/tester pipeline, 2 grouped moves) -- timing every sweep launch, and keeps
 * the fastest for the cycles that follow.  Every launch shape makes the same decisions, so the call is
 * rp_run_cycles(ctx, candidates x cycles_per_shape, 0, NULL) bit for bit, whatever shape it keeps.  Candidates: hard
 * spheres and spherocylinders all six tested shapes (the guarded FP32 precision only its grouped eight-wavefront
 * launch), polyspheres 1, 2, 4, 8 or 16 wavefronts; soft and overlap-counting sweeps keep their shape.  out (nullable,
 * RP_AUTOTUNE_MAX entries) receives each candidate and its mean sweep time per cycle, *n_out their number and
 * *chosen (nullable) the index kept. */
#define RP_AUTOTUNE_MAX 8
typedef struct rp_autotune_entry {
    int waves;
    int schedule;
    double sweep_us_per_cycle;
} rp_autotune_entry;
int rp_autotune(rp_ctx *ctx, size_t cycles_per_shape, rp_autotune_entry *out, size_t *n_out, size_t *chosen);

void rp_destroy(rp_ctx *ctx);

/* Batched pair decisions on the device: SpherocylinderTraits::overlapBetween (SpherocylinderTraits.cpp:28-46) or
 * SphereTraits::HardInteraction::overlapBetween (SphereTraits.cpp:56-65), called with a fixed cell translation as
 * Packing does (Packing.cpp:21-44, 381-398).  Each record is 27 doubles: pos1[3] O1[9] pos2[3] O2[9] T[3].
 * decisions[i] = 1 for an overlap. */
int rp_pair_decisions(const rp_shape_desc *shape, size_t n, const double *records, uint8_t *decisions);

/* Guarded FP32 form of rp_pair_decisions (no counterpart in the reference; its FP64 decisions are the target).  Each
 * pair's centre distance, and for spherocylinders within range the segment distance, is evaluated in FP32 with a
 * rigorous bound on its error against the FP64 value; a pair farther from the threshold than the bound is decided in
 * FP32, any other pair by the FP64 test, so every decision equals rp_pair_decisions'.  codes (n bytes, or NULL):
 * bits 0-1 = 0 FP32 no overlap, 1 FP32 overlap, 2 FP64 fallback; bit 2 = the segment stage ran; bit 3 = differs
 * from FP64 (never set).  stats (or NULL) counts the fallbacks and reports the largest FP32 error.  With codes and
 * stats both NULL only the undecided pairs are evaluated in FP64 (the form to time); otherwise every pair is, for the
 * statistics. */
typedef struct rp_guard_stats {
    uint64_t pairs;
    uint64_t fp64_fallback;      /* pairs inside a guard band, decided by the FP64 test */
    uint64_t segment_stage;      /* pairs that reached the FP32 segment-distance stage */
    uint64_t segment_fallback;   /* of those, pairs inside the segment guard band */
    uint64_t mismatches;         /* guarded decision != FP64 decision: 0 */
    double max_err_centre;       /* max |FP32 - FP64| squared centre distance */
    double max_err_segment;      /* max |FP32 - FP64| squared segment distance (FP32 at its own (s, t)) */
} rp_guard_stats;
int rp_pair_decisions_guarded(const rp_shape_desc *shape, size_t n, const double *records, uint8_t *decisions,
                              uint8_t *codes, rp_guard_stats *stats);

/* rp_pair_decisions_guarded on device arrays, allocating nothing: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * d_records (27 n doubles), d_decisions (n bytes) and d_codes (n bytes, or NULL) are device pointers on the current
 * device; the same decisions, codes and statistics as rp_pair_decisions_guarded.  The statistics' per-pair errors
 * (and the codes when d_codes is NULL) live in a fixed scratch area of the library on each device, which holds
 * RP_GUARD_NOALLOC_MAX_PAIRS pairs: a larger n is refused (RP_ERR_PRECONDITION).  The area is shared by every call
 * on that device, so calls that overlap on one device race.  The call is synchronous (the null stream is
 * synchronised before it returns). */
#define RP_GUARD_NOALLOC_MAX_PAIRS ((size_t)1 << 20)
int rp_pair_decisions_guarded_noalloc(const rp_shape_desc *shape, size_t n, const double *d_records,
                                      uint8_t *d_decisions, uint8_t *d_codes, rp_guard_stats *stats);

/* Precision of a context's sweeps (RP_PREC_*, as rp_create's argument), and the guarded sweep's statistics since the
 * context was created: pairs classified in FP32, fp64_fallback = pairs inside a guard band, and in verify mode
 * mismatches (0) and max_err_centre (largest |FP32 - FP64| squared centre distance); the segment fields stay 0 (the
 * segment stage is FP64). */
int rp_set_precision(rp_ctx *ctx, int precision);
int rp_get_guard_stats(rp_ctx *ctx, rp_guard_stats *out);

/* The box quantities the device path derives on the host (TriclinicBox, src/core/TriclinicBox.h:49-70 and
 * TriclinicBox.cpp:27-51): the inverse of the box matrix (Matrix::inverse, src/geometry/Matrix.tpp:258-271), the three
 * heights and the volume. */
int rp_box_info(const double box[9], double inv[9], double heights[3], double *volume);

/* The sweep's move arithmetic on explicit inputs, evaluated on the device by the functions the sweep kernels use:
 * for each of n moves the trial position pos + t wrapped into the box (Shape::translate with
 * PeriodicBoundaryConditions::getCorrection, src/core/Shape.cpp:9-12, PeriodicBoundaryConditions.cpp:11-22), the
 * rotation matrix of axis / |axis| and angle (Matrix<3,3>::rotation as RototranslationSampler::sampleMove calls it,
 * src/geometry/Matrix.tpp:438-450, RototranslationSampler.cpp:39-43), the trial orientation R O (Shape::rotate,
 * Shape.cpp:14-16) and the device's sin and cos of the angle.  Arrays: pos / t / axis / out_pos n x 3, rot / out_R /
 * out_rot n x 9 (row-major), angle n, out_sincos n x 2. */
int rp_apply_moves(const double box[9], size_t n, const double *pos, const double *rot, const double *t,
                   const double *axis, const double *angle, double *out_pos, double *out_R, double *out_rot,
                   double *out_sincos);

/* The learned candidate ordering (no counterpart in the reference; the reference tests candidates in cell order,
 * src/core/Packing.cpp:381-398).  A network 12 -> 64 -> 64 -> 64 -> 1 (GELU hidden units, tanh form; sigmoid
 * output) predicts from a candidate's pose relative to a trial pose -- its position in the trial body frame divided
 * by the total interaction range (3) and the relative rotation O_trial^T O_candidate (9, row-major) -- the
 * probability that the exact test finds an overlap.  RP_MLP_PARAMS floats of parameters: W1 [64][12], b1 [64],
 * W2 [64][64], b2 [64], W3 [64][64], b3 [64], w4 [64], b4.
 * rp_mlp_gradient: the device's FP32 forward and backward kernels on a batch of n pairs (x: n x 12, y: n labels 0/1):
 * the mean binary cross-entropy and its gradient (RP_MLP_PARAMS floats).
 * rp_mlp_infer: overlap probabilities of n pairs with BF16 MFMA layers and FP32 accumulation. */
#define RP_MLP_PARAMS 9217
int rp_mlp_gradient(const float *theta, size_t n, const float *x, const float *y, double *loss, float *grad);
int rp_mlp_infer(const float *theta, size_t n, const float *x, float *score);

/* Training and use of the ordering on a polysphere context.
 * rp_mlp_train: the run's own trajectory supplies the data -- `snapshots` times, cycles_between cycles of the run
 * (rp_run_cycles without adaptation), then moves_per_snapshot trial moves drawn as the sweep draws them with the first
 * sampler's steps, and every close candidate of each trial pose (within the total interaction range) as one pair
 * with the exact test's decision as its label.  Glorot-uniform weights, the output bias at the log-odds of the overlap
 * fraction, then Adam (learning_rate, batch pairs per step) for `epochs` passes in a seeded random order, all on the
 * device by the FP32 training kernels.
 * rp_mlp_order_stats: for `moves` fresh trial moves on the current state, the mean number of exact tests per move
 * that a search stopping at the first overlap makes in natural stencil order, in nearest-first order and in the
 * network's order (BF16 MFMA scores, most likely overlap first); the decisions (any overlap) are the same in every
 * order.  Moves that leave their cell or have no close candidate count 0 tests.
 * rp_mlp_pairs: the pairs of `moves` fresh trial moves (x: 12 features, y: exact decision, meta: move, natural rank,
 * nearest-first key) into caller arrays of `cap` pairs; *n receives the count (RP_ERR_PRECONDITION if above cap). */
typedef struct rp_mlp_stats {
    uint64_t pairs;              /* training pairs (rp_mlp_train) or evaluated pairs (rp_mlp_order_stats) */
    double overlap_fraction;     /* of those, the fraction whose exact test finds an overlap */
    double loss_initial, loss_final;  /* mean binary cross-entropy over the training pairs before / after training */
    uint64_t steps;              /* Adam steps */
    uint64_t moves, moves_rejected;
    double candidates_per_move;  /* close candidates per sampled move */
    double tests_natural, tests_nearest_first, tests_learned;  /* exact tests per sampled move, early exit */
} rp_mlp_stats;
int rp_mlp_train(rp_ctx *ctx, size_t snapshots, size_t cycles_between, size_t moves_per_snapshot, size_t epochs,
                 size_t batch, double learning_rate, rp_mlp_stats *out);
int rp_mlp_order_stats(rp_ctx *ctx, size_t moves, rp_mlp_stats *out);
int rp_mlp_get_weights(rp_ctx *ctx, float *theta);
int rp_mlp_set_weights(rp_ctx *ctx, const float *theta);
int rp_mlp_pairs(rp_ctx *ctx, size_t moves, size_t cap, float *x, float *y, int *meta, size_t *n);

/* Batched soft pair energies on the device: Interaction::calculateEnergyBetween with a fixed cell translation
 * (CentralInteraction.h:58-67 for RP_SOFT_WCA); records as in rp_pair_decisions. */
int rp_pair_energies(const rp_shape_desc *shape, size_t n, const double *records, double *energies);

const char *rp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* RAMPACK_AMD_H */
