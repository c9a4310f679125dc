"""Packings the reference runs without its neighbour grid (Packing::rebuildNeighbourGrid, Packing.cpp:862-883: fewer
than four interaction-range cells across every direction, or a box thinner than one cell): every pair meets through
the nearest image only (PeriodicBoundaryConditions::getTranslation / getDistance2), the moves run in order on any
particle, and the box moves count overlaps over every pair.

* Overlap counts of random snapshots in such boxes (spheres and spherocylinders, triclinic) equal the reference's
  own Packing::countTotalOverlaps (oracle/_ref), full and early-exit.
* Sweep decisions replay through the oracle's restatement with no mismatch.
* A small packing at high pressure compresses from a box with a grid into one without (as the small replicas of an
  equation-of-state scan do): both paths run, the state stays overlap-free and the last sweep replays.
* Packings that fit in LDS are staged there by the all-pairs sweep; a slab of 1500 rods, too many, is read in place:
  both replay."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice
from test_gpu_replay import check_record

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
HS = ra.Shape("sphere", radius=0.5)

def lying_rod_slab():
    """1500 rods lying along x in two layers of a 97.5 x 60 x 5.5 box (spacing 6.5 along the rods, 1.2 across, the
    layers 2.75 apart): thinner than one interaction range (6), so without a grid, and too many rods for the all-pairs
    sweep to stage in LDS."""
    box, pos = lattice.sc_positions((15, 50, 2), (6.5, 1.2, 2.75))
    rot = np.tile(np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]]).reshape(1, 9), (len(pos), 1))
    return box, pos, np.ascontiguousarray(rot)


BOXES = {
    "sphere": (HS, 0, 0.5, 0.0, np.array([[3.6, 0.3, -0.2], [0.0, 3.4, 0.25], [0.0, 0.0, 3.1]]), 40),
    "spherocylinder": (SC, 1, 0.5, 5.0, np.array([[9.0, 0.6, -0.4], [0.0, 8.5, 0.5], [0.0, 0.0, 16.0]]), 60),
}


@pytest.mark.parametrize("name", ["sphere", "spherocylinder"])
def test_nogrid_counts_equal_reference(name):
    shape, kind, r, l, box, n = BOXES[name]
    ref = oracle_lib.ref()
    for seed in (1, 2, 3):
        pos, rot = corpus.random_snapshot(n, seed, box, kind=name)
        full, _ = ref.snapshot_stats(kind, r, l, box, pos, rot)
        early, _ = ref.snapshot_stats(kind, r, l, box, pos, rot, early_exit=True)
        pk = ra.GpuPacking(shape, box, pos, rot, seed=seed)
        got_full, got_early = pk.count_total_overlaps(), pk.count_total_overlaps(early_exit=True)
        pk.close()
        print(f"{name} seed {seed}: {got_full} overlapping pairs (reference {full})", flush=True)
        assert full > 0
        assert got_full == full and got_early == early


@pytest.mark.parametrize("name", ["sphere", "spherocylinder"])
def test_nogrid_sweep_decisions_replay(name):
    shape = BOXES[name][0]
    if name == "sphere":
        box, pos, rot = lattice.hard_sphere_lattice(3, 0.30)  # 27 spheres, every height about 3.1
    else:
        box = np.array([[7.5, 0.4, 0.0], [0.0, 7.0, 0.3], [0.0, 0.0, 14.5]])
        pos = np.array([[x, y, z] for x in (0.8, 2.6, 4.4, 6.2) for y in (0.9, 2.7, 4.5, 6.3) for z in (3.2, 10.4)])
        rot = np.tile(np.eye(3).reshape(9), (len(pos), 1))
    pk = ra.GpuPacking(shape, box, pos, rot, seed=3)
    pk.set_environment(1.0, 2.0, 0.35, 0.25 if name != "sphere" else None, box_move=("delta_triclinic", 0.05))
    for _ in range(2):
        rec = pk.record_sweep()
        assert rec["cells"] == [1, 1, 1]
        check_record(shape, rec)
    pk.run_cycles(200, adapt_steps=True)
    check_record(shape, pk.record_sweep())
    b, p, r = pk.download()
    kind = 0 if name == "sphere" else 1
    assert oracle_lib.oracle().snapshot_stats(kind, 0.5, BOXES[name][3], b, p, r)[0] == 0
    pk.close()


def test_small_packings_compress_into_boxes_without_a_grid():
    """Parallel rods in two layers, the box 24.2 high along the rods (just above the reference's grid threshold of
    four interaction ranges, 24) at p* = 4: the first box moves take it below the threshold, as happens to the small
    high-pressure replicas of an equation-of-state scan.  Both the checkerboard (period 2) and the all-pairs path
    (period 1) run, the state stays overlap-free and the last sweep replays."""
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(7, 7, 2), spacing=(1.6, 1.6, 12.1))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=5)
    pk.set_environment(1.0, 4.0, 0.3, 0.1, box_move=("log", 0.002))
    periods = []
    for _ in range(30):
        pk.run_cycles(1, adapt_steps=True)
        periods.append(pk.timing()["sweep_period"])
    for _ in range(20):
        pk.run_cycles(50, adapt_steps=True)
        periods.append(pk.timing()["sweep_period"])
    b, q, r = pk.download()
    print(f"phi {pk.packing_fraction():.4f}, box diagonal {np.round(np.diag(b), 2)}, grid periods {periods}",
          flush=True)
    assert periods[0] == 2 and periods[-1] == 1  # the checkerboard first, the all-pairs path once compressed
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, q, r)[0] == 0
    check_record(SC, pk.record_sweep())
    pk.close()


def test_nogrid_orientation_fix_matches_oracle():
    """Packing::renormalizeOrientations in a box without a grid: orientations perturbed by 1e-7 are replaced by the
    oracle's fix_rotation bit for bit (no fix is rejected: the rods stand apart), and periodic fixes during a run
    keep the packing overlap-free."""
    box = np.array([[7.5, 0.4, 0.0], [0.0, 7.0, 0.3], [0.0, 0.0, 14.5]])
    pos = np.array([[x, y, z] for x in (0.8, 2.6, 4.4, 6.2) for y in (0.9, 2.7, 4.5, 6.3) for z in (3.2, 10.4)])
    rng = np.random.Generator(np.random.PCG64(13))
    rot = np.ascontiguousarray(np.tile(np.eye(3).reshape(9), (len(pos), 1)) + (rng.random((len(pos), 9)) - 0.5) * 2e-7)
    pk = ra.GpuPacking(SC, box, pos, rot, seed=1)
    assert pk.fix_orientations() == 0
    _, p, r = pk.download()
    expected = np.ascontiguousarray(rot.copy())
    oracle_lib.oracle().lib.orc_fix_rotation(len(expected), oracle_lib._d(expected))
    order = np.lexsort(p.T[::-1])
    assert np.array_equal(p[order], pos[np.lexsort(pos.T[::-1])])
    assert np.array_equal(r[order], expected[np.lexsort(pos.T[::-1])])
    pk.set_environment(1.0, 2.0, 0.3, 0.2, box_move=("delta_triclinic", 0.05))
    pk.set_orientation_fix_every(7)
    pk.run_cycles(60, adapt_steps=True)
    b, p, r = pk.download()
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r)[0] == 0
    pk.close()


def test_slab_too_large_for_lds_replays():
    """The slab of 1500 lying rods (no grid: thinner than one interaction range) holds too many rods for the all-pairs
    sweep to stage in LDS, which then reads them in place: its decisions replay, box moves run and no overlap forms."""
    box, pos, rot = lying_rod_slab()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=23)
    pk.set_environment(1.0, 1.0, 0.3, 0.2, box_move=("delta_triclinic", 0.01))
    check_record(SC, pk.record_sweep())
    c = pk.run_cycles(40, adapt_steps=True)
    t = pk.timing()
    b, p, r = pk.download()
    print(f"slab: moves {c['moves_accepted']}/{c['moves_attempted']}, box {c['box_accepted']}/{c['box_attempted']}, "
          f"staged {t['sweep_cap']}", flush=True)
    assert t["sweep_period"] == 1 and t["sweep_cap"] == 0
    assert 0 < c["moves_accepted"] < c["moves_attempted"]
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r)[0] == 0
    check_record(SC, pk.record_sweep())
    pk.close()
