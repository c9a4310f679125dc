"""Polyspheres on the device: total overlap counts over the grid of bead centres equal the reference build's counts
(committed fixtures, full and early exit), pair decisions equal the oracle's bead-pair test, and the packing
fraction uses the summed bead volume."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from test_poly_oracle import records_for

pytestmark = pytest.mark.gpu


def poly_shape(beads):
    c, r = corpus.bead_set(beads)
    return ra.Shape("polysphere", centres=c, radii=r)


def test_device_counts_match_reference_fixtures(golden):
    for s in golden["poly_snapshots"]:
        box = np.array(s["box"])
        pos, rot = corpus.random_snapshot(s["n"], s["seed"], box)
        pk = ra.GpuPacking(poly_shape(s["beads"]), box, pos, rot, seed=1)
        full = pk.count_total_overlaps(early_exit=False)
        early = pk.count_total_overlaps(early_exit=True)
        print(f"{s['name']}: {full} (reference {s['overlaps']}), early {early}", flush=True)
        assert full == s["overlaps"] and early == s["overlaps_early_exit"]
        pk.close()


@pytest.mark.parametrize("beads", ["kinked11", "dimer", "mixed"])
def test_device_pair_decisions_match_oracle(beads):
    centres, radii = corpus.bead_set(beads)
    ext = float(np.sqrt((centres * centres).sum(1).max()))
    rec = records_for(centres, 20000, 11, 2 * ext + 1.5)
    dev = ra.pair_decisions(poly_shape(beads), rec)
    orc = oracle_lib.poly_pair_decisions(centres, radii, rec)
    assert 0 < orc.sum() < len(rec)
    assert np.array_equal(dev, orc)


def test_polysphere_observables_and_validation():
    c, r = corpus.bead_set("kinked11")
    box = np.diag([30.0, 30.0, 30.0])
    pos, rot = corpus.random_snapshot(200, 5, box)
    pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=1)
    obs = pk.observables()
    assert np.isclose(obs["packing_fraction"], 200 * 11 * np.pi / 6 / 27000.0, rtol=1e-14)
    pk.close()
    with pytest.raises(ra.PreconditionError):
        ra.GpuPacking(ra.Shape("polysphere", centres=c, radii=-r), box, pos, rot)
    with pytest.raises(ra.PreconditionError):
        ra.GpuPacking(ra.Shape("polysphere", centres=np.zeros((17, 3)), radii=np.ones(17)), box, pos, rot)


def chain_start(n_side=(8, 8, 2), spacing=(3.0, 3.0, 11.5)):
    """Kinked 11-bead chains on an orthorhombic lattice, all in the body frame's orientation (overlap-free: the
    oracle confirms it in the tests)."""
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in n_side], indexing="ij"), -1).reshape(-1, 3).astype(float)
    pos = (idx + 0.5) * np.array(spacing)
    box = np.diag(np.array(n_side) * np.array(spacing))
    rot = np.tile(np.eye(3).reshape(1, 9), (len(pos), 1))
    return box, np.ascontiguousarray(pos), np.ascontiguousarray(rot)


def poly_overlaps(beads, box, pos, rot, early=False):
    c, r = corpus.bead_set(beads)
    return oracle_lib.oracle().snapshot_stats(2, 0.0, 0.0, box, pos, rot, early_exit=early, centres=c, radii=r)[0]


def test_polysphere_sweep_stays_overlap_free():
    box, pos, rot = chain_start()
    assert poly_overlaps("kinked11", box, pos, rot) == 0
    pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=21)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.05))
    for _ in range(3):
        c = pk.run_cycles(20, adapt_steps=True)
        b, p, r = pk.download()
        assert poly_overlaps("kinked11", b, p, r) == 0
        assert pk.count_total_overlaps() == 0
    print(c, flush=True)
    assert c["moves_accepted"] > 0 and c["box_accepted"] > 0
    assert abs(c["moves_attempted"] - 60 * len(pos)) <= 0.15 * 60 * len(pos)
    # orientations stay orthonormal and the particles move
    R = r.reshape(-1, 3, 3)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-12)
    assert not np.allclose(p, pos)
    pk.close()


@pytest.mark.parametrize("waves", ["1", "2", "8", "16"])
def test_polysphere_launch_shapes_give_identical_trajectories(waves, monkeypatch):
    box, pos, rot = chain_start()
    runs = []
    for w in (waves, "4"):
        monkeypatch.setenv("RAMPACK_SWEEP_WAVES", w)
        pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=5)
        pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.05))
        pk.run_cycles(15)
        runs.append(pk.download())
        pk.close()
    for x, y in zip(*runs):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("period", ["", "2", "3"])
def test_polysphere_colourings_are_overlap_free(period, monkeypatch):
    """Both checkerboard periods on a box wide enough for period 3 (cells of half the total range, at least six
    per axis); without RAMPACK_COLOURING polyspheres take period 3."""
    monkeypatch.setenv("RAMPACK_COLOURING", period)
    box, pos, rot = chain_start((12, 12, 3), (3.0, 3.0, 11.5))
    pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=31)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.02))
    pk.set_timing(True)
    c = pk.run_cycles(20, adapt_steps=True)
    assert pk.timing()["sweep_period"] == int(period or 3)
    b, p, r = pk.download()
    assert poly_overlaps("kinked11", b, p, r) == 0 and pk.count_total_overlaps() == 0
    assert c["moves_accepted"] > 0 and not np.allclose(p, pos)
    pk.close()


def test_polysphere_try_scaling():
    box, pos, rot = chain_start()
    pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=8)
    start = pk.download()
    for f in (1.02, 0.999, 0.7):
        dE = pk.try_scaling(box * f)
        b, p, r = pk.download()
        n = poly_overlaps("kinked11", b, p, r, early=True)
        assert dE == (np.inf if n else 0.0), f  # polyspheres are not convex: every scaling is counted
        pk.revert_scaling()
        assert all(np.array_equal(x, y) for x, y in zip(pk.download(), start))
    assert pk.try_scaling(box * 0.7) == np.inf
    pk.close()


def test_dimer_npt_packing_fraction_matches_reference(golden):
    """NpT hard dumbbells against the reference's own runs of the same state point (fixture from the reference
    build's Packing with PolysphereTraits, samplers and log scaler, several seeds): the mean packing fraction over
    independent device runs agrees with the reference's mean over its runs within 2 standard errors of the
    difference, each side's error taken from the spread between its runs.  Both sides thermalise for 4e4 cycles: the
    packing fraction climbs for tens of thousands of cycles after the start, and shorter runs compare two
    unequilibrated states."""
    st = next((s for s in golden.get("statistics", []) if s["name"] == "dimer_npt"), None)
    if st is None or "phi_runs" not in st:
        pytest.skip("no reference statistics in the fixture")
    box, pos, rot = corpus.dimer_start(st)
    means = []
    for seed in st.get("device_seeds", (97531, 1, 2, 3, 4, 5, 6, 7)):
        pk = ra.GpuPacking(poly_shape(st["beads"]), box, pos, rot, seed=seed)
        pk.set_environment(st["temperature"], st["pressure"], st["trans_step"], None,
                           box_move=("log", st["box_step"]))
        pk.run_cycles(st["therm"], adapt_steps=True)
        phi = []
        for _ in range(st["avg"] // st["every"]):
            pk.run_cycles(st["every"])
            phi.append(pk.packing_fraction())
        means.append(float(np.mean(phi)))
        print(f"device run {len(means)} (seed {seed}): phi {means[-1]:.5f}", flush=True)
        b, p, r = pk.download()
        assert poly_overlaps(st["beads"], b, p, r) == 0
        pk.close()
    ref = np.array(st["phi_runs"])
    err = np.hypot(np.std(means, ddof=1) / np.sqrt(len(means)), ref.std(ddof=1) / np.sqrt(len(ref)))
    print(f"gpu phi {np.mean(means)} (runs {np.round(means, 5)}), reference {ref.mean()} (runs {np.round(ref, 5)}), "
          f"error of the difference {err}", flush=True)
    assert abs(np.mean(means) - ref.mean()) <= 2 * err


def random_chains(n_side=(8, 8, 2), spacing=(3.0, 3.0, 11.5), seed=12):
    box, pos, _ = chain_start(n_side, spacing)
    rng = np.random.Generator(np.random.PCG64(seed))
    rot = corpus.random_rotations(rng, len(pos))
    return box, pos, np.ascontiguousarray(rot)


def test_polysphere_counting_mode_and_overlap_relaxation():
    """Overlap counting for polyspheres (Packing::toggleOverlapCounting, countParticleOverlaps over interaction
    centres): the cached count of overlapping bead pairs equals the oracle's full recount after sweeps and box moves,
    and Simulation::relaxOverlaps from randomly oriented chains ends with no overlap."""
    box, pos, rot = random_chains()
    start = poly_overlaps("kinked11", box, pos, rot)
    assert start > 0
    pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=13)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.05))
    pk.set_overlap_counting(True)
    assert pk.cached_overlaps() == start
    for _ in range(3):
        pk.run_cycles(5)
        b, p, r = pk.download()
        assert pk.cached_overlaps() == poly_overlaps("kinked11", b, p, r)
    pk.set_overlap_counting(False)
    cycles, left = pk.relax_overlaps(20000)
    print(f"polysphere relaxation: {start} overlapping bead pairs removed in {cycles} cycles", flush=True)
    assert left == 0
    b, p, r = pk.download()
    assert poly_overlaps("kinked11", b, p, r) == 0
    pk.run_cycles(10)
    pk.close()
