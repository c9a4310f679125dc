"""GPU checkerboard sweep: invariants (no overlaps, N trial moves per cycle, determinism), error paths, and the
reference's statistical validation cases (SimulationTest.cpp:49-101) against Carnahan-Starling."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
HS = ra.Shape("sphere", radius=0.5)


def overlaps_by_oracle(shape, box, pos, rot):
    kind = 0 if shape.kind == "sphere" else 1
    count, _ = oracle_lib.oracle().snapshot_stats(kind, shape.radius, shape.length, box, pos, rot)
    return count


def rod_start(n_side=(18, 18, 4)):
    return lattice.spherocylinder_lattice(n_side=n_side, spacing=(1.45, 1.45, 6.3))


def test_sweep_keeps_packing_overlap_free_and_moves_every_particle():
    box, pos, rot = rod_start()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=11)
    pk.set_environment(1.0, 2.0, 0.3, None, box_move=("delta_triclinic", 0.02))
    c = pk.run_cycles(30, adapt_steps=True)
    print(c, flush=True)
    # dense cells make round(N / cells) moves each (the reference's floor(N / domains) per domain)
    assert abs(c["moves_attempted"] - 30 * len(pos)) <= 0.1 * 30 * len(pos)
    assert c["unstaged_blocks"] == 0
    assert 0 < c["moves_accepted"] < c["moves_attempted"]
    b, p, r = pk.download()
    assert overlaps_by_oracle(SC, b, p, r) == 0
    assert pk.count_total_overlaps() == 0
    # positions stay in the box, orientations stay orthonormal
    rel = np.linalg.solve(b, p.T).T
    assert np.all(rel >= -1e-12) and np.all(rel < 1 + 1e-12)
    o = r.reshape(-1, 3, 3)
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", o, o), np.broadcast_to(np.eye(3), o.shape), atol=1e-10)


def test_two_cell_grid_reads_stencil_from_hbm():
    """A box two cells wide across the rods (21 x 21 x 25.6: tall enough for the reference's grid, which needs four
    interaction ranges along some direction): the 27-cell stencil repeats cells under different images, and the
    dense cells hold more rods than the own-cell table, so the blocks read candidates from HBM; the sweep must stay
    overlap-free there too."""
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(20, 20, 4), spacing=(1.05, 1.05, 6.4))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=2)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.02))
    c = pk.run_cycles(20)
    assert c["unstaged_blocks"] > 0
    assert c["moves_attempted"] == 20 * len(pos)
    b, p, r = pk.download()
    assert overlaps_by_oracle(SC, b, p, r) == 0


def test_sweep_moves_particles_across_cells_and_box_changes():
    box, pos, rot = rod_start()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=5)
    pk.set_environment(1.0, 1.5, 0.4, None, box_move=("delta_triclinic", 0.002))
    c = pk.run_cycles(150, adapt_steps=True)
    assert c["box_accepted"] > 0
    b, p, r = pk.download()
    assert not np.allclose(b, box)
    disp = p - pos
    assert np.max(np.abs(disp)) > 1.0


def test_same_seed_gives_identical_trajectories_and_other_seed_differs():
    box, pos, rot = rod_start()
    out = []
    for seed in (3, 3, 4):
        pk = ra.GpuPacking(SC, box, pos, rot, seed=seed)
        pk.set_environment(1.0, 2.0, 0.3, None, box_move=("delta_triclinic", 0.02))
        pk.run_cycles(15)
        out.append(pk.download())
        pk.close()
    assert all(np.array_equal(a, b) for a, b in zip(out[0], out[1]))
    assert not np.array_equal(out[0][1], out[2][1])


def test_hard_spheres_overlap_free_log_moves():
    box, pos, rot = lattice.hard_sphere_lattice(10, 0.30)
    pk = ra.GpuPacking(HS, box, pos, rot, seed=1234)
    pk.set_environment(1.0, 5.0, 0.1, None, box_move=("log", 0.01))
    c = pk.run_cycles(200, adapt_steps=True)
    assert c["moves_attempted"] == 200 * 1000
    b, p, r = pk.download()
    assert overlaps_by_oracle(HS, b, p, r) == 0


def test_overlapping_start_is_a_validation_error():
    box = np.diag([20.0, 20.0, 20.0])
    pos, rot = corpus.random_snapshot(300, 1, box)
    pk = ra.GpuPacking(SC, box, pos, rot)
    pk.set_environment(1.0, 1.0, 0.1)
    with pytest.raises(ra.ValidationError):
        pk.run_cycles(1)


def test_particle_outside_box_is_a_precondition_error():
    box, pos, rot = rod_start()
    pos = pos.copy()
    pos[3, 0] = -0.1
    with pytest.raises(ra.PreconditionError):
        ra.GpuPacking(SC, box, pos, rot)


def test_environment_preconditions():
    box, pos, rot = rod_start()
    pk = ra.GpuPacking(SC, box, pos, rot)
    with pytest.raises(ra.PreconditionError):
        pk.run_cycles(1)
    with pytest.raises(ra.PreconditionError):
        pk.set_environment(1.0, 1.0, -0.1)
    with pytest.raises(ra.PreconditionError):
        pk.set_environment(1.0, 1.0, 0.1, box_move=("delta_triclinic", 0.0))


def block_mean_error(samples):
    """Mean and its block-averaged standard error (the samples of one run are correlated)."""
    return ra.mean_and_error(samples, method="blocks")


def run_density(shape, box, pos, rot, temperature, pressure, trans, box_move, therm, avg, every, seed):
    pk = ra.GpuPacking(shape, box, pos, rot, seed=seed)
    pk.set_environment(temperature, pressure, trans, None, box_move=box_move)
    pk.run_cycles(therm, adapt_steps=True)
    samples = []
    for _ in range(avg // every):
        pk.run_cycles(every)
        b, _, _ = pk.download()
        samples.append(len(pos) / abs(np.linalg.det(b)))
    return block_mean_error(samples)


def test_dilute_hard_sphere_gas_carnahan_starling():
    """SimulationTest.cpp:49-75: 50 spheres of radius 0.05, T = 10, p = 1, delta_v scaling; CS density 0.0999791."""
    dims = (5000.0 ** (1 / 3),) * 3
    box, pos, rot = lattice.orthorhombic_arrange(50, dims)
    rho, err = run_density(ra.Shape("sphere", radius=0.05), box, pos, rot, 10.0, 1.0, 1.0, ("delta_v", 1.0),
                           5000, 10000, 100, 1234)
    print(f"dilute HS: {rho} +- {err} (CS 0.0999791)", flush=True)
    assert abs(rho - 0.0999791) <= 2 * err
    assert err / rho < 0.03


def test_degenerate_hard_sphere_gas_carnahan_starling():
    """SimulationTest.cpp:77-101: 50 spheres of radius 0.5, T = 1, p = 1; CS density 0.398574."""
    dims = (200.0 ** (1 / 3),) * 3
    box, pos, rot = lattice.orthorhombic_arrange(50, dims)
    rho, err = run_density(HS, box, pos, rot, 1.0, 1.0, 1.0, ("delta_v", 1.0), 5000, 10000, 100, 1234)
    print(f"degenerate HS: {rho} +- {err} (CS 0.398574)", flush=True)
    assert abs(rho - 0.398574) <= 2 * err
    assert err / rho < 0.03


@pytest.mark.parametrize("waves,pipe", [(1, 0), (2, 0), (4, 0), (2, 1), (4, 1), (2, 2), (4, 2), (8, 1), (8, 2)])
@pytest.mark.parametrize("kind", ["sphere", "spherocylinder"])
def test_every_launch_shape_is_overlap_free(kind, waves, pipe, monkeypatch):
    """Each workgroup size (1, 2, 4 or 8 wavefronts per active cell; plain, planner/tester split or grouped moves) for
    both shapes keeps the packing overlap-free."""
    monkeypatch.setenv("RAMPACK_SWEEP_WAVES", str(waves))
    monkeypatch.setenv("RAMPACK_SWEEP_PIPELINE", str(pipe))
    if kind == "sphere":
        shape = HS
        box, pos, rot = lattice.hard_sphere_lattice(10, 0.45)
        move = ("log", 0.001)
    else:
        shape = SC
        box, pos, rot = rod_start()
        move = ("delta_triclinic", 0.002)
    pk = ra.GpuPacking(shape, box, pos, rot, seed=100 + waves + 10 * pipe)
    pk.set_environment(1.0, 5.0 if kind == "sphere" else 2.0, 0.15, None, box_move=move)
    c = pk.run_cycles(100, adapt_steps=True)
    assert abs(c["moves_attempted"] - 100 * len(pos)) <= 0.1 * 100 * len(pos)
    b, p, r = pk.download()
    assert overlaps_by_oracle(shape, b, p, r) == 0


@pytest.mark.parametrize("period", [2, 3])
def test_both_colourings_are_overlap_free(period, monkeypatch):
    """Period-2 (cells >= range, 8 colours, 27-cell stencil) and period-3 (cells >= range/2, 27 colours, 125-cell
    stencil) checkerboards both keep the packing overlap-free, including the box-move scans on their grids."""
    monkeypatch.setenv("RAMPACK_COLOURING", str(period))
    box, pos, rot = rod_start()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=40 + period)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
    pk.set_timing(True)
    c = pk.run_cycles(120, adapt_steps=True)
    t = pk.timing()
    assert t["sweep_period"] == period
    assert t["sweep_launches"] == 120 * period ** 3
    assert c["box_accepted"] > 0
    b, p, r = pk.download()
    assert overlaps_by_oracle(SC, b, p, r) == 0


def test_sampled_timing_times_one_sweep_per_cycle():
    """Sampled timing: one sweep launch per cycle carries events, no other launch does, and the trajectory is the
    untimed one."""
    box, pos, rot = rod_start()
    runs = []
    for mode in ("sampled", False):
        pk = ra.GpuPacking(SC, box, pos, rot, seed=77)
        pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
        pk.set_timing(mode)
        pk.run_cycles(40)
        t = pk.timing()
        if mode:
            assert t["sweep_launches"] == 40 and t["sweep_ms"] > 0
            assert t["rebuild_launches"] == 0 and t["scan_launches"] == 0 and t["proposal_launches"] == 0
        runs.append(pk.download())
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("kind", ["sphere", "spherocylinder"])
def test_launch_shapes_give_identical_trajectories(kind, monkeypatch):
    """Proposals are drawn per (cycle, cell, move) and every decision is the same FP64 evaluation, so the
    single-wavefront, multi-wavefront, pipelined (planner + testers) and grouped (several moves tested at once,
    decided in order) sweeps produce bit-identical packings."""
    if kind == "sphere":
        shape = HS
        box, pos, rot = lattice.hard_sphere_lattice(10, 0.45)
        move = ("log", 0.001)
    else:
        shape = SC
        box, pos, rot = rod_start()
        move = ("delta_triclinic", 0.002)
    results = {}
    for waves, pipe in ((1, 0), (2, 0), (4, 0), (2, 1), (4, 1), (2, 2), (4, 2), (8, 0), (8, 1), (8, 2)):
        monkeypatch.setenv("RAMPACK_SWEEP_WAVES", str(waves))
        monkeypatch.setenv("RAMPACK_SWEEP_PIPELINE", str(pipe))
        pk = ra.GpuPacking(shape, box, pos, rot, seed=77)
        pk.set_environment(1.0, 5.0 if kind == "sphere" else 2.0, 0.15, None, box_move=move)
        c = pk.run_cycles(40, adapt_steps=True)
        results[(waves, pipe)] = (pk.download(), c["moves_accepted"])
        pk.close()
    (b0, p0, r0), acc0 = results[(1, 0)]
    assert acc0 > 0
    for key, ((b, p, r), acc) in results.items():
        assert acc == acc0, key
        assert np.array_equal(b, b0) and np.array_equal(p, p0) and np.array_equal(r, r0), key


def test_one_cell_axis_sweep_stays_overlap_free():
    """A box less than two interaction ranges high on one axis: the grid has one cell across it, the cell is its
    own neighbour under the images -1, 0, +1, and the sweep reads the stencil from HBM."""
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(3, 6, 3), spacing=(2.3, 2.2, 8.3))
    assert box[0, 0] < 2 * SC.range
    pk = ra.GpuPacking(SC, box, pos, rot, seed=8)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.01))
    c = pk.run_cycles(200, adapt_steps=True)
    assert c["unstaged_blocks"] > 0
    assert c["box_too_small"] == 0
    assert 0 < c["moves_accepted"] < c["moves_attempted"]
    b, p, r = pk.download()
    assert overlaps_by_oracle(SC, b, p, r) == 0


def test_degenerate_spherocylinder_gas_reference_value():
    """SimulationTest.cpp:103-148: 50 spherocylinders L = 3, r = 0.5 from a 5 x 5 x 2 arrangement (spacing 2, 2, 5)
    in a 10^3 box, T = 1, p = 0.5, rototranslation(0.5, 1), delta_v(10); the reference's long-run density is
    0.09454087.  The equilibrium box (about 8.1 high for a range of 4) moves between one- and two-cell grids.
    The volume of 50 particles decorrelates slowly: over 10^4 cycles the reference's own runs scatter by 5 %
    between seeds (0.0929, 0.0973), far beyond the naive standard error; over 6 x 10^4 cycles sampled every 100,
    three reference seeds give 0.0946, 0.0948, 0.0945 with block errors of 5-6 x 10^-4.  The same protocol is used
    here, with the block-averaged error (block_error), at 2 standard errors."""
    box, pos, rot = lattice.orthorhombic_arrange(50, (10, 10, 10), (5, 5, 2), (2, 2, 5))
    shape = ra.Shape("spherocylinder", radius=0.5, length=3.0)
    pk = ra.GpuPacking(shape, box, pos, rot, seed=1234)
    pk.set_environment(1.0, 0.5, 0.5, 1.0, box_move=("delta_v", 10.0))
    pk.run_cycles(5000, adapt_steps=True)
    rho = []
    for _ in range(600):
        pk.run_cycles(100)
        rho.append(pk.observables()["number_density"])
    c = pk.run_cycles(0)
    m = float(np.mean(rho))
    block_err = ra.block_error(rho)
    print(f"spherocylinder gas: {m} +- {block_err} (reference 0.09454087); box moves too small "
          f"{c['box_too_small']}", flush=True)
    assert c["box_too_small"] == 0
    assert abs(m - 0.09454087) <= 2 * block_err
    assert block_err / m < 0.01


@pytest.mark.parametrize("shape_name", ["hard", "wca"])
def test_fix_orientations_matches_oracle(shape_name):
    """Packing::renormalizeOrientations: orientations perturbed by 1e-7 are replaced by the oracle's fix_rotation
    bit for bit; in a packing with room to spare no fix is rejected.  A WCA shape has no hard part: every fix
    applies."""
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(10, 10, 3), spacing=(1.6, 1.6, 6.4))
    rng = np.random.Generator(np.random.PCG64(12))
    rot = np.ascontiguousarray(rot + (rng.random(rot.shape) - 0.5) * 2e-7)
    shape = SC if shape_name == "hard" else ra.Shape("spherocylinder", 0.5, 5.0, soft="wca_segment")
    pk = ra.GpuPacking(shape, box, pos, rot)
    assert pk.fix_orientations() == 0
    _, p, r = pk.download()
    expected = np.ascontiguousarray(rot.copy())
    oracle_lib.oracle().lib.orc_fix_rotation(len(expected), oracle_lib._d(expected))
    assert np.array_equal(p, pos)
    assert np.array_equal(r, expected)
    pk.close()


def random_rod_lattice(n_side, spacing, seed):
    """Rods on a simple-cubic lattice with uniformly random orientations: overlapping when spacing < L + 2r."""
    idx = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(float)
    pos = (idx + 0.5) * spacing
    rng = np.random.Generator(np.random.PCG64(seed))
    rot = corpus.random_rotations(rng, len(pos))
    return np.diag([n_side * spacing] * 3), np.ascontiguousarray(pos), np.ascontiguousarray(rot)


def test_cached_overlap_count_stays_exact_in_counting_mode():
    """Overlap-counting mode (Packing::toggleOverlapCounting): after particle and box moves the cached count equals
    a full recount by the oracle, exactly."""
    box, pos, rot = random_rod_lattice(8, 3.0, 4)
    start, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, box, pos, rot)
    assert start > 0
    pk = ra.GpuPacking(SC, box, pos, rot, seed=6)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.01))
    pk.set_overlap_counting(True)
    assert pk.cached_overlaps() == start
    for _ in range(4):
        c = pk.run_cycles(5)
        b, p, r = pk.download()
        full, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r)
        assert pk.cached_overlaps() == full
    assert c["box_attempted"] > 0 and c["moves_accepted"] > 0
    pk.set_overlap_counting(False)
    with pytest.raises(ra.PreconditionError):
        pk.cached_overlaps()


def test_overlap_relaxation_reaches_a_legal_packing():
    """Simulation::relaxOverlaps from randomly oriented rods on a lattice (the 20 x 20 x 20 random-orientation
    start at phi = 0.25 cut to 1000 rods at the same spacing): the relaxation ends with no overlap, confirmed by the oracle, and ordinary integration accepts
    the result."""
    box, pos, rot = random_rod_lattice(10, 52.22 / 20, 2024)
    before, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, box, pos, rot)
    assert before > 100
    pk = ra.GpuPacking(SC, box, pos, rot, seed=2024)
    pk.set_environment(1.0, 1.0, 0.5, None, box_move=("delta_triclinic", 0.01))
    cycles, left = pk.relax_overlaps(20000)
    print(f"relaxation: {before} overlaps removed in {cycles} cycles", flush=True)
    assert left == 0
    b, p, r = pk.download()
    assert overlaps_by_oracle(SC, b, p, r) == 0
    pk.run_cycles(10)
    pk.close()


def test_anisotropic_linear_and_one_side_box_moves():
    """AnisotropicVolumeScaler (AnisotropicVolumeScaler.cpp:13-31) through TriclinicAdapter: 'independent xy z'
    scales the x and y sides by one factor (their lengths stay equal bit for bit) and z on its own, 'anisotropic x y
    z' scales the three sides apart; a diagonal box stays diagonal (sides are only scaled), and the isotropic linear
    sampler keeps a cube a cube.  TriclinicDeltaScaler with scaleTogether false shears the box.  Every run stays
    overlap-free."""
    box, pos, rot = lattice.hard_sphere_lattice(8, 0.30)
    runs = [("log", "independent xy z"), ("log", "anisotropic x y z"), ("linear", "isotropic"),
            ("delta_triclinic", "one_side")]
    for i, (kind, spec) in enumerate(runs):
        pk = ra.GpuPacking(HS, box, pos, rot, seed=60 + i)
        # steps of about 0.1 % of the box: a 1 % compression of this fluid almost always creates a contact
        step = {"log": 0.001, "linear": 0.01, "delta_triclinic": 0.01}[kind]
        pk.set_environment(1.0, 5.0, 0.1, None, box_move=(kind, step, spec))
        c = pk.run_cycles(300, adapt_steps=True)
        b, p, r = pk.download()
        print(kind, spec, c["box_accepted"], np.diag(b), flush=True)
        assert c["box_accepted"] > 10
        assert overlaps_by_oracle(HS, b, p, r) == 0
        off = b - np.diag(np.diag(b))
        if kind == "delta_triclinic":
            assert np.abs(off).max() > 0
            continue
        assert not off.any()
        d = np.diag(b)
        if spec == "independent xy z":
            assert d[0] == d[1] and d[2] != d[0]
        elif spec == "anisotropic x y z":
            assert d[0] != d[1] and d[1] != d[2] and d[0] != d[2]
        else:
            assert d[0] == d[1] == d[2]
        pk.close()
    pk = ra.GpuPacking(HS, box, pos, rot)
    with pytest.raises(ValueError):
        pk.set_environment(1.0, 5.0, 0.1, None, box_move=("log", 0.01, "anisotropic xy"))


def test_spherocylinder_liquid_npt_matches_reference_runs(golden):
    """NpT hard spherocylinders L/D = 5 at p* = 1 (an isotropic liquid near packing fraction 0.31, 300 rods from an
    aligned lattice, delta_triclinic box moves) against 24 runs of the reference build (its Packing, samplers and
    scaler under the restated cycle loop): the mean packing fraction and nematic P2 over 16 independent device runs
    agree with the reference's means within 2 standard errors of the difference, each side's error taken from the
    spread between its runs (the run means scatter by about 0.002 in packing fraction: the box is small enough for
    the reference to run it without its neighbour grid, and so does the device, through every pair)."""
    st = next((s for s in golden.get("statistics", []) if s["name"] == "sc_npt_p1"), None)
    if st is None:
        pytest.skip("no reference statistics in the fixture")
    box, pos, rot = lattice.spherocylinder_lattice(n_side=tuple(st["n_side"]), spacing=tuple(st["spacing"]))
    phis, p2s = [], []
    for seed in range(1, 17):
        pk = ra.GpuPacking(SC, box, pos, rot, seed=seed)
        pk.set_environment(st["temperature"], st["pressure"], st["trans_step"], st["rot_step"],
                           box_move=("delta_triclinic", st["box_step"]))
        pk.run_cycles(st["therm"], adapt_steps=True)
        phi, p2 = [], []
        for _ in range(st["avg"] // st["every"]):
            pk.run_cycles(st["every"])
            o = pk.observables()
            phi.append(o["packing_fraction"])
            p2.append(o["nematic_p2"])
        phis.append(float(np.mean(phi)))
        p2s.append(float(np.mean(p2)))
        print(f"seed {seed}: phi {phis[-1]:.5f}, P2 {p2s[-1]:.4f}", flush=True)
        b, p, r = pk.download()
        assert overlaps_by_oracle(SC, b, p, r) == 0
        pk.close()
    for name, dev, ref in (("phi", phis, st["phi_runs"]), ("P2", p2s, st["p2_runs"])):
        dev, ref = np.array(dev), np.array(ref)
        err = np.hypot(dev.std(ddof=1) / np.sqrt(len(dev)), ref.std(ddof=1) / np.sqrt(len(ref)))
        print(f"{name}: device {dev.mean()} (runs {np.round(dev, 4)}), reference {ref.mean()} "
              f"(runs {np.round(ref, 4)}), error of the difference {err}", flush=True)
        assert abs(dev.mean() - ref.mean()) <= 2 * err
