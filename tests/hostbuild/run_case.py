"""One case on the CPU build of the library (run by tests/test_hostbuild.py in a process of its own, with the
AddressSanitizer runtime preloaded; any access outside an allocation aborts the process with a report).

    python tests/hostbuild/run_case.py LIB replay WAVES PIPE     # decision records of one launch shape vs the oracle
    python tests/hostbuild/run_case.py LIB replica I CYCLES      # replica I of the 32-replica scan, overlap checks
    python tests/hostbuild/run_case.py LIB nogrid NAME CYCLES    # a packing without a grid, overlap checks, replays
    python tests/hostbuild/run_case.py LIB biggrid CYCLES        # grids larger than one scan workgroup
    python tests/hostbuild/run_case.py LIB autotune CYCLES       # the launch-shape autotuner keeps the trajectory
    python tests/hostbuild/run_case.py LIB polywaves CYCLES      # polysphere sweeps on eight and four wavefronts
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "golden"))

import numpy as np  # noqa: E402

import rampack_amd as ra  # noqa: E402

ra.load_library(sys.argv[1])

import oracle_lib  # noqa: E402
from rampack_amd import lattice  # noqa: E402

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)


def replay(waves: str, pipe: str) -> None:
    from test_gpu_replay import check_record, rod_start
    os.environ["RAMPACK_SWEEP_WAVES"] = waves
    os.environ["RAMPACK_SWEEP_PIPELINE"] = pipe
    box, pos, rot = rod_start()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=31)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.01))
    pk.run_cycles(5, adapt_steps=True)
    for _ in range(2):
        check_record(SC, pk.record_sweep(), flags=not (waves == "8" and pipe != "2"))
    pk.close()


def replica(i: int, cycles: int) -> None:
    """The scan's replica i: N drawn uniformly from 1000..20000 (seed 7), p* = linspace(0.5, 4, 32)[i], parallel rods
    on a partly filled sc lattice at packing fraction 0.2; the state is checked for overlaps by the oracle every ten
    cycles, and the sweep's decisions are replayed through the oracle at the start and at the end."""
    from test_gpu_replay import check_record
    n = int(np.random.default_rng(7).integers(1000, 20001, 32)[i])
    p = float(np.linspace(0.5, 4.0, 32)[i])
    box, pos, rot = lattice.rod_replica_start(n, 0.2)
    pk = ra.GpuPacking(SC, box, pos, rot, seed=100 + i)
    pk.set_environment(1.0, p, 0.5, 0.2, box_move=("log", 0.002))
    check_record(SC, pk.record_sweep())
    done = 0
    while done < cycles:
        c = pk.run_cycles(min(10, cycles - done), adapt_steps=True)
        done += min(10, cycles - done)
        b, q, r = pk.download()
        overl = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, q, r)[0]
        print(f"replica {i} N={n} p*={p:.3f} cycle {c['cycles']}: moves {c['moves_accepted']}/{c['moves_attempted']}, "
              f"box {c['box_accepted']}/{c['box_attempted']}, unstaged blocks {c['unstaged_blocks']}, "
              f"phi {pk.packing_fraction():.4f}, overlaps {overl}", flush=True)
        assert overl == 0
    check_record(SC, pk.record_sweep())
    pk.close()


def nogrid(name: str, cycles: int) -> None:
    """A packing the reference runs without its neighbour grid: "liquid" (300 rods from the 10 x 10 x 3 lattice, p* = 1,
    triclinic box moves: the all-pairs sweep with the packing staged in LDS) or "slab" (1500 rods lying in two layers
    of a 97.5 x 60 x 5.5 box: too many for LDS, read in place).  The decision records replay through the oracle before
    and after the cycles, and the oracle finds no overlap."""
    from test_gpu_replay import check_record
    from test_gpu_nogrid import lying_rod_slab
    if name == "liquid":
        box, pos, rot = lattice.spherocylinder_lattice(n_side=(10, 10, 3), spacing=(1.6, 1.6, 6.5))
    else:
        box, pos, rot = lying_rod_slab()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=17)
    pk.set_environment(1.0, 1.0, 0.2, 0.2, box_move=("delta_triclinic", 0.01))
    rec = pk.record_sweep()
    assert rec["cells"] == [1, 1, 1]
    check_record(SC, rec)
    c = pk.run_cycles(cycles, adapt_steps=True)
    t = pk.timing()
    b, q, r = pk.download()
    overl = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, q, r)[0]
    print(f"{name}: N={len(pos)} moves {c['moves_accepted']}/{c['moves_attempted']}, box {c['box_accepted']}/"
          f"{c['box_attempted']}, sweep period {t['sweep_period']}, staged {t['sweep_cap']}, overlaps {overl}",
          flush=True)
    assert t["sweep_period"] == 1 and t["sweep_cap"] == (len(pos) if name == "liquid" else 0)
    assert overl == 0
    check_record(SC, pk.record_sweep())
    pk.close()


def big_grid(cycles: int) -> None:
    """A dilute packing whose grids have more cells than one workgroup scans (1728 spheres at packing fraction
    0.048 in a box 26.6 wide: 26^3 cells of the interaction range): the cell list is rebuilt through the tiled
    scan; the records come back unchanged, the overlap count is 0, and after a few sweep cycles the state is
    overlap-free and its last sweep replays."""
    from test_gpu_replay import check_record
    hs = ra.Shape("sphere", radius=0.5)
    box, pos, rot = lattice.hard_sphere_lattice(12, 0.048)
    pk = ra.GpuPacking(hs, box, pos, rot, seed=3)
    b, q, r = pk.download()
    assert np.array_equal(q, pos) and np.array_equal(r, rot)
    assert pk.count_total_overlaps() == 0
    pk.set_environment(1.0, 1.0, 0.3, 0.2, box_move=("log", 0.001))
    c = pk.run_cycles(cycles)
    t = pk.timing()
    b, q, r = pk.download()
    overl = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, b, q, r)[0]
    print(f"big grid: N={len(pos)} moves {c['moves_accepted']}/{c['moves_attempted']}, box {c['box_accepted']}/"
          f"{c['box_attempted']}, overlaps {overl}", flush=True)
    assert overl == 0 and c["moves_attempted"] == cycles * len(pos)
    check_record(hs, pk.record_sweep(), rejections=False)
    pk.close()


def poly_waves(cycles: int) -> None:
    """Kinked 11-bead chains (the polysphere sweep reading its stencil from HBM) with sixteen, eight and four
    wavefronts per active cell: the same trajectory, and no overlap by the oracle."""
    from test_gpu_poly import chain_start, poly_shape, poly_overlaps
    box, pos, rot = chain_start()
    states = []
    for w in ("16", "8", "4"):
        os.environ["RAMPACK_SWEEP_WAVES"] = w
        pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=5)
        pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.05))
        c = pk.run_cycles(cycles)
        assert pk.timing()["sweep_waves"] == int(w)
        states.append(pk.download())
        print(f"{w} wavefronts: moves {c['moves_accepted']}/{c['moves_attempted']}", flush=True)
        pk.close()
    del os.environ["RAMPACK_SWEEP_WAVES"]
    for other in states[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(states[0], other))
    b, q, r = states[0]
    assert poly_overlaps("kinked11", b, q, r) == 0


def autotune(cycles: int) -> None:
    """The autotuner's six launch shapes on rods: tuning over k cycles per shape, then two cycles, ends bit for bit
    where 6 k cycles and then two cycles without tuning end."""
    from test_gpu_replay import rod_start
    box, pos, rot = rod_start()
    states = []
    for tune in (True, False):
        pk = ra.GpuPacking(SC, box, pos, rot, seed=9)
        pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
        if tune:
            res = pk.autotune(cycles)
            assert len(res["candidates"]) == 6
            pk.run_cycles(2)
        else:
            pk.run_cycles(6 * cycles)
            pk.run_cycles(2)
        states.append(pk.download())
        pk.close()
    assert all(np.array_equal(a, b) for a, b in zip(*states))
    print("autotune: tuned and untuned runs end in the same state", flush=True)


if __name__ == "__main__":
    if sys.argv[2] == "replay":
        replay(sys.argv[3], sys.argv[4])
    elif sys.argv[2] == "nogrid":
        nogrid(sys.argv[3], int(sys.argv[4]))
    elif sys.argv[2] == "biggrid":
        big_grid(int(sys.argv[3]))
    elif sys.argv[2] == "autotune":
        autotune(int(sys.argv[3]))
    elif sys.argv[2] == "polywaves":
        poly_waves(int(sys.argv[3]))
    else:
        replica(int(sys.argv[3]), int(sys.argv[4]))
    print("case passed", flush=True)
