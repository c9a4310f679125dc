"""The library's kernels and host code on the CPU (tests/hostbuild): compiled by g++ against a host stand-in of the
HIP runtime, every workgroup's lanes run as fibers with wavefront operations and barriers emulated, and
AddressSanitizer checks every access against its allocation (device buffers and each launch's dynamic LDS have
exactly the requested size).

* Every launch shape's decision records replay through the oracle with no mismatch: the CPU build reaches the
  device's decisions, and the kernels' intra-wavefront ordering is explicit (a lane that runs ahead of the others of
  its wavefront, as lanes do here, sees nothing stale).
* Packings without a neighbour grid (the all-pairs sweep, staged in LDS and read in place) run with box moves:
  no access outside an allocation, no overlap, and their decision records replay with no mismatch.
* A grid with more cells than one scan workgroup takes (the tiled scan of the cell counts) rebuilds the records
  unchanged, and sweeps on it replay.
* Polysphere sweeps on sixteen and eight wavefronts per cell follow the four-wavefront trajectory.
* The launch-shape autotuner runs every candidate shape and the run ends where an untuned run ends.
* Replicas of the 32-replica equation-of-state scan (partly filled parallel lattices at packing fraction 0.2,
  p* from 0.5 to 4) run with box moves and step adaptation: no access outside an allocation, no overlap, and their
  decision records replay with no mismatch.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "hostbuild"))


@pytest.fixture(scope="module")
def host_lib():
    import host_build
    try:
        return host_build.build()
    except RuntimeError as e:
        pytest.fail(str(e))


def run_case(lib, *args, timeout=600):
    import host_build
    env = dict(os.environ, LD_PRELOAD=host_build.asan_runtime(), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    res = subprocess.run([sys.executable, os.path.join(HERE, "hostbuild", "run_case.py"), lib] + [str(a) for a in args],
                         capture_output=True, text=True, timeout=timeout, env=env)
    print(res.stdout[-3000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    assert "case passed" in res.stdout


@pytest.mark.parametrize("shape", ["1 0", "2 1", "4 1", "4 2", "8 1", "8 2"])
def test_launch_shapes_replay_on_cpu_build(host_lib, shape):
    run_case(host_lib, "replay", *shape.split())


@pytest.mark.parametrize("name,cycles", [("liquid", 5), ("slab", 2)])
def test_packings_without_grid_on_cpu_build(host_lib, name, cycles):
    run_case(host_lib, "nogrid", name, cycles)


def test_polysphere_sweep_on_sixteen_and_eight_wavefronts_on_cpu_build(host_lib):
    run_case(host_lib, "polywaves", 3)


def test_autotuner_keeps_the_trajectory_on_cpu_build(host_lib):
    run_case(host_lib, "autotune", 1)


def test_grid_larger_than_one_scan_workgroup_on_cpu_build(host_lib):
    run_case(host_lib, "biggrid", 3)


@pytest.mark.parametrize("replica", [13, 8, 1])
def test_scan_replicas_on_cpu_build(host_lib, replica):
    """Replica 13 (N = 1100, p* = 2.19: eight-wavefront grouped cells), 8 (N = 2055, p* = 1.40) and 1 (N = 12877,
    p* = 0.61: two-wavefront planner/tester cells, a fifth of the cells empty)."""
    run_case(host_lib, "replica", replica, 20 if replica != 1 else 10, timeout=900)
