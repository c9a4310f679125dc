"""Packing::tryScaling / revertScaling on the device-resident state (rp_try_scaling / rp_revert_scaling,
Packing.cpp:172-208, 363-371) against the oracle: the scaled positions are newBox * (oldBox^-1 * pos) bit for bit,
the result is +inf exactly when the oracle finds an overlap in the scaled packing (none looked for when the box only
grew), the soft energy change is the oracle's within the rounding of the sums, the counting-mode result follows
the oracle's full counts, and revertScaling restores the state bit for bit."""
import ctypes

import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice
from test_gpu_sweep import random_rod_lattice

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
WCA_SPHERE = ra.Shape("sphere", radius=0.5, soft="wca", epsilon=1.0, sigma=1.0)


def rescaled(old_box, new_box, pos):
    out = np.empty_like(pos)
    d = ctypes.POINTER(ctypes.c_double)
    ob = np.ascontiguousarray(old_box, dtype=np.float64).reshape(9)
    nb = np.ascontiguousarray(new_box, dtype=np.float64).reshape(9)
    p = np.ascontiguousarray(pos, dtype=np.float64)
    oracle_lib.oracle().lib.orc_rescale(ob.ctypes.data_as(d), nb.ctypes.data_as(d), len(p), p.ctypes.data_as(d),
                                        out.ctypes.data_as(d))
    return out


def same_state(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_try_scaling_hard_rods_against_oracle():
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(12, 12, 3), spacing=(1.6, 1.6, 6.4))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=3)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.01))
    pk.run_cycles(10)
    start = pk.download()
    b0, p0, r0 = start
    shear = np.array([[0.0, 0.02, 0.0], [0.0, 0.0, -0.01], [0.0, 0.0, 0.0]]) * b0[0, 0]
    for name, new_box in [("grow", b0 * 1.03), ("nudge", b0 * 0.9995), ("shrink", b0 * 0.97),
                          ("crush", b0 * 0.8), ("shear", b0 + shear)]:
        dE = pk.try_scaling(new_box)
        b, p, r = pk.download()
        assert np.array_equal(b, new_box), name
        assert np.array_equal(p, rescaled(b0, new_box, p0)), name
        assert np.array_equal(r, r0), name
        overlaps, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r, early_exit=True)
        print(f"{name}: dE {dE}, oracle overlaps {overlaps}", flush=True)
        if name == "grow":
            assert dE == 0.0  # a convex shape in a box that only grew cannot overlap: no scan
        else:
            assert dE == (np.inf if overlaps else 0.0), name
        if name == "crush":
            assert dE == np.inf
        pk.revert_scaling()
        assert same_state(pk.download(), start), name
    # the packing keeps working after a revert
    pk.run_cycles(10)
    b, p, r = pk.download()
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r)[0] == 0
    pk.close()


def test_revert_without_try_is_a_precondition_error():
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(8, 8, 2), spacing=(1.6, 1.6, 6.4))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=4)
    with pytest.raises(ra.PreconditionError):
        pk.revert_scaling()
    with pytest.raises(ra.PreconditionError):
        pk.try_scaling(np.zeros((3, 3)))
    pk.close()


def test_try_scaling_soft_energy_change():
    box, pos, rot = corpus.jittered_lattice(10, 1.08, 0.1, 45)
    pk = ra.GpuPacking(WCA_SPHERE, box, pos, rot, seed=5)
    start = pk.download()
    e0 = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=1.0)[1]
    for f in (0.99, 0.97, 1.01):
        dE = pk.try_scaling(box * f)
        b, p, r = pk.download()
        assert np.array_equal(p, rescaled(box, box * f, pos))
        e1 = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, b, p, r, soft=1, eps=1.0, sigma=1.0)[1]
        print(f"scale {f}: dE {dE}, oracle {e1 - e0}", flush=True)
        assert abs(dE - (e1 - e0)) <= 1e-10 * max(abs(e0), abs(e1), 1.0)
        assert abs(pk.total_energy() - e1) <= 1e-10 * max(abs(e1), 1.0)
        pk.revert_scaling()
        assert same_state(pk.download(), start)


def test_try_scaling_in_overlap_counting_mode():
    box, pos, rot = random_rod_lattice(8, 3.0, 4)
    start, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, box, pos, rot)
    pk = ra.GpuPacking(SC, box, pos, rot, seed=6)
    pk.set_environment(1.0, 1.0, 0.3, None, box_move=("delta_triclinic", 0.01))
    pk.set_overlap_counting(True)
    assert pk.cached_overlaps() == start
    for f in (1.02, 0.98, 1.0):
        dE = pk.try_scaling(box * f)
        b, p, r = pk.download()
        after, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r)
        print(f"scale {f}: dE {dE}, overlaps {start} -> {after}", flush=True)
        assert dE == (-np.inf if after < start else (np.inf if after > start else 0.0))
        assert pk.cached_overlaps() == after
        pk.revert_scaling()
        assert pk.cached_overlaps() == start
    pk.close()
