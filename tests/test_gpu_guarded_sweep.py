"""The guarded FP32 stencil scan inside the sweep (RP_PREC_GUARDED_FP32): the same decisions as the FP64 scan, hence
bit-identical same-seed trajectories on every state; in verify mode every pair is also evaluated in FP64, no class
differs, and the guard-band hit rate and the largest FP32 error are reported."""
import numpy as np
import pytest

import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
HS = ra.Shape("sphere", radius=0.5)


def run(shape, state, precision, pressure, move, cycles, seed=5):
    box, pos, rot = state
    pk = ra.GpuPacking(shape, box, pos, rot, seed=seed, precision=precision)
    pk.set_environment(1.0, pressure, 0.15, None, box_move=move)
    c = pk.run_cycles(cycles, adapt_steps=True)
    out = pk.download(), c, pk.guard_stats()
    pk.close()
    return out


SPHERE_OFF = ("switched off: the eight-wavefront lean and guarded kernels faulted the GPU on their first launch on "
              "spheres (sparse stencils) and the cause is not found; spheres keep the general FP64 kernel")


@pytest.mark.parametrize("kind", [pytest.param("sphere", marks=pytest.mark.skip(reason=SPHERE_OFF)),
                                  "spherocylinder"])
def test_guarded_trajectories_equal_fp64(kind, monkeypatch):
    """Spheres would run the guarded scan only on the eight-wavefront launch, forced for their sparse stencils."""
    if kind == "sphere":
        monkeypatch.setenv("RAMPACK_SWEEP_WAVES", "8")
        shape, state, p, move = HS, lattice.hard_sphere_lattice(20, 0.45), 5.0, ("log", 0.001)
    else:
        shape, state, p, move = SC, lattice.spherocylinder_smectic(20000, 0.45), 2.0, ("delta_triclinic", 0.002)
    (b0, p0, r0), c0, _ = run(shape, state, ra.PREC_FP64, p, move, 200)
    (b1, p1, r1), c1, st = run(shape, state, ra.PREC_GUARDED_FP32, p, move, 200)
    print(f"{kind}: {c1['moves_accepted']}/{c1['moves_attempted']} accepted, FP32 pairs {st['pairs']}, "
          f"guard band {st['fp64_fallback']} ({st['fp64_fallback'] / max(st['pairs'], 1):.2e})", flush=True)
    assert st["pairs"] > 0
    assert c0["moves_accepted"] == c1["moves_accepted"]
    assert np.array_equal(b0, b1) and np.array_equal(p0, p1) and np.array_equal(r0, r1)


def test_guarded_verify_mode_reports_errors_and_no_mismatch():
    state = lattice.spherocylinder_smectic(100000, 0.45)
    (_, _, _), c, st = run(SC, state, ra.PREC_GUARDED_FP32_VERIFY, 2.0, ("delta_triclinic", 0.002), 50)
    print(f"verify: pairs {st['pairs']}, guard band {st['fp64_fallback']} "
          f"({st['fp64_fallback'] / st['pairs']:.2e}), max |FP32 - FP64| d^2 {st['max_err_centre']:.3e}, "
          f"mismatches {st['mismatches']}", flush=True)
    assert st["pairs"] > 10 ** 7
    assert st["mismatches"] == 0
    assert 0 < st["max_err_centre"] < 1e-3
