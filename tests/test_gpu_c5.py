"""The large dense system: N = 1e6 hard spherocylinders L/D = 5 started from the AA-stacked smectic at packing fraction
0.50, triclinic NpT at p* = 2.5, and its soft variant in which every pair also carries a WCA repulsion (epsilon 1,
sigma 1, cut at 2^(1/6) sigma) on the segment-segment distance with Metropolis acceptance at kT = 1 (the reference's
CompoundInteraction of the hard spherocylinder and the WCA segment interaction).

* Hard: the plain FP64 sweep and the guarded FP32 sweep run the same seeded trajectory (same decisions, so the same
  end state bit for bit and the same counters); the end state is recounted by the oracle: no overlap.
* Soft variant: after the cycles the oracle recounts the state: no hard-core overlap, and the device's FP64 total
  energy equals the oracle's within 1e-10 relative."""
import numpy as np
import pytest

import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
C_ROD = ra.Shape("spherocylinder", radius=0.5, length=5.0, soft="wca_segment", epsilon=1.0, sigma=1.0,
                 hard_core=True)
HARD_CORE = 0x100
N = 1_000_000


@pytest.fixture(scope="module")
def smectic():
    return lattice.spherocylinder_smectic(N, 0.50)


def test_c5_hard_fp64_and_guarded_fp32_same_trajectory(smectic):
    box, pos, rot = smectic
    runs = []
    for precision in (ra.PREC_FP64, ra.PREC_GUARDED_FP32):
        pk = ra.GpuPacking(SC, box, pos, rot, seed=2025, precision=precision)
        pk.set_environment(1.0, 2.5, 0.1, 0.05, box_move=("delta_triclinic", 0.001))
        c = pk.run_cycles(20, adapt_steps=True)
        runs.append((pk.download(), c, pk.guard_stats(), pk.timing()))
        pk.close()
    (s0, c0, _, t0), (s1, c1, g1, _) = runs
    print(f"N = 1e6 smectic at phi 0.50, p* = 2.5: {c0['moves_accepted']}/{c0['moves_attempted']} moves accepted, "
          f"box {c0['box_accepted']}/{c0['box_attempted']}, sweep waves {t0['sweep_waves']}, unstaged blocks "
          f"{c0['unstaged_blocks']}; guarded: {g1['pairs']} FP32 pairs, {g1['fp64_fallback']} in the guard band",
          flush=True)
    assert abs(c0["moves_attempted"] - 20 * N) <= 0.1 * 20 * N
    assert 0 < c0["moves_accepted"] < c0["moves_attempted"]
    # the launches split cells between LDS and HBM stencils differently (the guarded scan's entries are larger), which
    # changes only the launch-shape counters
    shape = ("pair_tests", "unstaged_blocks")
    assert {k: v for k, v in c0.items() if k not in shape} == {k: v for k, v in c1.items() if k not in shape}
    assert all(np.array_equal(a, b) for a, b in zip(s0, s1))
    assert g1["pairs"] > 0
    b, p, r = s0
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r)[0] == 0


def test_c5_compound_wca_variant(smectic):
    box, pos, rot = smectic
    pk = ra.GpuPacking(C_ROD, box, pos, rot, seed=7)
    pk.set_environment(1.0, 2.5, 0.05, 0.02, box_move=("delta_triclinic", 0.0005))
    c = pk.run_cycles(5, adapt_steps=True)
    e = pk.total_energy()
    b, p, r = pk.download()
    pk.close()
    n, e_orc = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r, soft=2 | HARD_CORE, eps=1.0, sigma=1.0)
    print(f"N = 1e6 compound: {c['moves_accepted']}/{c['moves_attempted']} moves accepted, box "
          f"{c['box_accepted']}/{c['box_attempted']}; energy {e} (oracle {e_orc}); hard-core overlaps {n}", flush=True)
    assert 0 < c["moves_accepted"] < c["moves_attempted"]
    assert n == 0
    assert e > 0
    assert e == pytest.approx(e_orc, rel=1e-10)
