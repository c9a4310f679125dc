"""CompoundInteraction (src/core/interactions/CompoundInteraction.cpp:9-60) on the device: a hard core plus a soft
part.  Pair decisions are the hard part's and pair energies the soft part's; snapshot overlap counts and total
energies equal the reference's own CompoundInteraction of a hard and a WCA sphere; the soft sweep rejects every move
that makes the hard cores overlap and keeps the energy bookkeeping; hard spherocylinders that also carry WCA on the
segment distance (the large dense workload's soft variant) stay overlap-free under NpT moves."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

HARD_CORE = 0x100
ROD = ra.Shape("spherocylinder", radius=0.5, length=5.0)
C_ROD = ra.Shape("spherocylinder", radius=0.5, length=5.0, soft="wca_segment", epsilon=1.0, sigma=1.0,
                 hard_core=True)
C_SPHERE = ra.Shape("sphere", radius=0.5, soft="wca", epsilon=1.0, sigma=1.0, hard_core=True)


def test_compound_pair_entry_points():
    rec = corpus.spherocylinder_contact_corpus(50000, 7, radius=0.5, length=5.0, delta=0.2)
    assert np.array_equal(ra.pair_decisions(C_ROD, rec), oracle_lib.oracle().pair_decisions(1, 0.5, 5.0, rec))
    e = ra.pair_energies(C_ROD, rec)
    assert np.array_equal(e, oracle_lib.oracle().pair_energies(1, 0.5, 5.0, 2, 1.0, 1.0, rec))
    assert (e > 0).sum() > 1000
    assert C_ROD.range == pytest.approx(5.0 + 2 ** (1 / 6))


def test_compound_sphere_count_and_energy_equal_reference():
    box = np.array([[9.0, 0.4, -0.3], [0.0, 8.5, 0.2], [0.0, 0.0, 9.5]])
    pos, rot = corpus.random_snapshot(600, 21, box, kind="sphere")
    pk = ra.GpuPacking(C_SPHERE, box, pos, rot)
    n_ref, e_ref = oracle_lib.ref().snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1 | HARD_CORE, eps=1.0, sigma=1.0)
    assert pk.count_total_overlaps() == n_ref and n_ref > 0
    assert pk.total_energy() == pytest.approx(e_ref, rel=1e-11)


def test_compound_sphere_sweep_rejects_core_overlaps():
    """NVT compound spheres from a lattice: after Metropolis sweeps no hard cores overlap (the oracle's count), the
    energy equals the oracle's sum over the state, and the same seed gives the same trajectory."""
    box, pos, rot = lattice.hard_sphere_lattice(10, 0.45)
    out = []
    for _ in range(2):
        pk = ra.GpuPacking(C_SPHERE, box, pos, rot, seed=19)
        pk.set_environment(1.0, 1.0, 0.2, None)
        c = pk.run_cycles(150, adapt_steps=True)
        assert 0 < c["moves_accepted"] < c["moves_attempted"]
        b, p, r = pk.download()
        n, e = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, b, p, r, soft=1 | HARD_CORE, eps=1.0, sigma=1.0)
        assert n == 0 and pk.count_total_overlaps() == 0
        assert pk.total_energy() == pytest.approx(e, rel=1e-11)
        out.append(p)
        pk.close()
    assert np.array_equal(out[0], out[1])


def test_compound_rods_npt_overlap_free_and_softer_than_hard_rods():
    """Hard rods with WCA on the segment distance at p* = 2.5 (the large dense workload's soft variant, here on 4 000
    rods): overlap-free by the oracle's count, energy equal to the oracle's sum, and a lower packing fraction than
    plain hard rods at the same pressure (the repulsion widens every rod by up to 2^(1/6) - 1)."""
    box, pos, rot = lattice.spherocylinder_smectic(4000, 0.45)
    phis = {}
    for shape in (C_ROD, ROD):
        pk = ra.GpuPacking(shape, box, pos, rot, seed=23)
        pk.set_environment(1.0, 2.5, 0.1, None, box_move=("delta_triclinic", 0.05))
        pk.run_cycles(600, adapt_steps=True)
        vals = []
        for _ in range(20):
            pk.run_cycles(20)
            vals.append(pk.observables()["packing_fraction"])
        phis[shape.soft] = np.mean(vals)
        b, p, r = pk.download()
        n, e = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r,
                                                  soft=(2 | HARD_CORE) if shape.soft else 0, eps=1.0, sigma=1.0)
        assert n == 0
        if shape.soft:
            assert pk.total_energy() == pytest.approx(e, rel=1e-10) and e > 0
        pk.close()
    print(f"packing fraction at p* = 2.5: compound {phis['wca_segment']:.4f}, hard {phis[None]:.4f}", flush=True)
    assert phis["wca_segment"] < phis[None] - 0.01
