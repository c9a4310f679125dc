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


def test_compound_rods_npt_overlap_free_and_deterministic():
    """Hard rods with WCA on the segment distance (the large dense workload's soft variant) in NpT at p* = 1 from the
    aligned lattice: box moves are accepted, the packing stays overlap-free by the oracle's count, the energy equals
    the oracle's sum over the state, and the same seed gives the same trajectory."""
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(18, 18, 4), spacing=(1.45, 1.45, 6.3))
    out = []
    for _ in range(2):
        pk = ra.GpuPacking(C_ROD, box, pos, rot, seed=23)
        pk.set_environment(1.0, 1.0, 0.2, None, box_move=("delta_triclinic", 0.01))
        c = pk.run_cycles(600, adapt_steps=True)
        assert c["box_accepted"] > 10 and 0 < c["moves_accepted"] < c["moves_attempted"]
        b, p, r = pk.download()
        n, e = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, b, p, r, soft=2 | HARD_CORE, eps=1.0, sigma=1.0)
        assert n == 0 and pk.count_total_overlaps() == 0
        assert pk.total_energy() == pytest.approx(e, rel=1e-10) and e > 0
        out.append((b, p, r))
        pk.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def test_compound_sphere_npt_matches_reference_runs(golden):
    """NpT hard spheres with a WCA shell (CompoundInteraction of SphereTraits and RepulsiveLennardJonesInteraction)
    against independent runs of the reference's own CompoundInteraction (fixture `compound_sphere_npt`): the mean
    packing fraction and energy per particle over independent device runs within two standard errors of the
    difference, each side's error from the spread between its runs."""
    st = next((s for s in golden.get("statistics", []) if s["name"] == "compound_sphere_npt"), None)
    if st is None:
        pytest.skip("no reference statistics in the fixture")
    box, pos, rot = lattice.hard_sphere_lattice(st["n_side"], st["phi0"])
    phi, energy = [], []
    for seed in st["device_seeds"]:
        pk = ra.GpuPacking(C_SPHERE, box, pos, rot, seed=seed)
        pk.set_environment(st["temperature"], st["pressure"], st["trans_step"], None,
                           box_move=("delta_v", st["box_step"]))
        pk.run_cycles(st["therm"], adapt_steps=True)
        f, e = [], []
        for _ in range(st["avg"] // st["every"]):
            pk.run_cycles(st["every"])
            f.append(pk.observables()["packing_fraction"])
            e.append(pk.total_energy() / len(pos))
        phi.append(np.mean(f))
        energy.append(np.mean(e))
        pk.close()
    for name, dev, ref in (("phi", phi, st["phi_runs"]), ("energy", energy, st["energy_runs"])):
        md, sd = np.mean(dev), np.std(dev, ddof=1) / np.sqrt(len(dev))
        mr, sr = np.mean(ref), np.std(ref, ddof=1) / np.sqrt(len(ref))
        z = (md - mr) / np.hypot(sd, sr)
        print(f"{name}: device {md:.5f} +- {sd:.5f} ({len(dev)} runs), reference {mr:.5f} +- {sr:.5f} "
              f"({len(ref)} runs), difference {z:+.2f} standard errors", flush=True)
        assert abs(z) <= 2.0
