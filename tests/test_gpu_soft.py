"""Soft (WCA) interactions on the device: pair energies bit-identical to the reference build (WCA spheres) and to the
oracle (WCA on the spherocylinder segment distance, which has no reference implementation); total energies against
the oracle within the rounding of the sum; the soft sweep's Metropolis moves and box moves against the reference's
own statistics (committed fixture) and its invariants."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

CUT = 2 ** (1 / 6)
WCA_SPHERE = ra.Shape("sphere", radius=0.5, soft="wca", epsilon=1.0, sigma=1.0)
WCA_ROD = ra.Shape("spherocylinder", radius=0.5, length=5.0, soft="wca_segment", epsilon=1.0, sigma=1.0)


@pytest.mark.parametrize("eps,sigma,seed", [(1.0, 1.0, 1), (3.0, 2.0, 2), (0.7, 1.3, 3)])
def test_wca_sphere_pair_energies_are_bit_identical_to_oracle(eps, sigma, seed):
    rec = corpus.sphere_contact_corpus(50000, seed, radius=sigma * CUT / 2, delta=0.2 * sigma)
    shape = ra.Shape("sphere", radius=0.5, soft="wca", epsilon=eps, sigma=sigma)
    gpu = ra.pair_energies(shape, rec)
    cpu = oracle_lib.oracle().pair_energies(0, 0.5, 0.0, 1, eps, sigma, rec)
    assert np.array_equal(gpu, cpu)
    assert (gpu > 0).sum() > 10000


def test_wca_sphere_pair_energies_match_reference_fixture(golden):
    for c in golden.get("soft_corpora", []):
        rec = corpus.sphere_contact_corpus(c["n"], c["seed"], radius=c["sigma"] * CUT / 2, delta=c["delta"])
        assert corpus.records_sha256(rec) == c["sha256"], c["name"]
        shape = ra.Shape("sphere", radius=0.5, soft="wca", epsilon=c["eps"], sigma=c["sigma"])
        e = ra.pair_energies(shape, rec)
        assert corpus.records_sha256(e) == c["energies_sha256"], c["name"]


@pytest.mark.parametrize("seed,delta", [(4, 0.3), (5, 1e-6)])
def test_segment_wca_pair_energies_are_bit_identical_to_oracle(seed, delta):
    rec = corpus.spherocylinder_contact_corpus(50000, seed, radius=CUT / 2, length=5.0, delta=delta)
    gpu = ra.pair_energies(WCA_ROD, rec)
    cpu = oracle_lib.oracle().pair_energies(1, 0.5, 5.0, 2, 1.0, 1.0, rec)
    assert np.array_equal(gpu, cpu)


def test_total_energy_matches_oracle_and_reference_fixture(golden):
    for s in golden.get("soft_snapshots", []):
        box, pos, rot = corpus.jittered_lattice(s["n_side"], s["spacing"], s["jitter"], s["seed"])
        assert corpus.records_sha256(np.concatenate([pos.ravel(), rot.ravel()])) == s["sha256"], s["name"]
        _, orc = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=1.0)
        assert orc == pytest.approx(s["energy"], rel=1e-12)
        pk = ra.GpuPacking(WCA_SPHERE, box, pos, rot)
        e = pk.total_energy()
        print(f"{s['name']}: gpu {e!r} reference {s['energy']!r}", flush=True)
        # the sums run in different orders: agreement to the rounding of an N-pair sum of positive terms
        assert e == pytest.approx(s["energy"], rel=1e-11)
        pk.close()


def test_total_energy_matches_oracle_on_rods_and_is_zero_for_hard_shapes():
    box = np.array([[22.0, 0.5, -0.4], [0.0, 21.0, 0.3], [0.0, 0.0, 23.0]])
    pos, rot = corpus.random_snapshot(500, 31, box)
    pk = ra.GpuPacking(WCA_ROD, box, pos, rot)
    _, expected = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, box, pos, rot, soft=2, eps=1.0, sigma=1.0)
    assert pk.total_energy() == pytest.approx(expected, rel=1e-11)
    pk.close()
    hard = ra.GpuPacking(ra.Shape("spherocylinder", 0.5, 5.0), box, pos, rot)
    assert hard.total_energy() == 0.0
    hard.close()


def test_wca_sweep_keeps_energy_bookkeeping_consistent():
    """NVT soft spheres: after many Metropolis sweeps the total energy is finite and equals the oracle's sum over
    the downloaded state; same seed, same trajectory."""
    box, pos, rot = lattice.hard_sphere_lattice(10, 0.40)
    out = []
    for _ in range(2):
        pk = ra.GpuPacking(WCA_SPHERE, box, pos, rot, seed=9)
        pk.set_environment(1.0, 1.0, 0.15, None)
        c = pk.run_cycles(100, adapt_steps=True)
        assert 0 < c["moves_accepted"] < c["moves_attempted"]
        b, p, r = pk.download()
        e = pk.total_energy()
        _, expected = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, b, p, r, soft=1, eps=1.0, sigma=1.0)
        assert e == pytest.approx(expected, rel=1e-11)
        assert 0 < e / len(p) < 5
        out.append(p)
        pk.close()
    assert np.array_equal(out[0], out[1])


def test_wca_sphere_npt_density_matches_reference(golden):
    """NpT WCA spheres against the reference's own run of the same state point (fixture written from the reference
    build with the reference's Packing, samplers and scaler): mean density and energy per particle within 2
    combined standard errors, each side's error block-averaged over its run's correlated samples."""
    st = next((s for s in golden.get("statistics", []) if s["name"] == "wca_sphere_npt"), None)
    if st is None:
        pytest.skip("no reference statistics in the fixture")
    box, pos, rot = lattice.hard_sphere_lattice(st["n_side"], st["phi0"])
    pk = ra.GpuPacking(WCA_SPHERE, box, pos, rot, seed=4321)
    pk.set_environment(st["temperature"], st["pressure"], 0.1, None, box_move=("delta_v", st["box_step"]))
    pk.run_cycles(st["therm"], adapt_steps=True)
    rho, energy = [], []
    for _ in range(st["avg"] // st["every"]):
        pk.run_cycles(st["every"])
        o = pk.observables()
        rho.append(o["number_density"])
        energy.append(pk.total_energy() / len(pos))
    m_rho, e_rho = ra.mean_and_error(rho, method="blocks")
    m_e, e_e = ra.mean_and_error(energy, method="blocks")
    print(f"gpu rho {m_rho} +- {e_rho}, reference {st['rho']} +- {st['rho_err']}; "
          f"gpu E/N {m_e} +- {e_e}, reference {st['energy']} +- {st['energy_err']}", flush=True)
    assert abs(m_rho - st["rho"]) <= 2 * np.hypot(e_rho, st["rho_err"])
    assert abs(m_e - st["energy"]) <= 2 * np.hypot(e_e, st["energy_err"])


def test_soft_shape_validation():
    box, pos, rot = lattice.hard_sphere_lattice(4, 0.2)
    with pytest.raises(ra.RampackError):
        ra.GpuPacking(ra.Shape("sphere", 0.5, soft="wca_segment"), box, pos, rot)
    with pytest.raises(ra.PreconditionError):
        ra.GpuPacking(ra.Shape("sphere", 0.5, soft="wca", sigma=0.0), box, pos, rot)
    with pytest.raises(ra.RampackError):
        ra.pair_decisions(WCA_SPHERE, np.zeros((1, 27)))


@pytest.mark.parametrize("soft,name,eps,sigma,seed", [(3, "lj", 1.0, 1.0, 41), (3, "lj", 0.5, 1.4, 42),
                                                       (4, "sic", 2.0, 1.0, 43), (4, "sic", -1.5, 1.2, 44)])
def test_lj_and_sic_pair_energies_are_bit_identical_to_oracle(soft, name, eps, sigma, seed):
    """Interaction::calculateEnergyBetween per pair (Lennard-Jones untruncated), centre distances 0.5-3.5 sigma."""
    rec = corpus.sphere_contact_corpus(50000, seed, radius=sigma, delta=1.5 * sigma)
    shape = ra.Shape("sphere", radius=0.5, soft=name, epsilon=eps, sigma=sigma)
    gpu = ra.pair_energies(shape, rec)
    cpu = oracle_lib.oracle().pair_energies(0, 0.5, 0.0, soft, eps, sigma, rec)
    assert np.array_equal(gpu, cpu)
    assert (gpu != 0).sum() > 5000


@pytest.mark.parametrize("soft,name,eps,sigma,temperature", [(3, "lj", 1.0, 1.0, 1.5), (4, "sic", 1.0, 1.0, 1.0),
                                                              (4, "sic", -0.5, 1.0, 2.0)])
def test_lj_and_sic_totals_and_sweep_bookkeeping(soft, name, eps, sigma, temperature):
    """Total energy (truncated at the range) equals the oracle's on a jittered lattice and, after NVT Metropolis
    sweeps, on the downloaded state; same seed, same trajectory."""
    shape = ra.Shape("sphere", radius=0.5, soft=name, epsilon=eps, sigma=sigma)
    box, pos, rot = corpus.jittered_lattice(12, 14.0 / 12, 0.2, 20 + soft)
    pk = ra.GpuPacking(shape, box, pos, rot, seed=3)
    _, expected = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=soft, eps=eps, sigma=sigma)
    assert pk.total_energy() == pytest.approx(expected, rel=1e-11, abs=1e-9)
    out = []
    for _ in range(2):
        pk = ra.GpuPacking(shape, box, pos, rot, seed=9)
        pk.set_environment(temperature, 1.0, 0.15, None)
        c = pk.run_cycles(60, adapt_steps=True)
        assert 0 < c["moves_accepted"] < c["moves_attempted"]
        b, p, r = pk.download()
        e = pk.total_energy()
        _, expected = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, b, p, r, soft=soft, eps=eps, sigma=sigma)
        print(f"{name} eps {eps}: E/N {e / len(p)}", flush=True)
        assert np.isfinite(e) and e == pytest.approx(expected, rel=1e-10, abs=1e-8)
        out.append(p)
        pk.close()
    assert np.array_equal(out[0], out[1])
