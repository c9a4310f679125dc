"""Polyspheres (PolysphereTraits) in the oracle, checked on the CPU: total overlap counts over the grid of interaction
centres equal the reference build's (committed fixtures) and an all-images brute force; pair decisions follow the bead
geometry (touching beads, contact distances built by construction) and agree with the grid count of the same two
particles; the kinked-chain generator has touching beads and the stated kink."""
import numpy as np
import pytest

import corpus
import oracle_lib
from oracle_lib import _dp
from rampack_amd import lattice

POLY = 2


def count(box, pos, rot, centres, radii, early=False):
    return oracle_lib.oracle().snapshot_stats(POLY, 0.0, 0.0, box, pos, rot, early_exit=early, centres=centres,
                                              radii=radii)[0]


def brute(box, pos, rot, centres, radii):
    o = oracle_lib.oracle()
    c = np.ascontiguousarray(centres, dtype=np.float64)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64).ravel()
    p = np.ascontiguousarray(pos)
    q = np.ascontiguousarray(rot)
    return o.lib.orc_count_brute(POLY, 0.0, 0.0, len(c), c.ctypes.data_as(_dp), r.ctypes.data_as(_dp),
                                 b.ctypes.data_as(_dp), len(p), p.ctypes.data_as(_dp), q.ctypes.data_as(_dp))


def test_counts_match_reference_fixtures(golden):
    for s in golden["poly_snapshots"]:
        centres, radii = corpus.bead_set(s["beads"])
        box = np.array(s["box"])
        pos, rot = corpus.random_snapshot(s["n"], s["seed"], box)
        assert corpus.records_sha256(np.concatenate([pos.ravel(), rot.ravel()])) == s["sha256"]
        assert count(box, pos, rot, centres, radii) == s["overlaps"], s["name"]
        assert count(box, pos, rot, centres, radii, early=True) == s["overlaps_early_exit"], s["name"]


@pytest.mark.parametrize("beads,seed", [("kinked11", 1), ("dimer", 2), ("mixed", 3)])
def test_counts_match_brute_force(beads, seed):
    centres, radii = corpus.bead_set(beads)
    box = np.array([[16.0, 0.8, -0.4], [0.0, 15.0, 0.6], [0.0, 0.0, 17.0]])
    pos, rot = corpus.random_snapshot(150, seed, box)
    assert count(box, pos, rot, centres, radii) == brute(box, pos, rot, centres, radii)


def records_for(centres, n, seed, spread):
    """Records of two particles with random orientations, centres `spread`-uniform apart, T = 0 or a box vector."""
    rng = np.random.Generator(np.random.PCG64(seed))
    o1 = corpus.random_rotations(rng, n)
    o2 = corpus.random_rotations(rng, n)
    p1 = rng.random((n, 3)) * 20 + 10
    p2 = p1 + (rng.random((n, 3)) * 2 - 1) * spread
    T = np.zeros((n, 3))
    shift = rng.random(n) < 0.5
    T[shift] = [40.0, 0.0, 0.0]
    p2[shift] -= [40.0, 0.0, 0.0]
    return np.ascontiguousarray(np.concatenate([p1, o1, p2, o2, T], axis=1))


@pytest.mark.parametrize("beads", ["kinked11", "dimer", "mixed"])
def test_pair_decisions_agree_with_two_particle_counts(beads):
    """A record's decision is whether the two-particle snapshot (particle 2 placed at p2 + T in a large box) has any
    overlapping bead pair on the grid count, which the reference fixtures pin."""
    centres, radii = corpus.bead_set(beads)
    ext = float(np.sqrt((centres * centres).sum(1).max()))
    rec = records_for(centres, 400, 7, 2 * ext + 1.5)
    dec = oracle_lib.poly_pair_decisions(centres, radii, rec)
    assert 0 < dec.sum() < len(rec)
    box = np.diag([60.0, 60.0, 60.0])
    for i in range(len(rec)):
        pos = np.stack([rec[i, 0:3], rec[i, 12:15] + rec[i, 24:27]])
        rot = np.stack([rec[i, 3:12], rec[i, 15:24]])
        assert (count(box, pos, rot, centres, radii) > 0) == bool(dec[i]), i


def test_pair_decisions_at_constructed_contact():
    """Dimers (beads of radius 0.5 at z = +-0.5) end to end along z: the gap between the facing beads is d - 2, so
    the pair overlaps exactly when the centre distance d is below 2 (tested at 2 -+ 1e-9), whatever the image T."""
    centres, radii = corpus.bead_set("dimer")
    eye = np.eye(3).ravel()
    rows = []
    for d, T in [(2 - 1e-9, 0.0), (2 + 1e-9, 0.0), (2 - 1e-9, 50.0), (2 + 1e-9, 50.0), (1.0, 0.0)]:
        p1 = np.array([5.0, 5.0, 5.0])
        p2 = p1 + [0.0, 0.0, d] - [0.0, 0.0, T]
        rows.append(np.concatenate([p1, eye, p2, eye, [0.0, 0.0, T]]))
    dec = oracle_lib.poly_pair_decisions(centres, radii, np.array(rows))
    assert dec.tolist() == [1, 0, 1, 0, 1]
    # side by side at distance 1 -+ 1e-9 (beads touch sideways)
    rows = [np.concatenate([[5.0, 5.0, 5.0], eye, [5.0 + x, 5.0, 5.0], eye, [0.0, 0.0, 0.0]]) for x in
            (1 - 1e-9, 1 + 1e-9)]
    assert oracle_lib.poly_pair_decisions(centres, radii, np.array(rows)).tolist() == [1, 0]


def test_kinked_chain_geometry():
    c, r = lattice.kinked_chain(11, 1.0, 30.0)
    steps = np.linalg.norm(np.diff(c, axis=0), axis=1)
    assert np.allclose(steps, 1.0) and np.allclose(r, 0.5)
    arm1 = c[5] - c[0]
    arm2 = c[10] - c[5]
    cosang = arm1 @ arm2 / np.linalg.norm(arm1) / np.linalg.norm(arm2)
    assert np.isclose(np.degrees(np.arccos(cosang)), 30.0)
    assert np.allclose(c.mean(axis=0), 0.0, atol=1e-12)
