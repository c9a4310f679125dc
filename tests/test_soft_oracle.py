"""The oracle's soft (WCA) path, pinned before the device is checked against it.

* WCA on the centre distance (a sphere with RepulsiveLennardJonesInteraction): per-pair energies bit-identical to the
  reference build, the reference's own unit-test values (RepulsiveLennardJonesInteractionTest.cpp:10-25), and total
  energies against the reference build and an all-pairs brute force.
* WCA on the spherocylinder segment distance has no reference implementation: it is pinned by geometry whose segment
  distance is known in closed form (side by side, end to end, crossed) and by a brute-force minimum over sampled
  segment points, and its total energy by an all-images brute force."""
import math

import numpy as np
import pytest

import corpus
import oracle_lib

CUT = 2 ** (1 / 6)


def wca_closed(eps, sigma, d):
    """U(d) = 4 eps ((sigma/d)^12 - (sigma/d)^6) + eps for d < sigma 2^(1/6), else 0 (textbook WCA)."""
    if d >= sigma * CUT:
        return 0.0
    s = sigma / d
    return 4 * eps * (s ** 12 - s ** 6) + eps


def record(p1, o1, p2, o2, t=(0, 0, 0)):
    return np.concatenate([p1, np.ravel(o1), p2, np.ravel(o2), t])[None, :]


def need_ref():
    if not oracle_lib.ref_available():
        pytest.skip("reference build absent")
    return oracle_lib.ref()


@pytest.mark.parametrize("eps,sigma,seed", [(1.0, 1.0, 1), (3.0, 2.0, 2), (0.7, 1.3, 3)])
def test_wca_pair_energies_match_reference_bit_for_bit(eps, sigma, seed):
    ref = need_ref()
    rec = corpus.sphere_contact_corpus(20000, seed, radius=sigma * CUT / 2, delta=0.2 * sigma)
    e_orc = oracle_lib.oracle().pair_energies(0, 0.5, 0.0, 1, eps, sigma, rec)
    e_ref = ref.pair_energies(0, 0.5, 0.0, 1, eps, sigma, rec)
    assert np.array_equal(e_orc, e_ref)
    assert (e_orc > 0).sum() > 5000 and (e_orc == 0).sum() > 5000


def test_wca_reference_unit_values():
    """RepulsiveLennardJonesInteractionTest.cpp:10-25: epsilon 3, sigma 2; distance 2 gives 3, distance 4 gives 0."""
    o = np.eye(3)
    orc = oracle_lib.oracle()
    assert orc.pair_energies(0, 0.5, 0, 1, 3.0, 2.0, record([1, 0, 0], o, [3, 0, 0], o))[0] == 3.0
    assert orc.pair_energies(0, 0.5, 0, 1, 3.0, 2.0, record([1, 0, 0], o, [5, 0, 0], o))[0] == 0.0
    # across a periodic image: p2 + T lies 2 away
    assert orc.pair_energies(0, 0.5, 0, 1, 3.0, 2.0, record([1, 0, 0], o, [13, 0, 0], o, (-10, 0, 0)))[0] == 3.0


def rot_to(axis):
    """A rotation whose column 2 is the unit vector `axis`."""
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    t = np.array([1.0, 0, 0]) if abs(a[0]) < 0.9 else np.array([0, 1.0, 0])
    b1 = np.cross(t, a)
    b1 /= np.linalg.norm(b1)
    b2 = np.cross(a, b1)
    return np.stack([b1, b2, a], axis=1)


@pytest.mark.parametrize("d", [0.9, 1.0, 1.05, 1.12, 1.2])
def test_segment_wca_closed_form_geometries(d):
    orc = oracle_lib.oracle()
    L = 5.0
    z = rot_to([0, 0, 1])
    x = rot_to([1, 0, 0])
    # side by side, parallel: segment distance d
    e = orc.pair_energies(1, 0.5, L, 2, 1.0, 1.0, record([10, 10, 10], z, [10 + d, 10, 10], z))[0]
    assert e == pytest.approx(wca_closed(1.0, 1.0, d), rel=1e-12, abs=1e-14)
    # end to end on one line: gap d between the segment ends
    e = orc.pair_energies(1, 0.5, L, 2, 1.0, 1.0, record([10, 10, 10], z, [10, 10, 10 + L + d], z))[0]
    assert e == pytest.approx(wca_closed(1.0, 1.0, d), rel=1e-12, abs=1e-14)
    # crossed at right angles, centres offset by d along the common normal
    e = orc.pair_energies(1, 0.5, L, 2, 1.0, 1.0, record([10, 10, 10], z, [10, 10 + d, 10], x))[0]
    assert e == pytest.approx(wca_closed(1.0, 1.0, d), rel=1e-12, abs=1e-14)
    # T shape: the end of the second rod d away from the middle of the first
    e = orc.pair_energies(1, 0.5, L, 2, 1.0, 1.0, record([10, 10, 10], z, [10 + d + L / 2, 10, 10], x))[0]
    assert e == pytest.approx(wca_closed(1.0, 1.0, d), rel=1e-12, abs=1e-14)


def brute_segment_distance(p1, a1, p2, a2, half, k=801):
    """Minimum distance between the segments by dense sampling of one parameter and the exact point-segment
    distance for the other (long double is unavailable in numpy; float64 with 801 samples plus local refinement)."""
    s = np.linspace(-half, half, k)
    best = np.inf
    for _ in range(3):
        pts = p1[None, :] + s[:, None] * a1[None, :]
        w = pts - p2[None, :]
        t = np.clip(w @ a2, -half, half)
        d = np.linalg.norm(w - t[:, None] * a2[None, :], axis=1)
        i = int(np.argmin(d))
        best = min(best, d[i])
        lo, hi = s[max(i - 1, 0)], s[min(i + 1, k - 1)]
        s = np.linspace(lo, hi, k)
    return best


def test_segment_wca_matches_brute_force_minimum():
    rng = np.random.Generator(np.random.PCG64(5))
    orc = oracle_lib.oracle()
    L = 5.0
    rec = corpus.spherocylinder_contact_corpus(300, 6, radius=CUT / 2, length=L, delta=0.3)
    e = orc.pair_energies(1, 0.5, L, 2, 1.0, 1.0, rec)
    for r, ei in zip(rec, e):
        p1, p2 = r[0:3], r[12:15] + r[24:27]
        a1, a2 = r[3:12].reshape(3, 3)[:, 2], r[15:24].reshape(3, 3)[:, 2]
        d = brute_segment_distance(p1, a1, p2, a2, L / 2)
        if abs(d - CUT) < 1e-6:
            continue  # the cut-off decision itself is within the brute force's resolution
        assert ei == pytest.approx(wca_closed(1.0, 1.0, d), rel=1e-6, abs=1e-9)
    assert (e > 0).sum() > 50


def brute_total_energy(box, pos, rot, kind, eps, sigma, length):
    """All pairs over all 27 images (valid while at most one image of a pair is within range)."""
    n = len(pos)
    tr = np.array([box @ np.array([i, j, k]) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], float)
    axes = rot.reshape(-1, 3, 3)[:, :, 2]
    total = 0.0
    for i in range(n):
        for t in tr:
            p2 = pos[i + 1:] + t
            if kind == 0:
                d2 = ((p2 - pos[i]) ** 2).sum(1)
            else:
                m = len(p2)
                d2 = corpus.segment_distance2(np.repeat(pos[i][None], m, 0), np.repeat(axes[i][None], m, 0), p2,
                                              axes[i + 1:], length / 2)
            d2 = d2[d2 < (sigma * CUT) ** 2]
            x6 = (sigma * sigma / d2) ** 3
            total += float((4 * eps * (x6 * x6 - x6) + eps).sum())
    return total


def test_wca_sphere_total_energy_matches_reference_and_brute_force():
    ref = need_ref()
    box = np.array([[9.0, 0.4, -0.3], [0.0, 8.5, 0.2], [0.0, 0.0, 9.5]])
    pos, rot = corpus.random_snapshot(600, 11, box, kind="sphere")
    _, e_orc = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=1.0)
    n_ref, e_ref = ref.snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=1.0)
    assert n_ref == 0  # the WCA sphere has no hard part
    assert e_orc == pytest.approx(e_ref, rel=1e-12)
    assert e_orc == pytest.approx(brute_total_energy(box, pos, rot, 0, 1.0, 1.0, 0.0), rel=1e-9)
    assert e_orc > 0


def test_segment_wca_total_energy_matches_brute_force():
    box = np.array([[16.0, 0.5, -0.4], [0.0, 15.0, 0.3], [0.0, 0.0, 17.0]])
    pos, rot = corpus.random_snapshot(120, 12, box)
    n, e = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, box, pos, rot, soft=2, eps=1.0, sigma=1.0)
    assert n == 0
    assert e == pytest.approx(brute_total_energy(box, pos, rot, 1, 1.0, 1.0, 5.0), rel=1e-9)
    assert e > 0


def test_oracle_total_energy_matches_reference_fixture(golden):
    for s in golden.get("soft_snapshots", []):
        box, pos, rot = corpus.jittered_lattice(s["n_side"], s["spacing"], s["jitter"], s["seed"])
        assert corpus.records_sha256(np.concatenate([pos.ravel(), rot.ravel()])) == s["sha256"], s["name"]
        _, e = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=1.0)
        assert e == pytest.approx(s["energy"], rel=1e-12)
        assert e > 0


# Lennard-Jones (soft 3) and square-inverse core (soft 4) on spheres
@pytest.mark.parametrize("soft,eps,sigma,seed", [(3, 1.0, 1.0, 4), (3, 0.5, 1.4, 5), (4, 2.0, 1.0, 6),
                                                  (4, -1.5, 1.2, 7)])
def test_lj_and_sic_pair_energies_match_reference_bit_for_bit(soft, eps, sigma, seed):
    """Per pair, Interaction::calculateEnergyBetween of the reference build (LennardJonesInteraction,
    SquareInverseCoreInteraction on a sphere) at centre distances from 0.5 to 3.5 sigma, untruncated."""
    ref = need_ref()
    rec = corpus.sphere_contact_corpus(20000, seed, radius=sigma, delta=1.5 * sigma)
    e_orc = oracle_lib.oracle().pair_energies(0, 0.5, 0.0, soft, eps, sigma, rec)
    e_ref = ref.pair_energies(0, 0.5, 0.0, soft, eps, sigma, rec)
    assert np.array_equal(e_orc, e_ref)
    assert (e_orc != 0).sum() > 2500  # SIC: nonzero within sigma only, a sixth of the corpus
    if soft == 3:
        d = np.linalg.norm(rec[:, 12:15] + rec[:, 24:27] - rec[:, 0:3], axis=1)
        assert ((d > 3 * sigma) & (e_orc != 0)).sum() > 1000  # the pair formula has no cut-off


def test_lj_and_sic_closed_forms():
    """LJ: U(sigma) = 0 and the minimum U(2^(1/6) sigma) = -eps; SIC: eps (sigma^2 / d^2 - 1) inside sigma (eps at
    d = sigma / sqrt 2, attractive for eps < 0), 0 outside."""
    o = np.eye(3)
    orc = oracle_lib.oracle()
    e = lambda soft, eps, sigma, d: orc.pair_energies(0, 0.5, 0, soft, eps, sigma,
                                                      record([1, 1, 1], o, [1 + d, 1, 1], o))[0]
    assert e(3, 2.0, 1.5, 1.5) == 0.0
    assert e(3, 2.0, 1.5, 1.5 * CUT) == pytest.approx(-2.0, rel=1e-13)
    assert e(4, 3.0, 2.0, 2.0 / math.sqrt(2)) == pytest.approx(3.0, rel=1e-14)
    assert e(4, -3.0, 2.0, 2.0 / math.sqrt(2)) == pytest.approx(-3.0, rel=1e-14)
    assert e(4, 3.0, 2.0, 2.0) == 0.0 and e(4, 3.0, 2.0, 2.5) == 0.0


def brute_central_total(box, pos, soft, eps, sigma):
    """All pairs over all 27 images, each pair's energy truncated at the interaction range (LJ 3 sigma, SIC
    sigma), valid while at most one image of a pair is within range."""
    rng2 = (3 * sigma) ** 2 if soft == 3 else sigma ** 2
    tr = np.array([box @ np.array([i, j, k]) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], float)
    total = 0.0
    for i in range(len(pos)):
        for t in tr:
            d2 = ((pos[i + 1:] + t - pos[i]) ** 2).sum(1)
            d2 = d2[d2 < rng2]
            if soft == 3:
                x6 = (sigma * sigma / d2) ** 3
                total += float((4 * eps * (x6 * x6 - x6)).sum())
            else:
                total += float((eps * np.maximum(0.0, sigma ** 2 / d2 - 1)).sum())
    return total


@pytest.mark.parametrize("soft,eps,sigma", [(3, 1.0, 1.0), (4, 1.5, 1.0), (4, -0.8, 1.1)])
def test_lj_and_sic_total_energies(soft, eps, sigma):
    """Total energies: the oracle (truncated at the range) equals an all-images brute force; for the square-inverse
    core, which vanishes beyond its range, also the reference build's getTotalEnergy."""
    box = np.array([[14.0, 0.5, -0.3], [0.0, 13.5, 0.4], [0.0, 0.0, 14.5]])
    box_, pos, rot = corpus.jittered_lattice(12, 14.0 / 12, 0.2, 20 + soft)
    _, e_orc = oracle_lib.oracle().snapshot_stats(0, 0.5, 0.0, box_, pos, rot, soft=soft, eps=eps, sigma=sigma)
    assert e_orc == pytest.approx(brute_central_total(box_, pos, soft, eps, sigma), rel=1e-9)
    assert e_orc != 0
    if soft == 4 and oracle_lib.ref_available():
        _, e_ref = oracle_lib.ref().snapshot_stats(0, 0.5, 0.0, box_, pos, rot, soft=soft, eps=eps, sigma=sigma)
        assert e_orc == pytest.approx(e_ref, rel=1e-12)


HARD_CORE = 0x100


@pytest.mark.parametrize("radius,sigma", [(0.5, 1.0), (0.45, 1.0), (0.5, 0.8)])
def test_compound_sphere_counts_and_energies_match_reference(radius, sigma):
    """CompoundInteraction of a hard sphere and a WCA sphere (CompoundInteraction.cpp:9-60) in the oracle against the
    reference's own CompoundInteraction: the overlap count is the hard part's, the energy the soft part's (also over
    overlapping pairs, which getTotalEnergy sums like any other), the range the larger one."""
    ref = need_ref()
    box = np.array([[9.0, 0.4, -0.3], [0.0, 8.5, 0.2], [0.0, 0.0, 9.5]])
    pos, rot = corpus.random_snapshot(600, 21, box, kind="sphere")
    soft = 1 | HARD_CORE
    n_orc, e_orc = oracle_lib.oracle().snapshot_stats(0, radius, 0.0, box, pos, rot, soft=soft, eps=1.0, sigma=sigma)
    n_ref, e_ref = ref.snapshot_stats(0, radius, 0.0, box, pos, rot, soft=soft, eps=1.0, sigma=sigma)
    n_hard, _ = ref.snapshot_stats(0, radius, 0.0, box, pos, rot)
    _, e_soft = ref.snapshot_stats(0, radius, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=sigma)
    assert n_orc == n_ref == n_hard and n_ref > 0
    assert e_orc == pytest.approx(e_ref, rel=1e-12)
    assert e_ref == pytest.approx(e_soft, rel=1e-12)


def test_compound_segment_wca_rods_count_and_energy():
    """Hard spherocylinders that also carry WCA on the segment distance (no reference counterpart of the soft part):
    the count equals the plain hard rods', the energy the segment WCA's, which is pinned by brute force."""
    box = np.array([[16.0, 0.5, -0.4], [0.0, 15.0, 0.3], [0.0, 0.0, 17.0]])
    pos, rot = corpus.random_snapshot(120, 12, box)
    orc = oracle_lib.oracle()
    n_c, e_c = orc.snapshot_stats(1, 0.5, 5.0, box, pos, rot, soft=2 | HARD_CORE, eps=1.0, sigma=1.0)
    n_h, _ = orc.snapshot_stats(1, 0.5, 5.0, box, pos, rot)
    assert n_c == n_h and n_h > 0
    assert e_c == pytest.approx(brute_total_energy(box, pos, rot, 1, 1.0, 1.0, 5.0), rel=1e-9)
    if oracle_lib.ref_available():
        n_ref, _ = oracle_lib.ref().snapshot_stats(1, 0.5, 5.0, box, pos, rot)
        assert n_c == n_ref
