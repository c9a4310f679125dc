"""Guarded FP32 pair decisions, checked on the CPU before the device is checked against them: the oracle's
restatement of the guarded classifier (FP32 distances with error bounds, FP64 for the pairs inside the guard band)
must give the reference's FP64 decision on every pair, including the committed corpora whose decisions come from the
reference build (within 1e-6 and within 2e-15 of contact) and geometries chosen to stress the bounds (nearly parallel,
crossed, end-to-end, far).  Bounds planted at zero must produce mismatches on the near-contact corpora, so these
checks would catch a bound that is too small."""
import numpy as np
import pytest

import corpus
import oracle_lib
from test_golden import corpus_records, golden_decisions


def check(kind, radius, length, rec, expect=None):
    ref = oracle_lib.oracle().pair_decisions(kind, radius, length, rec)
    if expect is not None:
        assert np.array_equal(ref, expect)
    dec, codes = oracle_lib.pair_guarded(kind, radius, length, rec)
    assert int((dec != ref).sum()) == 0
    assert not (codes & 8).any()
    return codes


def test_guarded_decisions_equal_reference_fixtures(golden):
    for c in golden["corpora"]:
        codes = check(c["kind"], c["radius"], c["length"], corpus_records(c), golden_decisions(c))
        fallback = float(((codes & 3) == 2).mean())
        print(f"{c['name']}: FP64 fallback {fallback:.4f}", flush=True)


@pytest.mark.parametrize("delta,seed", [(1e-3, 1), (1e-2, 2), (0.3, 3), (2.0, 4)])
def test_guarded_spherocylinder_corpora_and_fallback_rate(delta, seed):
    rec = corpus.spherocylinder_contact_corpus(300000, seed, radius=0.5, length=5.0, delta=delta)
    codes = check(1, 0.5, 5.0, rec)
    fallback = float(((codes & 3) == 2).mean())
    print(f"delta {delta}: FP64 fallback {fallback:.5f}, segment stage {float(((codes & 4) > 0).mean()):.3f}")
    # the guard band is narrow: away from contact almost every pair is decided in FP32
    if delta >= 0.3:
        assert fallback < 0.01


def special_records(n, seed, mode):
    """Pairs of rods at random places with designed relative poses: 'parallel' (axes equal up to 1e-7),
    'crossed' (perpendicular axes, centres 1 +- 1e-4 apart along the common normal), 'end' (collinear, gap
    1 +- 1e-4 between the caps)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rot = corpus.random_rotations(rng, n)
    pos1 = rng.random((n, 3)) * 40
    rec = np.zeros((n, 27))
    for i in range(n):
        R = rot[i].reshape(3, 3)
        a, b, c = R[:, 0], R[:, 1], R[:, 2]
        if mode == "parallel":
            tilt = rng.normal(size=3) * 1e-7
            a2 = c + tilt
            a2 /= np.linalg.norm(a2)
            sep = a * (1 + rng.uniform(-1e-4, 1e-4)) + c * rng.uniform(-6, 6)
        elif mode == "crossed":
            a2 = a
            sep = b * (1 + rng.uniform(-1e-4, 1e-4)) + c * rng.uniform(-2, 2) + a * rng.uniform(-2, 2)
        else:
            a2 = c
            sep = c * (5 + 1 + rng.uniform(-1e-4, 1e-4))
        # second orientation: any rotation whose third column is a2
        t = np.array([1.0, 0, 0]) if abs(a2[0]) < 0.9 else np.array([0, 1.0, 0])
        e1 = np.cross(t, a2)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(a2, e1)
        R2 = np.stack([e1, e2, a2], axis=1)
        rec[i, 0:3] = pos1[i]
        rec[i, 3:12] = rot[i]
        rec[i, 12:15] = pos1[i] + sep
        rec[i, 15:24] = R2.ravel()
    return rec


@pytest.mark.parametrize("mode", ["parallel", "crossed", "end"])
def test_guarded_special_geometries(mode):
    rec = special_records(20000, {"parallel": 7, "crossed": 8, "end": 9}[mode], mode)
    codes = check(1, 0.5, 5.0, rec)
    if mode == "parallel":
        # nearly parallel pairs with the centre distance in range always take the FP64 path
        assert (((codes & 4) > 0) <= ((codes & 3) == 2)).all()


def test_guarded_spheres_and_random_far_pairs():
    check(0, 0.5, 0.0, corpus.sphere_contact_corpus(300000, 10, radius=0.5, delta=1e-3))
    check(0, 0.5, 0.0, corpus.sphere_contact_corpus(300000, 11, radius=0.5, delta=0.4))
    check(1, 0.5, 2.0, corpus.spherocylinder_contact_corpus(200000, 12, radius=0.5, length=2.0, delta=0.05))


def test_bounds_planted_at_zero_are_caught(golden):
    mismatches = 0
    for c in golden["corpora"]:
        rec = corpus_records(c)
        dec, _ = oracle_lib.pair_guarded(c["kind"], c["radius"], c["length"], rec, bound_scale=0.0)
        mismatches += int((dec != golden_decisions(c)).sum())
    assert mismatches > 0


def test_guarded_decisions_on_1e7_near_contact_pairs(golden):
    """The full decision corpus: 1e7 spherocylinder pairs within +-1e-6 of contact; the FP64 oracle and the guarded
    restatement both reproduce the reference's decision bitmap (compared through its SHA-256)."""
    import hashlib
    c = golden["large_corpora"][0]
    rec = corpus.chunked_contact_corpus(c["n"], c["seed"], radius=c["radius"], length=c["length"], delta=c["delta"],
                                        chunk=c["chunk"])
    assert corpus.records_sha256(rec) == c["sha256"]
    ref = oracle_lib.oracle().pair_decisions(c["kind"], c["radius"], c["length"], rec)
    assert int(ref.sum()) == c["overlaps"]
    assert hashlib.sha256(np.packbits(ref).tobytes()).hexdigest() == c["decisions_sha256"]
    dec, codes = oracle_lib.pair_guarded(c["kind"], c["radius"], c["length"], rec)
    assert np.array_equal(dec, ref)
    assert not (codes & 8).any()
