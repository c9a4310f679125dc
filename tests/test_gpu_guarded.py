"""Guarded FP32 pair decisions on the device (rp_pair_decisions_guarded): every decision equals the FP64 decision
(the reference's, through the committed corpora, and the device's FP64 entry point), the per-pair classification
codes are bit for bit the oracle's restatement of the guarded classifier, and the statistics report the guard-band
hit rate and the largest FP32 error.  The production form (no verification) gives the same decisions."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from test_golden import corpus_records, golden_decisions

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
HS = ra.Shape("sphere", radius=0.5)


def run(shape, kind, radius, length, rec, expect=None):
    dec, codes, st = ra.pair_decisions_guarded(shape, rec)
    fp64 = ra.pair_decisions(shape, rec)
    assert np.array_equal(dec, fp64)
    if expect is not None:
        assert np.array_equal(dec, expect)
    assert st["mismatches"] == 0 and st["pairs"] == len(rec)
    _, ocodes = oracle_lib.pair_guarded(kind, radius, length, rec)
    assert np.array_equal(codes, ocodes)
    fast = ra.pair_decisions_guarded(shape, rec, verify=False)
    assert np.array_equal(fast, fp64)
    return st


def test_guarded_decisions_on_reference_fixtures(golden):
    for c in golden["corpora"]:
        shape = SC if c["kind"] == 1 else HS
        st = run(shape, c["kind"], c["radius"], c["length"], corpus_records(c), golden_decisions(c))
        print(f"{c['name']}: {st}", flush=True)
        assert st["fp64_fallback"] > 0  # near contact: the guard band is used


@pytest.mark.parametrize("delta,seed", [(1e-3, 21), (0.3, 22)])
def test_guarded_statistics_away_from_contact(delta, seed):
    rec = corpus.spherocylinder_contact_corpus(500000, seed, radius=0.5, length=5.0, delta=delta)
    st = run(SC, 1, 0.5, 5.0, rec)
    rate = st["fp64_fallback"] / st["pairs"]
    print(f"delta {delta}: FP64 fallback {rate:.5f}, max FP32 error centre {st['max_err_centre']:.3g} "
          f"segment {st['max_err_segment']:.3g}", flush=True)
    assert st["max_err_centre"] < 1e-4 and st["max_err_segment"] < 1e-3
    if delta >= 0.3:
        assert rate < 0.01


def test_guarded_spheres():
    rec = corpus.sphere_contact_corpus(500000, 23, radius=0.5, delta=1e-3)
    st = run(HS, 0, 0.5, 0.0, rec)
    assert st["segment_stage"] == 0


def test_device_decisions_on_1e7_near_contact_pairs(golden):
    """1e7 spherocylinder pairs within +-1e-6 of contact: the device's FP64 and guarded FP32 decisions both
    reproduce the reference's decision bitmap (SHA-256 of the packed bits), with zero guarded/FP64 mismatches."""
    import hashlib
    c = golden["large_corpora"][0]
    rec = corpus.chunked_contact_corpus(c["n"], c["seed"], radius=c["radius"], length=c["length"], delta=c["delta"],
                                        chunk=c["chunk"], threads=16)
    assert corpus.records_sha256(rec) == c["sha256"]
    print("corpus ready", flush=True)
    fp64 = ra.pair_decisions(SC, rec)
    assert hashlib.sha256(np.packbits(fp64).tobytes()).hexdigest() == c["decisions_sha256"]
    dec, codes, st = ra.pair_decisions_guarded(SC, rec)
    assert np.array_equal(dec, fp64)
    assert st["mismatches"] == 0 and not (codes & 8).any()
    fast = ra.pair_decisions_guarded(SC, rec, verify=False)
    assert np.array_equal(fast, fp64)
    print(f"1e7 near-contact pairs: {st}", flush=True)
