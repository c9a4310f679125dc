"""The oracle against the reference: committed golden fixtures (written by tests/golden/make_golden.py from the
reference's own compiled hot path) and, where it is built, oracle/_ref itself."""
import base64
import ctypes
import os
import subprocess

import numpy as np
import pytest

import corpus
import oracle_lib


def corpus_records(c):
    if c["kind"] == 1:
        return corpus.spherocylinder_contact_corpus(c["n"], c["seed"], radius=c["radius"], length=c["length"],
                                                    delta=c["delta"])
    return corpus.sphere_contact_corpus(c["n"], c["seed"], radius=c["radius"], delta=c["delta"])


def golden_decisions(c):
    bits = np.frombuffer(base64.b64decode(c["decisions"]), dtype=np.uint8)
    return np.unpackbits(bits)[: c["n"]]


def snapshot_arrays(s):
    kind = "sphere" if s["kind"] == 0 else "spherocylinder"
    return corpus.random_snapshot(s["n"], s["seed"], np.array(s["box"]), kind=kind, parallel=s["parallel"])


def test_corpus_inputs_are_reproducible(golden):
    for c in golden["corpora"]:
        assert corpus.records_sha256(corpus_records(c)) == c["sha256"], c["name"]
    for s in golden["snapshots"]:
        pos, rot = snapshot_arrays(s)
        assert corpus.records_sha256(np.concatenate([pos.ravel(), rot.ravel()])) == s["sha256"], s["name"]


def test_oracle_decisions_match_reference_fixtures(golden):
    o = oracle_lib.oracle()
    for c in golden["corpora"]:
        dec = o.pair_decisions(c["kind"], c["radius"], c["length"], corpus_records(c))
        ref = golden_decisions(c)
        assert int((dec != ref).sum()) == 0, c["name"]
        assert 0.3 < dec.mean() < 0.7  # near contact: both outcomes occur


def test_oracle_counts_match_reference_fixtures(golden):
    o = oracle_lib.oracle()
    for s in golden["snapshots"]:
        pos, rot = snapshot_arrays(s)
        full, _ = o.snapshot_stats(s["kind"], s["radius"], s["length"], np.array(s["box"]), pos, rot)
        early, _ = o.snapshot_stats(s["kind"], s["radius"], s["length"], np.array(s["box"]), pos, rot, True)
        assert full == s["overlaps"], s["name"]
        assert early == s["overlaps_early_exit"], s["name"]


def test_ulp_corpus_detects_operation_order():
    """A plausible mistake -- letting the compiler fuse multiply-adds -- must change decisions on the ULP-level
    corpus; otherwise that corpus would pin nothing.  The mistake is planted in a CPU build of the oracle."""
    src = os.path.join(oracle_lib.ROOT, "oracle", "rp_oracle.cpp")
    out = "/tmp/liboracle_fma_test.so"
    res = subprocess.run(["g++", "-std=c++17", "-O2", "-march=x86-64-v3", "-ffp-contract=fast", "-fPIC", "-shared",
                          "-o", out, src], capture_output=True)
    if res.returncode != 0:
        pytest.skip("g++ with x86-64-v3 unavailable")
    lib = ctypes.CDLL(out)
    oracle_lib._setup_common(lib)
    fused = oracle_lib.Checker(lib, "orc_")
    rec = corpus.spherocylinder_contact_corpus(50000, 99, delta=2e-15)
    exact = oracle_lib.oracle().pair_decisions(1, 0.5, 5.0, rec)
    assert int((fused.pair_decisions(1, 0.5, 5.0, rec) != exact).sum()) > 0


@pytest.mark.skipif(not oracle_lib.ref_available(), reason="oracle/_ref is built only where the reference is")
def test_oracle_matches_reference_build_on_fresh_corpora():
    o, r = oracle_lib.oracle(), oracle_lib.ref()
    for kind, radius, length, seed, delta in [(1, 0.5, 5.0, 4242, 1e-15), (1, 0.37, 2.2, 4343, 1e-6),
                                              (1, 1.0, 0.0, 4444, 1e-15), (0, 0.5, 0.0, 4545, 1e-16)]:
        if kind == 1:
            rec = corpus.spherocylinder_contact_corpus(30000, seed, radius=radius, length=length, delta=delta)
        else:
            rec = corpus.sphere_contact_corpus(30000, seed, radius=radius, delta=delta)
        assert np.array_equal(o.pair_decisions(kind, radius, length, rec), r.pair_decisions(kind, radius, length, rec))


@pytest.mark.skipif(not oracle_lib.ref_available(), reason="oracle/_ref is built only where the reference is")
def test_oracle_nematic_matches_reference():
    rng = np.random.Generator(np.random.PCG64(8))
    rot = corpus.random_rotations(rng, 500)
    ref = oracle_lib.ref().lib.ref_nematic_p2(500, oracle_lib._d(rot))
    orc = oracle_lib.oracle().lib.orc_nematic_p2(500, oracle_lib._d(rot))
    assert orc == pytest.approx(ref, rel=1e-12, abs=1e-14)
