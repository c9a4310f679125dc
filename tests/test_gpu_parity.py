"""GPU parity: device pair decisions and total overlap counts against the reference's fixtures and the oracle.
Bit-exact: every decision and every count must be identical."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from test_golden import corpus_records, golden_decisions, snapshot_arrays

pytestmark = pytest.mark.gpu


def shape_of(kind, radius, length):
    return ra.Shape("sphere" if kind == 0 else "spherocylinder", radius=radius, length=length)


def test_pair_decisions_match_reference_fixtures(golden):
    for c in golden["corpora"]:
        rec = corpus_records(c)
        assert corpus.records_sha256(rec) == c["sha256"], c["name"]
        dec = ra.pair_decisions(shape_of(c["kind"], c["radius"], c["length"]), rec)
        mism = int((dec != golden_decisions(c)).sum())
        print(f"{c['name']}: n={c['n']} overlaps={int(dec.sum())} mismatches={mism}", flush=True)
        assert mism == 0, c["name"]


@pytest.mark.parametrize("radius,length,delta,seed", [(0.5, 5.0, 1e-6, 1), (0.5, 5.0, 1e-15, 2), (0.31, 1.7, 1e-15, 3),
                                                      (1.0, 0.0, 1e-15, 4)])
def test_pair_decisions_match_oracle_large_corpus(radius, length, delta, seed):
    rec = corpus.spherocylinder_contact_corpus(400000, seed, radius=radius, length=length, delta=delta)
    gpu = ra.pair_decisions(ra.Shape("spherocylinder", radius, length), rec)
    cpu = oracle_lib.oracle().pair_decisions(1, radius, length, rec)
    assert int((gpu != cpu).sum()) == 0


def test_pair_decisions_empty_and_far():
    assert len(ra.pair_decisions(ra.Shape("spherocylinder", 0.5, 5.0), np.zeros((0, 27)))) == 0
    rec = corpus.spherocylinder_contact_corpus(1000, 9, delta=0.0)
    rec[:, 12:15] += 100.0  # far apart: no overlap
    assert ra.pair_decisions(ra.Shape("spherocylinder", 0.5, 5.0), rec).sum() == 0


def test_total_overlap_counts_match_reference_fixtures(golden):
    for s in golden["snapshots"]:
        pos, rot = snapshot_arrays(s)
        pk = ra.GpuPacking(shape_of(s["kind"], s["radius"], s["length"]), np.array(s["box"]), pos, rot)
        full = pk.count_total_overlaps(False)
        early = pk.count_total_overlaps(True)
        print(f"{s['name']}: n={s['n']} reference={s['overlaps']} gpu={full} early={early}", flush=True)
        assert full == s["overlaps"], s["name"]
        assert early == s["overlaps_early_exit"], s["name"]
        pk.close()


def test_total_overlap_count_matches_oracle_on_dense_triclinic():
    box = np.array([[40.0, 2.0, -1.0], [0.0, 38.0, 1.5], [0.0, 0.0, 41.0]])
    pos, rot = corpus.random_snapshot(30000, 21, box)
    pk = ra.GpuPacking(ra.Shape("spherocylinder", 0.5, 5.0), box, pos, rot)
    expected, _ = oracle_lib.oracle().snapshot_stats(1, 0.5, 5.0, box, pos, rot)
    assert pk.count_total_overlaps() == expected
