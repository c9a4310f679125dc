"""Writes tests/golden/golden.json from the reference's own hot path (oracle/_ref/libref.so, compiled from the
reference sources by oracle/Makefile).  Run in the container that holds the reference tree:

    make -C oracle ref oracle && python tests/golden/make_golden.py

The fixture stores only data: generator parameters, SHA-256 of the generated inputs, the reference's decision
bitmaps and the reference's overlap counts / energies for deterministic snapshots.
"""
from __future__ import annotations

import base64
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import corpus  # noqa: E402
import oracle_lib  # noqa: E402

CORPORA = [
    # name, kind, radius, length, n, seed, delta
    ("sc_contact_1e-6", 1, 0.5, 5.0, 100000, 20261202, 1e-6),
    ("sc_contact_ulp", 1, 0.5, 5.0, 200000, 99, 2e-15),
    ("sphere_contact_1e-6", 0, 0.5, 0.0, 100000, 20261203, 1e-6),
    ("sphere_contact_ulp", 0, 0.5, 0.0, 200000, 5, 2e-16),
]

TRICLINIC = np.array([[30.0, 1.5, -0.5], [0.0, 29.0, 0.75], [0.0, 0.0, 31.0]])
CUBIC = np.diag([30.0, 30.0, 30.0])

SNAPSHOTS = [
    # name, kind, radius, length, box, n, seed, parallel
    ("sc_triclinic_dense", 1, 0.5, 5.0, TRICLINIC, 3000, 11, False),
    ("sc_cubic_sparse", 1, 0.5, 5.0, CUBIC, 800, 12, False),
    ("sc_cubic_parallel", 1, 0.5, 5.0, CUBIC, 2000, 13, True),
    ("sc_small_box", 1, 0.5, 2.0, np.diag([14.0, 15.0, 16.0]), 300, 14, False),
    ("sc_thin_box", 1, 0.5, 5.0, np.diag([7.0, 13.0, 25.0]), 150, 17, False),
    ("sphere_triclinic", 0, 0.5, 0.0, TRICLINIC, 20000, 15, False),
    ("sphere_cubic", 0, 0.5, 0.0, np.diag([12.0, 12.0, 12.0]), 1500, 16, False),
]


def corpus_records(kind, radius, length, n, seed, delta):
    if kind == 1:
        return corpus.spherocylinder_contact_corpus(n, seed, radius=radius, length=length, delta=delta)
    return corpus.sphere_contact_corpus(n, seed, radius=radius, delta=delta)


def snapshot_arrays(kind, box, n, seed, parallel):
    return corpus.random_snapshot(n, seed, box, kind="sphere" if kind == 0 else "spherocylinder", parallel=parallel)


def main():
    ref = oracle_lib.ref()
    out = {"corpora": [], "snapshots": []}
    for name, kind, radius, length, n, seed, delta in CORPORA:
        rec = corpus_records(kind, radius, length, n, seed, delta)
        dec = ref.pair_decisions(kind, radius, length, rec)
        out["corpora"].append(dict(name=name, kind=kind, radius=radius, length=length, n=n, seed=seed, delta=delta,
                                   sha256=corpus.records_sha256(rec), overlaps=int(dec.sum()),
                                   decisions=base64.b64encode(np.packbits(dec).tobytes()).decode()))
        print(name, int(dec.sum()), "/", n)
    for name, kind, radius, length, box, n, seed, parallel in SNAPSHOTS:
        pos, rot = snapshot_arrays(kind, box, n, seed, parallel)
        full, _ = ref.snapshot_stats(kind, radius, length, box, pos, rot, early_exit=False)
        early, _ = ref.snapshot_stats(kind, radius, length, box, pos, rot, early_exit=True)
        out["snapshots"].append(dict(name=name, kind=kind, radius=radius, length=length, box=box.tolist(), n=n,
                                     seed=seed, parallel=parallel, sha256=corpus.records_sha256(
                                         np.concatenate([pos.ravel(), rot.ravel()])),
                                     overlaps=full, overlaps_early_exit=early))
        print(name, full, early)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
