"""Polyspheres on the device: total overlap counts over the grid of bead centres equal the reference build's counts
(committed fixtures, full and early exit), pair decisions equal the oracle's bead-pair test, and the packing
fraction uses the summed bead volume."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from test_poly_oracle import records_for

pytestmark = pytest.mark.gpu


def poly_shape(beads):
    c, r = corpus.bead_set(beads)
    return ra.Shape("polysphere", centres=c, radii=r)


def test_device_counts_match_reference_fixtures(golden):
    for s in golden["poly_snapshots"]:
        box = np.array(s["box"])
        pos, rot = corpus.random_snapshot(s["n"], s["seed"], box)
        pk = ra.GpuPacking(poly_shape(s["beads"]), box, pos, rot, seed=1)
        full = pk.count_total_overlaps(early_exit=False)
        early = pk.count_total_overlaps(early_exit=True)
        print(f"{s['name']}: {full} (reference {s['overlaps']}), early {early}", flush=True)
        assert full == s["overlaps"] and early == s["overlaps_early_exit"]
        pk.close()


@pytest.mark.parametrize("beads", ["kinked11", "dimer", "mixed"])
def test_device_pair_decisions_match_oracle(beads):
    centres, radii = corpus.bead_set(beads)
    ext = float(np.sqrt((centres * centres).sum(1).max()))
    rec = records_for(centres, 20000, 11, 2 * ext + 1.5)
    dev = ra.pair_decisions(poly_shape(beads), rec)
    orc = oracle_lib.poly_pair_decisions(centres, radii, rec)
    assert 0 < orc.sum() < len(rec)
    assert np.array_equal(dev, orc)


def test_polysphere_observables_and_validation():
    c, r = corpus.bead_set("kinked11")
    box = np.diag([30.0, 30.0, 30.0])
    pos, rot = corpus.random_snapshot(200, 5, box)
    pk = ra.GpuPacking(poly_shape("kinked11"), box, pos, rot, seed=1)
    obs = pk.observables()
    assert np.isclose(obs["packing_fraction"], 200 * 11 * np.pi / 6 / 27000.0, rtol=1e-14)
    pk.close()
    with pytest.raises(ra.PreconditionError):
        ra.GpuPacking(ra.Shape("polysphere", centres=c, radii=-r), box, pos, rot)
    with pytest.raises(ra.PreconditionError):
        ra.GpuPacking(ra.Shape("polysphere", centres=np.zeros((17, 3)), radii=np.ones(17)), box, pos, rot)
