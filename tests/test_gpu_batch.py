"""Replicas run side by side (rp_run_cycles_batch): every replica of a small equation-of-state scan ends bit for
bit where its own run_cycles call ends, counters included; a context listed twice is refused."""
import numpy as np
import pytest

import rampack_amd as ra
from rampack_amd import lattice

pytestmark = [pytest.mark.gpu, pytest.mark.skip(
    reason="the side-by-side batch entry point was removed: its first GPU run (this test, then a 32-replica "
           "equation-of-state scan) left the GPU faulted and no cause was found in the code")]

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
REPLICAS = [(1000, 0.5, 11), (2500, 2.0, 12), (1800, 4.0, 13), (4000, 1.0, 14)]


def make(n, pressure, seed):
    box, pos, rot = lattice.spherocylinder_smectic(n, 0.2)
    pk = ra.GpuPacking(SC, box, pos, rot, seed=seed)
    pk.set_environment(1.0, pressure, 0.3, None, box_move=("delta_triclinic", 0.05))
    return pk


def test_batched_replicas_match_separate_runs():
    alone = []
    for n, p, seed in REPLICAS:
        pk = make(n, p, seed)
        c = pk.run_cycles(25, adapt_steps=True)
        c2 = pk.run_cycles(10)
        alone.append((pk.download(), c, c2))
        pk.close()
    pks = [make(n, p, seed) for n, p, seed in REPLICAS]
    c_batch = ra.run_cycles_batch(pks, 25, adapt_steps=True)  # noqa: F821 -- entry point removed, test skipped
    c2_batch = ra.run_cycles_batch(pks, 10)
    for i, pk in enumerate(pks):
        (b, pos, rot), c, c2 = alone[i]
        b2, pos2, rot2 = pk.download()
        print(f"replica {i}: accepted {c2['moves_accepted']} box {c2['box_accepted']}", flush=True)
        assert c_batch[i] == c and c2_batch[i] == c2, i
        assert np.array_equal(b, b2) and np.array_equal(pos, pos2) and np.array_equal(rot, rot2), i
        assert c2["moves_accepted"] > 0
        assert pk.count_total_overlaps() == 0
    with pytest.raises(ra.PreconditionError):
        ra.run_cycles_batch([pks[0], pks[1], pks[0]], 1)
    for pk in pks:
        pk.close()
