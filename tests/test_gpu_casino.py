"""The casino driver end to end on the device: a PYON input with an overlap relaxation of randomly oriented rods
followed by an NpT integration writes the reference's output files (averages, observable snapshots, RAMSNAP,
RAMTRJ), the final snapshot has no overlap by the oracle's count, and the trajectory holds every snapshot."""
import os

import numpy as np
import pytest

import oracle_lib
import rampack_amd as ra
from rampack_amd import casino, pyon

pytestmark = pytest.mark.gpu

INPUT = '''
rampack(
    version = "1.2.0",
    arrangement = lattice(cell=sc, cell_dim=[2.0, 2.0, 4.5], n_cells=[6, 6, 3],
                          transformations=[randomize_rotation(seed=7)]),
    shape = spherocylinder(l=3, r=0.5),
    seed = 1234,
    temperature = 1,
    pressure = 0.5,
    move_types = [translation(step=0.3), rotation(step=0.5)],
    box_move_type = log("isotropic", 0.01),
    domain_divisions = [2, 2, 1],
    runs = [
        overlap_relaxation(run_name="relax", snapshot_every=50, output_last_snapshot=[ramsnap("relaxed.ramsnap")]),
        integration(
            run_name = "gas",
            thermalization_cycles = 300,
            averaging_cycles = 600,
            snapshot_every = 50,
            averaging_every = 50,
            box_move_type = delta_triclinic(0.01),
            output_last_snapshot = [ramsnap("gas.ramsnap")],
            record_trajectory = [ramtrj("gas.ramtrj")],
            averages_out = "out.txt",
            observables = [number_density, scoped(packing_fraction, averaging=True), compressibility_factor,
                           nematic_order],
            observables_out = "obs_gas.txt"
        )
    ]
)
'''


def test_casino_runs_relaxation_and_integration(tmp_path):
    (tmp_path / "in.pyon").write_text(INPUT)
    logs = []
    casino.Casino(pyon.parse(INPUT), str(tmp_path), logs.append).run()
    print("\n".join(logs[-4:]), flush=True)
    box, pos, rot, _ = ra.ramsnap_read(str(tmp_path / "relaxed.ramsnap"))
    assert len(pos) == 108
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 3.0, box, pos, rot)[0] == 0
    box, pos, rot, _ = ra.ramsnap_read(str(tmp_path / "gas.ramsnap"))
    assert oracle_lib.oracle().snapshot_stats(1, 0.5, 3.0, box, pos, rot)[0] == 0
    assert box[0, 1] != 0 or box[1, 2] != 0  # delta_triclinic moved the off-diagonal entries
    avg = (tmp_path / "out.txt").read_text().splitlines()
    assert avg[0].split() == ["temperature", "pressure", "rho", "drho", "theta", "dtheta", "Z", "dZ", "P2", "dP2"]
    row = [float(x) for x in avg[1].split()]
    assert row[:2] == [1.0, 0.5] and len(row) == 10
    v0 = np.pi * 0.25 * 3 + 4 / 3 * np.pi * 0.125
    assert row[4] == pytest.approx(row[2] * v0, rel=1e-9)
    obs = (tmp_path / "obs_gas.txt").read_text().splitlines()
    assert obs[0].split() == ["cycle", "rho", "Z", "P2"]  # packing_fraction is scoped to averaging only
    assert [int(r.split()[0]) for r in obs[1:]] == list(range(50, 901, 50))
    with ra.RamtrjPlayer(str(tmp_path / "gas.ramtrj")) as pl:
        assert pl.get_total_cycles() == 900 and pl.get_cycle_step() == 50
        n = 0
        while pl.has_next():
            pl.next_snapshot()
            n += 1
    assert n == 18
    assert casino.main(["-q", "-i", str(tmp_path / "in.pyon")]) == 0
