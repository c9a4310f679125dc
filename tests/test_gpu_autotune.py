"""The launch-shape autotuner (rp_autotune).  Every candidate launch shape makes the same decisions, so tuning over
k cycles per shape is rp_run_cycles over all those cycles (no step adaptation) bit for bit, and the run goes on as
it would have without tuning; the shape kept is the fastest measured, and the guarded FP32 precision keeps its one
launch shape."""
import numpy as np
import pytest

import rampack_amd as ra
from rampack_amd import lattice
from test_gpu_replay import rod_start

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
HS = ra.Shape("sphere", radius=0.5)
SHAPE_COUNTERS = ("pair_tests", "unstaged_blocks")  # depend on the launch shape, not on the decisions


def start(kind):
    if kind == "sphere":
        box, pos, rot = lattice.hard_sphere_lattice(12, 0.45)
        return HS, box, pos, rot, (1.0, 5.0, 0.1, None, ("log", 0.001))
    box, pos, rot = rod_start()
    return SC, box, pos, rot, (1.0, 2.0, 0.2, None, ("delta_triclinic", 0.002))


@pytest.mark.parametrize("kind", ["sphere", "spherocylinder"])
def test_tuned_run_ends_where_an_untuned_run_ends(kind):
    shape, box, pos, rot, env = start(kind)
    runs = []
    for tune in (True, False):
        pk = ra.GpuPacking(shape, box, pos, rot, seed=9)
        pk.set_environment(*env[:4], box_move=env[4])
        pk.run_cycles(20, adapt_steps=True)
        if tune:
            res = pk.autotune(4)
            times = [c["sweep_us_per_cycle"] for c in res["candidates"]]
            print(f"{kind}: " + ", ".join(f"{c['waves']} {c['schedule']} {c['sweep_us_per_cycle']:.1f} us"
                                          for c in res["candidates"]) + f"; kept {res['chosen']}", flush=True)
            assert len(times) == 6 and all(t > 0 for t in times)
            assert res["chosen"] == int(np.argmin(times))
            c = pk.run_cycles(10)
        else:
            pk.run_cycles(6 * 4)
            c = pk.run_cycles(10)
        runs.append((pk.download(), {k: v for k, v in c.items() if k not in SHAPE_COUNTERS}))
        pk.close()
    (s0, c0), (s1, c1) = runs
    assert all(np.array_equal(a, b) for a, b in zip(s0, s1))
    assert c0 == c1


def test_guarded_precision_keeps_its_launch():
    box, pos, rot = rod_start()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=4, precision=ra.PREC_GUARDED_FP32)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
    res = pk.autotune(3)
    assert [(c["waves"], c["schedule"]) for c in res["candidates"]] == [(8, "grouped")] and res["chosen"] == 0
    pk.run_cycles(5)
    assert pk.guard_stats()["pairs"] > 0
    pk.close()
