"""Equation of state of hard spheres at p* = 5 (the plain method's state point: N = 1000 spheres of diameter 1,
rototranslation step 0.1 adapted during thermalisation, log box moves every cycle) against the reference's own runs
(fixture `hs_npt_c1`, 16 runs of the reference's hot path) and against Carnahan-Starling.

Errors: each side's standard error is the spread of its independent run means over sqrt(runs); the test asserts
|device - reference| <= 2 standard errors of the difference.  The block-averaged error of each run (block_error) is
reported beside it: volume fluctuations decorrelate over thousands of cycles here, so within one 3e4-cycle run the
block estimate stays below the scatter of the run means, and only independent runs give a sound error.  The
Carnahan-Starling packing fraction at p* = 5 is 0.39325 ((1 + e + e^2 - e^3) / (1 - e)^3 * 6 e / pi = 5); the
reference's runs themselves lie 0.25 % below it (finite N, slow volume relaxation), so that comparison uses a stated
relative tolerance of 0.5 %."""
import numpy as np
import pytest

import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

HS = ra.Shape("sphere", radius=0.5)
ETA_CS = 0.39325


def test_carnahan_starling_value():
    e = ETA_CS
    z = (1 + e + e * e - e ** 3) / (1 - e) ** 3
    assert abs(z * 6 * e / np.pi - 5.0) < 2e-4


def test_hard_sphere_eos_matches_reference_runs(golden):
    st = next((s for s in golden.get("statistics", []) if s["name"] == "hs_npt_c1"), None)
    if st is None:
        pytest.skip("no reference statistics in the fixture")
    means, block_errs = [], []
    for seed in st["device_seeds"]:
        box, pos, rot = lattice.hard_sphere_lattice(st["n_side"], st["phi0"])
        pk = ra.GpuPacking(HS, box, pos, rot, seed=seed)
        pk.set_environment(st["temperature"], st["pressure"], st["trans_step"], None, box_move=("log", st["box_step"]))
        pk.run_cycles(st["therm"], adapt_steps=True)
        phi = []
        for _ in range(st["avg"] // st["every"]):
            pk.run_cycles(st["every"])
            phi.append(pk.packing_fraction())
        pk.close()
        means.append(float(np.mean(phi)))
        block_errs.append(ra.block_error(phi))
    dev, ref = np.array(means), np.array(st["phi_runs"])
    se_dev, se_ref = dev.std(ddof=1) / np.sqrt(len(dev)), ref.std(ddof=1) / np.sqrt(len(ref))
    se = np.hypot(se_dev, se_ref)
    print(f"hard spheres p* = 5: device {dev.mean():.5f} +- {se_dev:.5f} ({len(dev)} runs), reference "
          f"{ref.mean():.5f} +- {se_ref:.5f} ({len(ref)} runs), difference {(dev.mean() - ref.mean()) / se:+.2f} SE; "
          f"Carnahan-Starling {ETA_CS} ({(dev.mean() - ETA_CS) / se_dev:+.2f} device SE); run spread "
          f"{dev.std(ddof=1):.5f}, mean block error per run {np.mean(block_errs):.5f} (reference "
          f"{np.mean(st['phi_block_err_runs']):.5f})", flush=True)
    assert abs(dev.mean() - ref.mean()) <= 2 * se
    assert abs(dev.mean() / ETA_CS - 1) < 0.005
