"""Writes tests/golden/golden.json from the reference's own hot path (oracle/_ref/libref.so, compiled from the
reference sources by oracle/Makefile).  Run in the container that holds the reference tree:

    make -C oracle ref oracle && python tests/golden/make_golden.py [--large-only] [--poly-only] [--sc-npt-only] [--compound-only] [--c1-only]
        [--wca-only] [--dimer-only]

The fixture stores only data: generator parameters, SHA-256 of the generated inputs, the reference's decision
bitmaps and the reference's overlap counts / energies for deterministic snapshots.
"""
from __future__ import annotations

import base64
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import corpus  # noqa: E402
import oracle_lib  # noqa: E402

CORPORA = [
    # name, kind, radius, length, n, seed, delta
    ("sc_contact_1e-6", 1, 0.5, 5.0, 100000, 20261202, 1e-6),
    ("sc_contact_ulp", 1, 0.5, 5.0, 200000, 99, 2e-15),
    ("sphere_contact_1e-6", 0, 0.5, 0.0, 100000, 20261203, 1e-6),
    ("sphere_contact_ulp", 0, 0.5, 0.0, 200000, 5, 2e-16),
]

# name, kind, radius, length, n, seed, delta, chunk: corpora too large to commit as bitmaps; the fixture keeps the
# SHA-256 of the inputs and of the reference's packed decision bitmap (corpus.chunked_contact_corpus)
LARGE_CORPORA = [
    ("sc_contact_1e7", 1, 0.5, 5.0, 10000000, 20261202, 1e-6, 250000),
]

# name, bead set, box, n, seed: random polysphere snapshots (positions uniform in the box, random orientations);
# the reference counts overlapping bead pairs of different particles on its grid of interaction centres
POLY_SNAPSHOTS = [
    ("kinked11_cubic", "kinked11", np.diag([30.0, 30.0, 30.0]), 600, 31),
    ("kinked11_triclinic", "kinked11", np.array([[26.0, 2.0, -1.5], [0.0, 25.0, 1.25], [0.0, 0.0, 27.0]]), 500, 32),
    ("dimer_triclinic", "dimer", np.array([[20.0, 1.5, -0.5], [0.0, 19.0, 0.75], [0.0, 0.0, 21.0]]), 3000, 33),
    ("mixed_radii_cubic", "mixed", np.diag([18.0, 18.0, 18.0]), 1500, 34),
]


def poly_snapshots(ref):
    out = []
    for name, beads, box, n, seed in POLY_SNAPSHOTS:
        centres, radii = corpus.bead_set(beads)
        pos, rot = corpus.random_snapshot(n, seed, box)
        full, _ = ref.snapshot_stats(2, 0.0, 0.0, box, pos, rot, centres=centres, radii=radii)
        early, _ = ref.snapshot_stats(2, 0.0, 0.0, box, pos, rot, early_exit=True, centres=centres, radii=radii)
        out.append(dict(name=name, beads=beads, box=box.tolist(), n=n, seed=seed,
                        sha256=corpus.records_sha256(np.concatenate([pos.ravel(), rot.ravel()])),
                        overlaps=full, overlaps_early_exit=early))
        print(name, full, early)
    return out


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


CUT = 2 ** (1 / 6)

SOFT_CORPORA = [
    # name, n, seed, eps, sigma, delta: WCA sphere pairs with centre distances within +-delta of the cut-off
    ("wca_sphere_pairs", 100000, 41, 1.0, 1.0, 0.25),
    ("wca_sphere_pairs_scaled", 50000, 42, 3.0, 2.0, 0.5),
]

SOFT_SNAPSHOTS = [
    # name, n_side, spacing, jitter, seed: WCA spheres (epsilon = sigma = 1) on a jittered simple-cubic lattice
    ("wca_sphere_lattice", 12, 1.08, 0.12, 43),
    ("wca_sphere_lattice_dense", 10, 1.0, 0.05, 44),
]

# NpT WCA spheres from a simple-cubic start, delta_v box moves (the reference's cycle loop restated on its Packing)
WCA_NPT = dict(name="wca_sphere_npt", n_side=8, phi0=0.30, temperature=1.0, pressure=2.0, box_step=5.0,
               therm=4000, avg=12000, every=20, seed=1234)


# NpT hard dumbbells (two touching beads of diameter 1) from an aligned orthorhombic start, log isotropic box moves
# (independent runs: samples within one run are correlated far beyond the sampling interval, so the fixture keeps the
# mean over runs and the spread between them)
# The packing fraction still climbs for tens of thousands of cycles after the start (reference means over 16 runs:
# 0.275 after 4e3 thermalisation cycles, 0.288 after 3e4, 0.290 after 6e4, 0.289 after 1.2e5), so the runs thermalise
# for 4e4 cycles.
DIMER_NPT = dict(name="dimer_npt", beads="dimer", n_side=6, spacing=(1.6, 1.6, 2.05), temperature=1.0,
                 pressure=1.5, trans_step=0.2, box_step=0.02, therm=40000, avg=20000, every=20,
                 seeds=[2468, 11, 12, 13, 14, 15, 16, 17] + list(range(200, 224)),
                 device_seeds=[97531, 1, 2, 3, 4, 5, 6, 7])


# NpT hard spherocylinders L/D = 5 at p* = 1 (an isotropic liquid near packing fraction 0.31) from an aligned
# simple-cubic start, delta_triclinic box moves; packing fraction and nematic P2 per independent run
SC_NPT = dict(name="sc_npt_p1", n_side=(10, 10, 3), spacing=(1.6, 1.6, 6.5), temperature=1.0, pressure=1.0,
              trans_step=0.2, rot_step=0.2, box_step=0.01, therm=5000, avg=6000, every=50,
              seeds=list(range(11, 35)))


def sc_run(args):
    """One reference NpT run of the spherocylinder state point; returns its mean packing fraction and P2."""
    st, seed = args
    from rampack_amd import lattice
    ref = oracle_lib.ref()
    box, pos, rot = lattice.spherocylinder_lattice(n_side=st["n_side"], spacing=st["spacing"])
    out = oracle_lib.ref_mc_run(ref, 1, 0.5, 5.0, box, pos, rot, st["temperature"], st["pressure"], st["trans_step"],
                                st["rot_step"], 1, st["box_step"], seed, therm=st["therm"], avg=st["avg"],
                                avg_every=st["every"])
    return float(out["samples"][:, 0].mean()), float(out["samples"][:, 1].mean())


def reference_sc_statistics(ref):
    st = dict(SC_NPT)
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        runs = list(ex.map(sc_run, [(st, seed) for seed in st["seeds"]]))
    st.update(phi_runs=[r[0] for r in runs], p2_runs=[r[1] for r in runs])
    return st


def reference_dimer_statistics(ref):
    st = dict(DIMER_NPT)
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        means = list(ex.map(dimer_run, [(st, seed) for seed in st["seeds"]]))
    st.update(phi_runs=means, phi=float(np.mean(means)), phi_run_sd=float(np.std(means, ddof=1)))
    return st


def dimer_run(args):
    """One reference NpT run of the dumbbell state point; returns its mean packing fraction."""
    st, seed = args
    ref = oracle_lib.ref()
    centres, radii = corpus.bead_set(st["beads"])
    c = np.ascontiguousarray(centres)
    r = np.ascontiguousarray(radii)
    ref.lib.ref_set_beads(len(c), c.ctypes.data_as(oracle_lib._dp), r.ctypes.data_as(oracle_lib._dp))
    box, pos, rot = corpus.dimer_start(st)
    out = oracle_lib.ref_mc_run(ref, 2, 0.0, 0.0, box, pos, rot, st["temperature"], st["pressure"], st["trans_step"],
                                0.0, 2, st["box_step"], seed, therm=st["therm"], avg=st["avg"], avg_every=st["every"])
    return float(out["samples"][:, 0].mean())


# NpT hard spheres (r = 0.5) with a WCA shell (eps = 1, sigma = 1): the reference's CompoundInteraction of SphereTraits
# and RepulsiveLennardJonesInteraction; delta_V box moves; packing fraction and energy per particle per independent run
COMPOUND_NPT = dict(name="compound_sphere_npt", n_side=8, phi0=0.3, temperature=1.0, pressure=2.0, trans_step=0.1,
                    box_step=5.0, therm=3000, avg=9000, every=20, seeds=list(range(300, 316)),
                    device_seeds=list(range(400, 412)))


def compound_run(args):
    """One reference NpT run of the compound-sphere state point: its mean packing fraction and energy per particle."""
    st, seed = args
    from rampack_amd import lattice
    ref = oracle_lib.ref()
    box, pos, rot = lattice.hard_sphere_lattice(st["n_side"], st["phi0"])
    out = oracle_lib.ref_mc_run(ref, 0, 0.5, 0.0, box, pos, rot, st["temperature"], st["pressure"], st["trans_step"],
                                0.0, 3, st["box_step"], seed, therm=st["therm"], avg=st["avg"], avg_every=st["every"],
                                soft=1 | 0x100, eps=1.0, sigma=1.0)
    return float(out["samples"][:, 0].mean()), float(out["samples"][:, 2].mean())


def reference_compound_statistics(ref):
    st = dict(COMPOUND_NPT)
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        runs = list(ex.map(compound_run, [(st, seed) for seed in st["seeds"]]))
    st.update(phi_runs=[r[0] for r in runs], energy_runs=[r[1] for r in runs])
    return st


# The hard-sphere NpT state point of the plain method: N = 1000 spheres of diameter 1 at p* = 5, rototranslation step
# 0.1 adapted during thermalisation, log (isotropic) box moves every cycle.  The equation of state needs an
# equilibrated density: from the 10x10x10 sc lattice at packing fraction 0.30 the small box steps random-walk the
# volume and the packing fraction still climbs after 4e4 cycles, so the runs start from the sc lattice at 0.39 (near
# the Carnahan-Starling value 0.39325), with box step 0.001; 1e4 thermalisation + 3e4 averaging cycles, the packing
# fraction sampled every 10 cycles.  Per independent run: the mean packing fraction and its block-averaged error.
HS_C1 = dict(name="hs_npt_c1", n_side=10, phi0=0.39, temperature=1.0, pressure=5.0, trans_step=0.1, box_step=0.001,
             therm=10000, avg=30000, every=10, seeds=list(range(500, 516)), device_seeds=list(range(600, 608)))


def hs_c1_run(args):
    st, seed = args
    from rampack_amd import lattice
    import rampack_amd as ra
    ref = oracle_lib.ref()
    box, pos, rot = lattice.hard_sphere_lattice(st["n_side"], st["phi0"])
    out = oracle_lib.ref_mc_run(ref, 0, 0.5, 0.0, box, pos, rot, st["temperature"], st["pressure"], st["trans_step"],
                                0.0, 2, st["box_step"], seed, therm=st["therm"], avg=st["avg"], avg_every=st["every"])
    phi = out["samples"][:, 0]
    return float(phi.mean()), ra.block_error(phi)


def reference_hs_c1_statistics(ref):
    st = dict(HS_C1)
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        runs = list(ex.map(hs_c1_run, [(st, seed) for seed in st["seeds"]]))
    st.update(phi_runs=[r[0] for r in runs], phi_block_err_runs=[r[1] for r in runs])
    return st


def reference_statistics(ref):
    from rampack_amd import lattice
    st = dict(WCA_NPT)
    box, pos, rot = lattice.hard_sphere_lattice(st["n_side"], st["phi0"])
    r = oracle_lib.ref_mc_run(ref, 0, 0.5, 0.0, box, pos, rot, st["temperature"], st["pressure"], 0.1, 0.0, 3,
                              st["box_step"], st["seed"], therm=st["therm"], avg=st["avg"], avg_every=st["every"],
                              soft=1, eps=1.0, sigma=1.0)
    s = r["samples"]
    rho = s[:, 0] / (np.pi / 6)   # packing fraction / v0 with the shape radius 0.5
    e = s[:, 2]
    import rampack_amd as ra
    st.update(rho=float(rho.mean()), rho_err=ra.block_error(rho), energy=float(e.mean()), energy_err=ra.block_error(e),
              samples=int(len(s)), errors="block-averaged")
    return st


def corpus_records(kind, radius, length, n, seed, delta):
    if kind == 1:
        return corpus.spherocylinder_contact_corpus(n, seed, radius=radius, length=length, delta=delta)
    return corpus.sphere_contact_corpus(n, seed, radius=radius, delta=delta)


def snapshot_arrays(kind, box, n, seed, parallel):
    return corpus.random_snapshot(n, seed, box, kind="sphere" if kind == 0 else "spherocylinder", parallel=parallel)


def large_corpora(ref):
    out = []
    for name, kind, radius, length, n, seed, delta, chunk in LARGE_CORPORA:
        rec = corpus.chunked_contact_corpus(n, seed, radius=radius, length=length, delta=delta, chunk=chunk)
        dec = ref.pair_decisions(kind, radius, length, rec)
        out.append(dict(name=name, kind=kind, radius=radius, length=length, n=n, seed=seed, delta=delta, chunk=chunk,
                        sha256=corpus.records_sha256(rec), overlaps=int(dec.sum()),
                        decisions_sha256=hashlib.sha256(np.packbits(dec).tobytes()).hexdigest()))
        print(name, int(dec.sum()), "/", n)
    return out


def main():
    ref = oracle_lib.ref()
    path = os.path.join(HERE, "golden.json")
    if "--sc-npt-only" in sys.argv:
        with open(path) as f:
            out = json.load(f)
        out["statistics"] = [s for s in out["statistics"] if s["name"] != SC_NPT["name"]]
        out["statistics"].append(reference_sc_statistics(ref))
        print(out["statistics"][-1])
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        return
    if "--wca-only" in sys.argv or "--dimer-only" in sys.argv:
        with open(path) as f:
            out = json.load(f)
        name = WCA_NPT["name"] if "--wca-only" in sys.argv else DIMER_NPT["name"]
        out["statistics"] = [s for s in out["statistics"] if s["name"] != name]
        out["statistics"].append(reference_statistics(ref) if "--wca-only" in sys.argv else
                                 reference_dimer_statistics(ref))
        print(out["statistics"][-1])
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        return
    if "--c1-only" in sys.argv:
        with open(path) as f:
            out = json.load(f)
        out["statistics"] = [s for s in out["statistics"] if s["name"] != HS_C1["name"]]
        out["statistics"].append(reference_hs_c1_statistics(ref))
        print(out["statistics"][-1])
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        return
    if "--compound-only" in sys.argv:
        with open(path) as f:
            out = json.load(f)
        out["statistics"] = [s for s in out["statistics"] if s["name"] != COMPOUND_NPT["name"]]
        out["statistics"].append(reference_compound_statistics(ref))
        print(out["statistics"][-1])
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        return
    if "--large-only" in sys.argv or "--poly-only" in sys.argv:
        with open(path) as f:
            out = json.load(f)
        if "--large-only" in sys.argv:
            out["large_corpora"] = large_corpora(ref)
        if "--poly-only" in sys.argv:
            out["poly_snapshots"] = poly_snapshots(ref)
            out["statistics"] = [s for s in out["statistics"] if s["name"] != DIMER_NPT["name"]]
            out["statistics"].append(reference_dimer_statistics(ref))
            print(out["statistics"][-1])
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        return
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
    out["soft_corpora"] = []
    for name, n, seed, eps, sigma, delta in SOFT_CORPORA:
        rec = corpus.sphere_contact_corpus(n, seed, radius=sigma * CUT / 2, delta=delta)
        e = ref.pair_energies(0, 0.5, 0.0, 1, eps, sigma, rec)
        out["soft_corpora"].append(dict(name=name, n=n, seed=seed, eps=eps, sigma=sigma, delta=delta,
                                        sha256=corpus.records_sha256(rec), energies_sha256=corpus.records_sha256(e),
                                        nonzero=int((e != 0).sum())))
        print(name, int((e != 0).sum()), "/", n)
    out["soft_snapshots"] = []
    for name, n_side, spacing, jitter, seed in SOFT_SNAPSHOTS:
        box, pos, rot = corpus.jittered_lattice(n_side, spacing, jitter, seed)
        _, energy = ref.snapshot_stats(0, 0.5, 0.0, box, pos, rot, soft=1, eps=1.0, sigma=1.0)
        out["soft_snapshots"].append(dict(name=name, n_side=n_side, spacing=spacing, jitter=jitter, seed=seed,
                                          sha256=corpus.records_sha256(np.concatenate([pos.ravel(), rot.ravel()])),
                                          energy=energy))
        print(name, energy)
    out["large_corpora"] = large_corpora(ref)
    out["poly_snapshots"] = poly_snapshots(ref)
    out["statistics"] = [reference_statistics(ref), reference_dimer_statistics(ref), reference_sc_statistics(ref),
                         reference_compound_statistics(ref)]
    print(out["statistics"])
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
