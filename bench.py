"""Benchmark: MC trial moves/s of the GPU sweep on N = 1e5 hard spherocylinders (L/D = 5) at p* = 2.0.

One step = one MC cycle of rampack's Simulation::performCycle (src/core/Simulation.cpp:325-363): N trial
rototranslations (eight checkerboard colour passes) plus one delta_triclinic box move, the same unit the reference
reports as "Cycles per second" (src/frontend/modes/CasinoMode.cpp:405-412).  Trial moves/s = N * cycles / time.

Input T: AA-stacked layers of parallel rods (packing fraction 0.45) relaxed at p* = 2 by a fixed number of cycles
with step adaptation (--equilibrate, default 3000), seeded, so every run starts the timed region from the same state
bit for bit (its SHA-256 is printed) whatever --warmup is; the warm-up and timed cycles then run with the adapted
steps frozen.  The state is generated on the device rather than loaded from a file (a RAMSNAP file of 1e5 rods is
about 20 MB).  Inputs are resident in HBM before timing.

Roofline: the dominant kernel is the sweep (k_sweep).  Its algorithmic traffic per trial move is
(27 + a) * 96 B (the reference's 27-cell stencil of FP64 particle records plus the write-back of accepted moves, a =
acceptance rate); achieved GB/s = that figure times the moves of one launch divided by the launch's average duration,
measured with HIP events on the launch stream over the timed cycles (events are read once, after them).  traffic is
the HBM bytes one sweep launch moves, from rocprofv3 FETCH_SIZE and WRITE_SIZE counters on this workload scaled by
the same counters on a 256 MiB device copy (SWEEP_TRAFFIC below).

CPU baseline: the reference's own hot path (oracle/_ref/libref_perf*.so: Packing, NeighbourGrid,
SpherocylinderTraits, RototranslationSampler, TriclinicDeltaScaler and DomainDecomposition compiled from the
reference sources with -O3, for x86-64-v4 where the host has AVX-512, else x86-64-v3) driven by the restated cycle
loop with OpenMP domain decomposition over every core this process may run on, on the same state, timed for a
bounded number of cycles.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC trial moves/s @N=1e5 spherocyl.; plain FP64 vs guarded-FP32+learned order"
HBM_PEAK_GBS = 8000.0
RECORD_BYTES = 96  # FP64 position + FP64 3x3 orientation, the reference's particle record

# HBM traffic of one k_sweep launch on this workload (profiles/k_sweep_hard_spherocylinders_N1e5_hbm_traffic.txt):
# rocprofv3 FETCH_SIZE / WRITE_SIZE means in KiB per dispatch, and the same counters on a 256 MiB device-to-device
# copy, which reads and writes exactly 256 MiB.
SWEEP_TRAFFIC = dict(fetch_kib=11161.49, write_kib=1402.26, copy_fetch_kib=131084.92, copy_write_kib=262144.0,
                     copy_bytes=268435456)


def sweep_traffic_bytes() -> float:
    t = SWEEP_TRAFFIC
    fetch = t["fetch_kib"] * t["copy_bytes"] / t["copy_fetch_kib"]
    write = t["write_kib"] * t["copy_bytes"] / t["copy_write_kib"]
    return fetch + write


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def near_cubic_divisions(cores: int):
    best = (cores, 1, 1)
    for a in range(1, cores + 1):
        if cores % a:
            continue
        for b in range(1, cores // a + 1):
            if (cores // a) % b:
                continue
            c = cores // a // b
            cand = tuple(sorted((a, b, c), reverse=True))
            if max(cand) - min(cand) < max(best) - min(best):
                best = cand
    return best


def cpu_baseline(box, pos, rot, trans_step, rot_step, box_step, cycles, threads, seed, wl):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib
    if not os.path.exists(oracle_lib.REF_PERF_PATH):
        return None
    ref = oracle_lib.ref(perf=True)
    if wl["beads"] is not None:
        c, r = (np.ascontiguousarray(x) for x in wl["beads"])
        ref.lib.ref_set_beads(len(c), c.ctypes.data_as(oracle_lib._dp), r.ctypes.data_as(oracle_lib._dp))
    dd = wl["divisions"](threads)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    res = oracle_lib.ref_mc_run(ref, wl["kind"], wl["radius"], wl["length"], box, pos, rot, 1.0, wl["pressure"],
                                trans_step, rot_step, 1, box_step, seed, dd=dd, threads=threads, therm=0, avg=cycles,
                                avg_every=0)
    rate = len(pos) * cycles / res["seconds"]
    build = "x86-64-v4" if "v4" in os.path.basename(oracle_lib.REF_PERF_V4_PATH) and \
        os.path.exists(oracle_lib.REF_PERF_V4_PATH) and oracle_lib.host_has_avx512() else "x86-64-v3"
    return {"value": rate, "unit": "trial moves/s", "cores": threads, "kind": "reference",
            "cpu_model": cpu_model(), "build": f"-O3 -march={build} -fopenmp",
            "sample": f"{cycles} cycles of the same N={len(pos)} state (the GPU's state after its timed cycles), "
                      f"NpT p*={wl['pressure']}, rototranslation + delta_triclinic, OpenMP domain divisions "
                      f"{list(dd)}, {res['seconds']:.1f} s"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((line.split(":", 1)[1].strip() for line in f if line.startswith("model name")), "unknown")
    except OSError:
        return "unknown"


def host_cores() -> int:
    """The cores this process may run on (its affinity mask), not the machine's count."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def state_sha256(box, pos, rot) -> str:
    import hashlib
    import numpy as np
    h = hashlib.sha256()
    for a in (box, pos, rot):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def workload(name: str, n: int):
    """The packing, its environment and what the roofline and CPU legs need for one bench workload."""
    import numpy as np
    import rampack_amd as ra
    from rampack_amd import lattice
    if name == "spherocylinders":
        box, pos, rot = lattice.spherocylinder_smectic(n, 0.45)
        return dict(shape=ra.Shape("spherocylinder", radius=0.5, length=5.0), state=(box, pos, rot), pressure=2.0,
                    trans_step=0.1, box_step=0.002, kind=1, radius=0.5, length=5.0, beads=None, centres=1,
                    divisions=near_cubic_divisions,
                    text=f"hard spherocylinders L/D=5 (r=0.5, L=5), N={len(pos)}, NpT p*=2.0 T=1, rototranslation "
                         f"+ delta_triclinic box move per cycle, plain FP64 checkerboard sweep")
    # rigid chains of 11 touching beads of diameter 1 with a 30 degree kink at the middle bead, aligned on a
    # 37 x 37 x 37 lattice (N = 50 653, phi = 0.425), NpT p* = 3
    c, r = lattice.kinked_chain(11, 1.0, 30.0)
    box, pos = lattice.sc_positions((37, 37, 37), (1.1, 1.1, 11.2))
    rot = np.tile(np.eye(3).reshape(1, 9), (len(pos), 1))
    return dict(shape=ra.Shape("polysphere", centres=c, radii=r), state=(box, pos, rot), pressure=3.0,
                trans_step=0.1, box_step=0.01, kind=2, radius=0.0, length=0.0, beads=(c, r), centres=len(c),
                divisions=lambda t: (2, 2, max(1, t // 4)),
                text=f"hard polyspheres: rigid chains of 11 touching beads (d=1) kinked 30 degrees at bead 5, "
                     f"N={len(pos)} from an aligned lattice at phi=0.425 in a 40.7 x 40.7 x 414 box, NpT p*=3.0 T=1, "
                     f"rototranslation + delta_triclinic box move per cycle, plain FP64 checkerboard sweep")


def copy_probe(nbytes: int):
    """A plain device-to-device copy of nbytes (the runtime's copy kernel), used to scale the FETCH_SIZE /
    WRITE_SIZE counters of the sweep against a kernel of known traffic."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy probe: 3 x {nbytes} B device-to-device")


def timed_leg(ra, wl, input_t, eq, precision, args) -> dict:
    """W untimed then K timed cycles from input T with the adapted steps frozen, at one precision of the sweep's
    stencil scan; the packing stays open for the CPU leg."""
    box, pos, rot = input_t
    name = {ra.PREC_FP64: "fp64", ra.PREC_GUARDED_FP32: "guarded_fp32"}[precision]
    pk = ra.GpuPacking(wl["shape"], box, pos, rot, seed=args.seed + 1, precision=precision)
    pk.set_environment(1.0, wl["pressure"], eq["trans_step"], eq["rot_step"],
                       box_move=("delta_triclinic", eq["box_step"]))
    pk.run_cycles(args.warmup)
    phi_start = pk.packing_fraction()
    before = pk.run_cycles(0)
    g0 = pk.guard_stats()
    # HIP events on the context's stream around one sweep launch per timed cycle (the colour pass whose index is
    # the cycle number modulo 8), read once after the timed cycles; events around every launch would add about 8 %
    # to the cycle time
    pk.set_timing("sampled")
    t0 = time.perf_counter()
    after = pk.run_cycles(args.steps)  # returns after the device finished (counters are read back)
    elapsed = time.perf_counter() - t0
    tm = pk.timing()
    pk.set_timing(False)
    g1 = pk.guard_stats()
    # per-kernel diagnostics (rebuild, overlap scan, proposals) from a short pass with every launch timed
    diag_cycles = min(100, args.steps)
    pk.set_timing(True)
    pk.run_cycles(diag_cycles)
    td = pk.timing()
    pk.set_timing(False)
    moves = after["moves_attempted"] - before["moves_attempted"]
    accepted = after["moves_accepted"] - before["moves_accepted"]
    rate = moves / elapsed
    phi_end = pk.packing_fraction()
    avg_sweep_us = tm["sweep_ms"] / max(tm["sweep_launches"], 1) * 1e3
    log(f"[{name}] timed {args.steps} cycles in {elapsed:.3f} s: {rate:.4g} trial moves/s, acceptance "
        f"{accepted / moves:.3f}, phi {phi_end:.4f}, box moves {after['box_accepted'] - before['box_accepted']}/"
        f"{after['box_attempted'] - before['box_attempted']}, unstaged blocks {after['unstaged_blocks']}")
    log(f"[{name}] launch shape: colour period {tm['sweep_period']}, {tm['sweep_waves']} wavefronts per cell, "
        f"{tm['sweep_cap']} staged entries, LDS limit {tm['lds_limit']} B; sweep (sampled) {tm['sweep_launches']} "
        f"launches, avg {avg_sweep_us:.1f} us")
    log(f"[{name}] diagnostic pass of {diag_cycles} cycles, every launch timed: sweep avg "
        f"{td['sweep_ms'] / max(td['sweep_launches'], 1) * 1e3:.1f} us x {td['sweep_launches']}; proposals avg "
        f"{td['proposal_ms'] / max(td['proposal_launches'], 1) * 1e3:.1f} us; rebuild avg "
        f"{td['rebuild_ms'] / max(td['rebuild_launches'], 1) * 1e3:.1f} us x {td['rebuild_launches']}; scan avg "
        f"{td['scan_ms'] / max(td['scan_launches'], 1) * 1e3:.1f} us x {td['scan_launches']}")
    out = dict(rate=rate, elapsed=elapsed, moves=moves, acceptance=accepted / moves, timing=tm, counters=after,
               packing=pk, phi_start=phi_start, phi_end=phi_end, sweep_us=avg_sweep_us)
    if precision != ra.PREC_FP64:
        pairs = g1["pairs"] - g0["pairs"]
        guard = g1["fp64_fallback"] - g0["fp64_fallback"]
        out["guard"] = {"fp32_pairs": pairs, "guard_band_pairs": guard, "guard_band_rate": guard / max(pairs, 1)}
        log(f"[{name}] FP32 pairs {pairs}, guard band {guard} ({guard / max(pairs, 1):.2e})")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--equilibrate", type=int, default=3000,
                    help="cycles with step adaptation that turn the lattice into input T (not timed, seeded)")
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-cycles", type=int, default=-1, help="-1: sized to ~20 s of CPU work; 0: skip")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", choices=["spherocylinders", "kinked-chains"], default="spherocylinders",
                    help="kinked-chains: the polysphere sweep on 50 653 kinked 11-bead chains (not the headline "
                         "metric's workload)")
    ap.add_argument("--precision", choices=["both", "fp64", "guarded"], default="both",
                    help="stencil scan of the timed legs: plain FP64, guarded FP32, or both on the same trajectory")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the sweep shards by spatial domains across "
              f"GPUs by design but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import rampack_amd as ra

    wl = workload(args.workload, args.n)
    box, pos, rot = wl["state"]
    pk = ra.GpuPacking(wl["shape"], box, pos, rot, seed=args.seed)
    pk.set_environment(1.0, wl["pressure"], wl["trans_step"], None, box_move=("delta_triclinic", wl["box_step"]))
    t0 = time.perf_counter()
    eq = pk.run_cycles(args.equilibrate, adapt_steps=True)
    input_t = pk.download()
    sha = state_sha256(*input_t)
    log(f"input T: {args.equilibrate} adapting cycles in {time.perf_counter() - t0:.2f} s: "
        f"phi={pk.packing_fraction():.4f} steps trans={eq['trans_step']:.4f} rot={eq['rot_step']:.4f} "
        f"box={eq['box_step']:.5f}, box moves {eq['box_accepted']}/{eq['box_attempted']}, sha256 {sha[:16]}")
    pk.close()

    # the legs run the same trajectory from input T (same seed and frozen steps; the guarded scan makes the FP64
    # scan's decisions), so they time the same work
    precisions = [("fp64", ra.PREC_FP64)]
    if wl["kind"] in (0, 1) and args.precision in ("both", "guarded"):
        precisions.append(("guarded_fp32", ra.PREC_GUARDED_FP32))
    if args.precision == "guarded":
        precisions = precisions[1:]
    legs = {}
    for name, prec in precisions:
        legs[name] = timed_leg(ra, wl, input_t, eq, prec, args)
    best = max(legs, key=lambda k: legs[k]["rate"])
    leg = legs[best]
    rate, elapsed, moves, acc_rate = leg["rate"], leg["elapsed"], leg["moves"], leg["acceptance"]
    tm, tb, pk = leg["timing"], leg["counters"], leg["packing"]
    phi_start, phi_end = leg["phi_start"], leg["phi_end"]
    avg_sweep_ms = tm["sweep_ms"] / max(tm["sweep_launches"], 1)
    # every cycle launches one sweep per colour (period^3 colours; all hold cells at this size)
    moves_per_launch = moves / (args.steps * tm["sweep_period"] ** 3)
    # algorithmic bytes per trial move: the reference's 27-cell stencil of particle records, plus the write-back of
    # accepted moves; with nc > 1 centres, nc (24 B centre + 72 B orientation) per stencil entry and nc 48 B
    # bead records written back
    nc = wl["centres"]
    if nc == 1:
        bytes_per_move = (27 + acc_rate) * RECORD_BYTES
    else:
        bytes_per_move = nc * 27 * (24 + 72) + acc_rate * (RECORD_BYTES + nc * 48)
    bytes_per_launch = bytes_per_move * moves_per_launch
    achieved = bytes_per_launch / (avg_sweep_ms * 1e-3) / 1e9

    cpu = None
    if args.cpu_cycles != 0:
        b, p, r = pk.download()
        threads = args.cpu_threads or host_cores()
        cycles = args.cpu_cycles
        if cycles < 0:
            probe = cpu_baseline(b, p, r, tb["trans_step"], tb["rot_step"], tb["box_step"], 2, threads, args.seed,
                                 wl)
            if probe is not None:
                per_cycle = len(p) / probe["value"]
                cycles = max(3, min(400, int(20.0 / per_cycle)))
        if cycles > 0:
            cpu = cpu_baseline(b, p, r, tb["trans_step"], tb["rot_step"], tb["box_step"], cycles, threads, args.seed,
                               wl)
        if cpu:
            log(f"cpu baseline: {cpu['value']:.4g} trial moves/s on {cpu['cores']} threads ({cpu['sample']})")

    out = {
        "metric": METRIC if args.workload == "spherocylinders" else
        "MC trial moves/s, kinked 11-bead polysphere chains N=50653 (not the headline workload)",
        "value": rate,
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f64 (f32 centre distances with FP64 guard)",
        "data": "synthetic",
        "config": {
            "workload": wl["text"],
            "n_particles": len(pos),
            "pressure": wl["pressure"],
            "input_T": {"lattice": "AA smectic, packing fraction 0.45", "equilibration_cycles": args.equilibrate,
                        "seed": args.seed, "sha256": sha, "trans_step": eq["trans_step"], "rot_step": eq["rot_step"],
                        "box_step": eq["box_step"]},
            "packing_fraction_timed_start": round(phi_start, 5),
            "packing_fraction_timed_end": round(phi_end, 5),
            "acceptance": round(acc_rate, 4),
            "precision": best,
            "legs": {k: {"trial_moves_per_s": v["rate"], "ms_per_step": v["elapsed"] / args.steps * 1e3,
                         "k_sweep_avg_us": v["sweep_us"], **({"guard": v["guard"]} if "guard" in v else {})}
                     for k, v in legs.items()},
            "guarded_over_fp64": (legs["guarded_fp32"]["rate"] / legs["fp64"]["rate"]
                                  if "guarded_fp32" in legs and "fp64" in legs else None),
            "legs_note": "same seed, input T and frozen steps: both legs run the same trajectory (the guarded "
                         "scan makes the FP64 scan's decisions); value is the faster leg",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": sweep_traffic_bytes() if args.workload == "spherocylinders" and args.n == 100000 else None,
            "kernel": "k_sweep",
            "avg_launch_us": avg_sweep_ms * 1e3,
            "timed_launches": tm["sweep_launches"],
            "timing": f"HIP events carried by one k_sweep dispatch per timed cycle (colour pass = cycle mod "
                      f"{tm['sweep_period'] ** 3}) on the context's stream",
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
    }
    for v in legs.values():
        v["packing"].close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
