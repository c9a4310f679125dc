"""ctypes access to the test-only checkers: oracle/lib/liboracle.so (the CPU restatement) and, where it was built,
oracle/_ref/libref.so (the reference's own hot path compiled from its sources)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref.so")
REF_PERF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_perf.so")
REF_PERF_V4_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_perf_v4.so")


def host_has_avx512() -> bool:
    """x86-64-v4 (AVX-512 F, BW, CD, DQ, VL) on this host's CPU."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((line.split(":", 1)[1].split() for line in f if line.startswith("flags")), [])
    except OSError:
        return False
    return all(x in flags for x in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))

_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def _setup_common(lib):
    lib.orc_pair_decisions.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_size_t, _dp, _u8p]
    lib.orc_snapshot_stats.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_size_t, _dp, _dp,
                                       ctypes.c_int, _u64p, _dp]
    lib.orc_count_brute.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp,
                                    ctypes.c_size_t, _dp, _dp]
    lib.orc_count_brute.restype = ctypes.c_uint64
    lib.orc_pair_energies.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_size_t, _dp, _dp]
    lib.orc_fix_rotation.argtypes = [ctypes.c_size_t, _dp]
    lib.orc_fix_rotation.restype = None
    lib.orc_segment_distance2.argtypes = [ctypes.c_size_t, _dp, _dp]
    lib.orc_segment_distance2.restype = None
    lib.orc_box_info.argtypes = [_dp, _dp, _dp, _dp]
    lib.orc_box_info.restype = None
    lib.orc_pbc.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_pbc.restype = None
    lib.orc_is_box_upscaled.argtypes = [_dp, _dp]
    lib.orc_translate_wrapped.argtypes = [_dp, ctypes.c_size_t, _dp, _dp, _dp]
    lib.orc_translate_wrapped.restype = None
    lib.orc_poly_pair_decisions.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_size_t, _dp, _u8p]
    lib.orc_rescale.argtypes = [_dp, _dp, ctypes.c_size_t, _dp, _dp]
    lib.orc_rescale.restype = None
    lib.orc_rotation.argtypes = [_dp, ctypes.c_double, _dp]
    lib.orc_rotation.restype = None
    lib.orc_rotation_sc.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_rotation_sc.restype = None
    lib.orc_rotate.argtypes = [ctypes.c_size_t, _dp, _dp, _dp]
    lib.orc_rotate.restype = None
    _fp = ctypes.POINTER(ctypes.c_float)
    lib.orc_mlp_loss_grad.argtypes = [_fp, ctypes.c_size_t, _fp, _fp, _dp, _dp]
    lib.orc_mlp_loss_grad.restype = ctypes.c_double
    lib.orc_sincos.argtypes = [ctypes.c_size_t, _dp, _dp]
    lib.orc_sincos.restype = None
    lib.orc_nematic_p2.argtypes = [ctypes.c_size_t, _dp]
    lib.orc_nematic_p2.restype = ctypes.c_double
    lib.orc_pair_guarded.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_size_t, _dp, _u8p,
                                     _u8p, ctypes.c_double]
    _ip = ctypes.POINTER(ctypes.c_int)
    lib.orc_replay_sweep.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_size_t, _dp, _dp,
                                     _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_size_t, _ip, _dp, _dp, _u8p, _dp, _dp]


class Checker:
    """Common interface of the restatement (prefix orc_) and the reference build (prefix ref_)."""

    def __init__(self, lib, prefix):
        self.lib = lib
        self.prefix = prefix

    def pair_decisions(self, kind: int, radius: float, length: float, rec: np.ndarray) -> np.ndarray:
        rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 27)
        out = np.zeros(len(rec), dtype=np.uint8)
        f = getattr(self.lib, self.prefix + "pair_decisions")
        f(kind, radius, length, len(rec), rec.ctypes.data_as(_dp), out.ctypes.data_as(_u8p))
        return out

    def pair_energies(self, kind: int, radius: float, length: float, soft: int, eps: float, sigma: float,
                      rec: np.ndarray) -> np.ndarray:
        rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 27)
        out = np.zeros(len(rec))
        f = getattr(self.lib, self.prefix + "pair_energies")
        rc = f(kind, radius, length, soft, eps, sigma, len(rec), rec.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}pair_energies does not support soft={soft} on kind={kind}")
        return out

    def snapshot_stats(self, kind, radius, length, box, pos, rot, early_exit=False, soft=0, eps=1.0, sigma=1.0,
                       centres=None, radii=None):
        box = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rot = np.ascontiguousarray(rot, dtype=np.float64)
        nc = 0 if centres is None else len(centres)
        cen = np.zeros(3) if centres is None else np.ascontiguousarray(centres, dtype=np.float64)
        rad = np.zeros(1) if radii is None else np.ascontiguousarray(radii, dtype=np.float64)
        ov = ctypes.c_uint64()
        en = ctypes.c_double()
        f = getattr(self.lib, self.prefix + "snapshot_stats")
        rc = f(kind, radius, length, nc, cen.ctypes.data_as(_dp), rad.ctypes.data_as(_dp), soft, eps, sigma,
               box.ctypes.data_as(_dp), len(pos), pos.ctypes.data_as(_dp), rot.ctypes.data_as(_dp),
               1 if early_exit else 0, ctypes.byref(ov), ctypes.byref(en))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}snapshot_stats failed ({rc})")
        return int(ov.value), float(en.value)


_oracle = None
_ref = None


def oracle() -> Checker:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run __graft_entry__.build()")
        lib = ctypes.CDLL(ORACLE_PATH)
        _setup_common(lib)
        _oracle = Checker(lib, "orc_")
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def _setup_ref(lib):
    lib.ref_pair_decisions.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_size_t, _dp, _u8p]
    lib.ref_set_beads.argtypes = [ctypes.c_int, _dp, _dp]
    lib.ref_set_beads.restype = None
    lib.ref_snapshot_stats.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_size_t, _dp, _dp,
                                       ctypes.c_int, _u64p, _dp]
    lib.ref_pair_energies.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_size_t, _dp, _dp]
    lib.ref_fix_rotation.argtypes = [ctypes.c_size_t, _dp]
    lib.ref_fix_rotation.restype = ctypes.c_int
    lib.ref_ramsnap_write.argtypes = [ctypes.c_char_p, _dp, ctypes.c_size_t, _dp, _dp, ctypes.c_char_p]
    lib.ref_ramsnap_read.argtypes = [ctypes.c_char_p, _dp, ctypes.c_size_t, _dp, _dp]
    lib.ref_ramsnap_read.restype = ctypes.c_long
    lib.ref_nematic_p2.argtypes = [ctypes.c_size_t, _dp]
    lib.ref_nematic_p2.restype = ctypes.c_double
    lib.ref_ramtrj_write.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _dp, _dp,
                                     _dp, ctypes.c_int, ctypes.c_size_t]
    lib.ref_ramtrj_read.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                    ctypes.c_size_t, _dp, _dp, _dp, _u64p]
    lib.ref_ramtrj_read.restype = ctypes.c_long
    lib.ref_packing_fraction.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_size_t]
    lib.ref_packing_fraction.restype = ctypes.c_double
    lib.ref_mc_run.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, _dp, ctypes.c_size_t, _dp, _dp,
                               ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.c_double, ctypes.c_ulong, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.ref_mc_run.restype = ctypes.c_long
    lib.ref_box_info.argtypes = [_dp, _dp, _dp, _dp]
    lib.ref_box_info.restype = None
    lib.ref_translate.argtypes = [_dp, ctypes.c_size_t, _dp, _dp, _dp]
    lib.ref_translate.restype = None
    lib.ref_rotation.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_int, _dp]
    lib.ref_rotation.restype = None
    lib.ref_rotate.argtypes = [ctypes.c_size_t, _dp, _dp, _dp]
    lib.ref_rotate.restype = None
    lib.ref_try_scaling.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_size_t, _dp,
                                    _dp, _dp, _dp]


def ref(perf: bool = False) -> Checker:
    """The reference build: the parity build, or (perf) the Release-optimised one, for AVX-512 hosts where one was
    built and the host supports it."""
    global _ref
    path = REF_PATH
    if perf:
        path = REF_PERF_V4_PATH if os.path.exists(REF_PERF_V4_PATH) and host_has_avx512() else REF_PERF_PATH
    if perf:
        lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        _setup_ref(lib)
        return Checker(lib, "ref_")
    if _ref is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing (built only where the reference sources are present)")
        lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        _setup_ref(lib)
        _ref = Checker(lib, "ref_")
    return _ref


def ref_mc_run(checker: Checker, kind, radius, length, box, pos, rot, temperature, pressure, trans_step, rot_step,
               box_kind, box_step, seed, dd=(1, 1, 1), threads=1, therm=0, avg=0, avg_every=1, soft=0, eps=1.0,
               sigma=1.0):
    """Runs the reference hot path with the restated cycle loop; returns dict(samples, box, pos, rot, seconds, steps)
    with samples rows (packing fraction, P2, total energy / N)."""
    n = len(pos)
    nsamp = (avg // avg_every) if avg_every else 0
    samples = np.zeros(3 * max(nsamp, 1))
    obox = np.zeros(9); opos = np.zeros((n, 3)); orot = np.zeros((n, 9))
    secs = ctypes.c_double()
    steps = np.zeros(3)
    box = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rot = np.ascontiguousarray(rot, dtype=np.float64)
    r = checker.lib.ref_mc_run(kind, radius, length, soft, eps, sigma, box.ctypes.data_as(_dp), n,
                               pos.ctypes.data_as(_dp),
                               rot.ctypes.data_as(_dp), temperature, pressure, trans_step, rot_step, box_kind,
                               box_step, seed, dd[0], dd[1], dd[2], threads, therm, avg, avg_every,
                               samples.ctypes.data_as(_dp), obox.ctypes.data_as(_dp), opos.ctypes.data_as(_dp),
                               orot.ctypes.data_as(_dp), ctypes.byref(secs), steps.ctypes.data_as(_dp))
    if r < 0:
        raise RuntimeError(f"ref_mc_run failed ({r})")
    return dict(samples=samples[:3 * r].reshape(-1, 3), box=obox.reshape(3, 3), pos=opos, rot=orot,
                seconds=secs.value, steps=steps)


def ref_ramtrj_write(checker: Checker, path, cycle_step, boxes, pos, rot, append=False, first_cycle=None):
    """The reference's RamtrjRecorder: snapshots k of (boxes[k], pos[k], rot[k]) at cycles first_cycle + k * step
    (first_cycle defaults to cycle_step); returns the harness status (0 ok, -1 exception)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 9)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(len(boxes), -1, 3)
    rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(len(boxes), -1, 9)
    first = cycle_step if first_cycle is None else first_cycle
    return checker.lib.ref_ramtrj_write(os.fsencode(str(path)), pos.shape[1], cycle_step, len(boxes),
                                        boxes.ctypes.data_as(_dp), pos.ctypes.data_as(_dp), rot.ctypes.data_as(_dp),
                                        int(append), first)


def ref_ramtrj_read(checker: Checker, path, n, max_snap, expected_n=0, jump_cycle=0):
    """The reference's RamtrjPlayer (AutoFix when expected_n > 0): (count or -1, boxes, pos, rot, info) with info =
    particles, snapshots, cycle step, total cycles, fixing needed, fixing successful, inferred, remainder."""
    boxes = np.zeros((max(max_snap, 1), 9))
    pos = np.zeros((max(max_snap, 1), n, 3))
    rot = np.zeros((max(max_snap, 1), n, 9))
    info = np.zeros(8, dtype=np.uint64)
    got = checker.lib.ref_ramtrj_read(os.fsencode(str(path)), expected_n, jump_cycle, n, max_snap,
                                      boxes.ctypes.data_as(_dp), pos.ctypes.data_as(_dp), rot.ctypes.data_as(_dp),
                                      info.ctypes.data_as(_u64p))
    return got, boxes[:max(got, 0)], pos[:max(got, 0)], rot[:max(got, 0)], info


def pair_guarded(kind: int, radius: float, length: float, rec, bound_scale: float = 1.0):
    """The oracle's guarded FP32 pair decisions: (decisions, codes) as rp_pair_decisions_guarded returns them."""
    rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 27)
    dec = np.zeros(len(rec), dtype=np.uint8)
    codes = np.zeros(len(rec), dtype=np.uint8)
    oracle().lib.orc_pair_guarded(kind, radius, length, len(rec), rec.ctypes.data_as(_dp), dec.ctypes.data_as(_u8p),
                                  codes.ctypes.data_as(_u8p), bound_scale)
    return dec, codes


def poly_pair_decisions(centres, radii, rec) -> np.ndarray:
    """The oracle's polysphere pair decisions (records of 27 doubles as for the other shapes)."""
    c = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
    rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 27)
    out = np.zeros(len(rec), dtype=np.uint8)
    oracle().lib.orc_poly_pair_decisions(len(c), c.ctypes.data_as(_dp), r.ctypes.data_as(_dp), len(rec),
                                         rec.ctypes.data_as(_dp), out.ctypes.data_as(_u8p))
    return out


def replay_sweep(kind: int, radius: float, length: float, rec: dict):
    """Replays a device decision record (GpuPacking.record_sweep) through the restated Packing::tryMove: returns
    (decisions per proposal: 1 accepted, 0 overlap, 2 left its cell, 3 not run; positions after; orientations after)."""
    lib = oracle().lib
    ip = ctypes.POINTER(ctypes.c_int)
    n = len(rec["pre_pos"])
    npr = rec["n_proposals"]
    dec = np.zeros(npr, dtype=np.uint8)
    post_pos = np.zeros((n, 3))
    post_rot = np.zeros((n, 9))
    arr = {k: np.ascontiguousarray(rec[k]) for k in ("pre_pos", "pre_rot", "prop_t", "prop_rot")}
    box = np.ascontiguousarray(rec["box"], dtype=np.float64).reshape(9)
    cells = np.ascontiguousarray(rec["cells"], dtype=np.int32)
    shift = np.ascontiguousarray(rec["shift"], dtype=np.float64)
    order = np.ascontiguousarray(rec["colour_order"], dtype=np.int32)
    cell_of = np.ascontiguousarray(rec["pre_cell"], dtype=np.int32)
    prop_k = np.ascontiguousarray(rec["prop_k"], dtype=np.int32)
    rc = lib.orc_replay_sweep(kind, radius, length, box.ctypes.data_as(_dp), n, arr["pre_pos"].ctypes.data_as(_dp),
                              arr["pre_rot"].ctypes.data_as(_dp), cell_of.ctypes.data_as(ip), cells.ctypes.data_as(ip),
                              shift.ctypes.data_as(_dp), rec["period"], len(order), order.ctypes.data_as(ip),
                              rec["moves_per_cell"], rec["n_samplers"], npr, prop_k.ctypes.data_as(ip),
                              arr["prop_t"].ctypes.data_as(_dp), arr["prop_rot"].ctypes.data_as(_dp),
                              dec.ctypes.data_as(_u8p), post_pos.ctypes.data_as(_dp), post_rot.ctypes.data_as(_dp))
    if rc != 0:
        raise RuntimeError(f"orc_replay_sweep failed ({rc})")
    return dec, post_pos, post_rot


def mlp_loss_grad(theta, x, y):
    """The oracle's network: (mean binary cross-entropy, gradient (9217), logits) in double precision."""
    fp = ctypes.POINTER(ctypes.c_float)
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
    y = np.ascontiguousarray(y, dtype=np.float32)
    grad = np.zeros(len(theta))
    z = np.zeros(len(x))
    loss = oracle().lib.orc_mlp_loss_grad(theta.ctypes.data_as(fp), len(x), x.ctypes.data_as(fp),
                                          y.ctypes.data_as(fp), grad.ctypes.data_as(_dp), z.ctypes.data_as(_dp))
    return loss, grad, z
