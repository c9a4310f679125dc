"""MI355X-native rampack Monte Carlo sweep.

Python access to the C ABI declared in include/rampack_amd.h (librampack_amd.so, built for gfx950).  The
classes mirror the reference's Packing / Simulation surface for the sweep (src/core/Packing.h,
src/core/Simulation.h): shapes are uploaded once, cycles run on the GPU, and the state comes back on request.

There is no CPU fallback: if the HIP library is missing or no GPU is visible, every entry point raises.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "librampack_amd.so")

SHAPE_SPHERE = 0
SHAPE_SPHEROCYLINDER = 1
SHAPE_POLYSPHERE = 2
BOX_LINEAR = 4
MOVE_ROTOTRANSLATION = 0
MOVE_TRANSLATION = 1
MOVE_ROTATION = 2
MAX_CENTRES = 16
SOFT_NONE = 0
SOFT_WCA = 1
SOFT_WCA_SEGMENT = 2
SOFT_LJ = 3
SOFT_SIC = 4
SOFT_HARD_CORE = 0x100
PREC_FP64 = 0
PREC_GUARDED_FP32 = 1
PREC_GUARDED_FP32_VERIFY = 2
BOX_NONE = 0
BOX_DELTA_TRICLINIC = 1
BOX_LOG_ISOTROPIC = 2
BOX_DELTA_V = 3

RP_OK = 0
RP_ERR_PRECONDITION = -1
RP_ERR_VALIDATION = -2
RP_ERR_DEVICE = -3
RP_ERR_UNSUPPORTED = -4

# Every symbol include/rampack_amd.h declares.
EXPORTED_SYMBOLS = (
    "rp_create", "rp_set_environment", "rp_run_cycles", "rp_count_total_overlaps", "rp_download",
    "rp_set_timing", "rp_get_timing", "rp_destroy", "rp_pair_decisions", "rp_last_error", "rp_get_observables",
    "rp_total_energy", "rp_pair_energies", "rp_fix_orientations", "rp_set_orientation_fix_every",
    "rp_ramsnap_write", "rp_ramsnap_save", "rp_ramsnap_info", "rp_ramsnap_read",
    "rp_set_overlap_counting", "rp_cached_overlaps", "rp_relax_overlaps",
    "rp_ramtrj_recorder_open", "rp_ramtrj_record", "rp_ramtrj_record_packing", "rp_ramtrj_recorder_last_cycle",
    "rp_ramtrj_recorder_close", "rp_ramtrj_player_open", "rp_ramtrj_player_header", "rp_ramtrj_next_snapshot",
    "rp_ramtrj_jump_to_snapshot", "rp_ramtrj_player_close", "rp_pair_decisions_guarded",
    "rp_pair_decisions_guarded_noalloc",
    "rp_try_scaling", "rp_revert_scaling", "rp_set_move_samplers", "rp_get_move_samplers", "rp_run_replicas",
    "rp_record_sweep", "rp_box_info", "rp_apply_moves", "rp_set_precision", "rp_get_guard_stats", "rp_autotune",
    "rp_mlp_gradient", "rp_mlp_infer", "rp_mlp_train", "rp_mlp_order_stats", "rp_mlp_get_weights",
    "rp_mlp_set_weights", "rp_mlp_pairs",
)


class ShapeDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("radius", ctypes.c_double), ("length", ctypes.c_double),
                ("soft", ctypes.c_int), ("epsilon", ctypes.c_double), ("sigma", ctypes.c_double),
                ("n_centres", ctypes.c_int), ("centres", ctypes.POINTER(ctypes.c_double)),
                ("radii", ctypes.POINTER(ctypes.c_double))]


class MoveDesc(ctypes.Structure):
    _fields_ = [("trans_step", ctypes.c_double), ("rot_step", ctypes.c_double),
                ("max_trans_step", ctypes.c_double), ("kind", ctypes.c_int)]


class BoxDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("step", ctypes.c_double), ("groups", ctypes.c_int * 3),
                ("independent", ctypes.c_int), ("one_side", ctypes.c_int)]


class Counters(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "cycles", "moves_attempted", "moves_accepted", "moves_left_cell", "box_attempted", "box_accepted",
        "box_scans", "box_too_small", "unstaged_blocks", "pair_tests")] + [
        ("trans_step", ctypes.c_double), ("rot_step", ctypes.c_double), ("box_step", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Timing(ctypes.Structure):
    _fields_ = [("sweep_launches", ctypes.c_uint64), ("sweep_ms", ctypes.c_double),
                ("rebuild_launches", ctypes.c_uint64), ("rebuild_ms", ctypes.c_double),
                ("scan_launches", ctypes.c_uint64), ("scan_ms", ctypes.c_double),
                ("proposal_launches", ctypes.c_uint64), ("proposal_ms", ctypes.c_double),
                ("lds_limit", ctypes.c_uint64), ("sweep_cap", ctypes.c_uint64), ("sweep_waves", ctypes.c_uint64),
                ("sweep_period", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AutotuneEntry(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int), ("schedule", ctypes.c_int), ("sweep_us_per_cycle", ctypes.c_double)]


AUTOTUNE_MAX = 8
SCHEDULES = ("plain", "pipelined", "grouped")


class Observables(ctypes.Structure):
    _fields_ = [("packing_fraction", ctypes.c_double), ("number_density", ctypes.c_double),
                ("volume", ctypes.c_double), ("nematic_p2", ctypes.c_double), ("q_tensor", ctypes.c_double * 6)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "q_tensor"}
        d["q_tensor"] = list(self.q_tensor)
        return d


class RamtrjHeader(ctypes.Structure):
    _fields_ = [("version_major", ctypes.c_uint8), ("version_minor", ctypes.c_uint8),
                ("num_particles", ctypes.c_uint64), ("num_snapshots", ctypes.c_uint64),
                ("cycle_step", ctypes.c_uint64), ("current_snapshot", ctypes.c_uint64)]


class RamtrjFix(ctypes.Structure):
    _fields_ = [("fixing_needed", ctypes.c_int), ("fixing_successful", ctypes.c_int),
                ("header_snapshots", ctypes.c_uint64), ("inferred_snapshots", ctypes.c_uint64),
                ("bytes_remainder", ctypes.c_uint64), ("bytes_per_snapshot", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GuardStats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_uint64), ("fp64_fallback", ctypes.c_uint64), ("segment_stage", ctypes.c_uint64),
                ("segment_fallback", ctypes.c_uint64), ("mismatches", ctypes.c_uint64),
                ("max_err_centre", ctypes.c_double), ("max_err_segment", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SweepRecord(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 9), ("shift", ctypes.c_double * 3), ("cells", ctypes.c_int * 3),
                ("period", ctypes.c_int), ("n_colours", ctypes.c_int), ("colour_order", ctypes.c_int * 27),
                ("moves_per_cell", ctypes.c_int), ("n_samplers", ctypes.c_int), ("n_proposals", ctypes.c_size_t)]


class MlpStats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_uint64), ("overlap_fraction", ctypes.c_double), ("loss_initial", ctypes.c_double),
                ("loss_final", ctypes.c_double), ("steps", ctypes.c_uint64), ("moves", ctypes.c_uint64),
                ("moves_rejected", ctypes.c_uint64), ("candidates_per_move", ctypes.c_double),
                ("tests_natural", ctypes.c_double), ("tests_nearest_first", ctypes.c_double),
                ("tests_learned", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class RampackError(RuntimeError):
    """Error returned through the C ABI (PreconditionException / ValidationException analogues)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class PreconditionError(RampackError):
    pass


class ValidationError(RampackError):
    pass


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads librampack_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() to compile the HIP library")
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    lib.rp_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ShapeDesc), dp, ctypes.c_size_t, dp, dp,
                              ctypes.c_uint64, ctypes.c_int]
    lib.rp_set_environment.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(MoveDesc),
                                       ctypes.POINTER(BoxDesc)]
    lib.rp_run_cycles.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(Counters)]
    lib.rp_run_replicas.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                    ctypes.POINTER(Counters)]
    lib.rp_count_total_overlaps.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    lib.rp_download.argtypes = [vp, dp, dp, dp]
    lib.rp_set_timing.argtypes = [vp, ctypes.c_int]
    lib.rp_get_timing.argtypes = [vp, ctypes.POINTER(Timing)]
    lib.rp_autotune.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(AutotuneEntry), ctypes.POINTER(ctypes.c_size_t),
                                ctypes.POINTER(ctypes.c_size_t)]
    lib.rp_destroy.argtypes = [vp]
    lib.rp_destroy.restype = None
    lib.rp_pair_decisions.argtypes = [ctypes.POINTER(ShapeDesc), ctypes.c_size_t, dp,
                                      ctypes.POINTER(ctypes.c_uint8)]
    lib.rp_get_observables.argtypes = [vp, ctypes.POINTER(Observables)]
    lib.rp_total_energy.argtypes = [vp, dp]
    lib.rp_try_scaling.argtypes = [vp, dp, dp]
    lib.rp_set_move_samplers.argtypes = [vp, ctypes.POINTER(MoveDesc), ctypes.c_size_t]
    lib.rp_get_move_samplers.argtypes = [vp, ctypes.POINTER(MoveDesc), ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_size_t)]
    lib.rp_revert_scaling.argtypes = [vp]
    lib.rp_set_overlap_counting.argtypes = [vp, ctypes.c_int]
    lib.rp_cached_overlaps.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.rp_relax_overlaps.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_uint64)]
    lib.rp_ramsnap_write.argtypes = [ctypes.c_char_p, dp, ctypes.c_size_t, dp, dp, ctypes.c_char_p]
    lib.rp_ramsnap_save.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
    lib.rp_ramsnap_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    lib.rp_ramsnap_read.argtypes = [ctypes.c_char_p, dp, ctypes.c_size_t, dp, dp, ctypes.c_char_p, ctypes.c_size_t]
    sz = ctypes.c_size_t
    lib.rp_ramtrj_recorder_open.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p, sz, sz, ctypes.c_int]
    lib.rp_ramtrj_record.argtypes = [vp, dp, sz, dp, dp, sz]
    lib.rp_ramtrj_record_packing.argtypes = [vp, vp, sz]
    lib.rp_ramtrj_recorder_last_cycle.argtypes = [vp, ctypes.POINTER(sz)]
    lib.rp_ramtrj_recorder_close.argtypes = [vp]
    lib.rp_ramtrj_player_open.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p, sz, ctypes.POINTER(RamtrjFix)]
    lib.rp_ramtrj_player_header.argtypes = [vp, ctypes.POINTER(RamtrjHeader)]
    lib.rp_ramtrj_next_snapshot.argtypes = [vp, dp, sz, dp, dp]
    lib.rp_ramtrj_jump_to_snapshot.argtypes = [vp, sz, dp, sz, dp, dp]
    lib.rp_ramtrj_player_close.argtypes = [vp]
    lib.rp_ramtrj_player_close.restype = None
    lib.rp_fix_orientations.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.rp_set_orientation_fix_every.argtypes = [vp, ctypes.c_uint64]
    lib.rp_pair_energies.argtypes = [ctypes.POINTER(ShapeDesc), ctypes.c_size_t, dp, dp]
    lib.rp_pair_decisions_guarded.argtypes = [ctypes.POINTER(ShapeDesc), ctypes.c_size_t, dp,
                                              ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint8),
                                              ctypes.POINTER(GuardStats)]
    lib.rp_pair_decisions_guarded_noalloc.argtypes = [ctypes.POINTER(ShapeDesc), ctypes.c_size_t, vp, vp, vp,
                                                      ctypes.POINTER(GuardStats)]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.rp_record_sweep.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(SweepRecord), dp, dp, ip, ip, ip, dp, dp,
                                    ctypes.POINTER(ctypes.c_uint8), dp, dp]
    lib.rp_box_info.argtypes = [dp, dp, dp, dp]
    lib.rp_apply_moves.argtypes = [dp, ctypes.c_size_t, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    lib.rp_set_precision.argtypes = [vp, ctypes.c_int]
    lib.rp_get_guard_stats.argtypes = [vp, ctypes.POINTER(GuardStats)]
    fp = ctypes.POINTER(ctypes.c_float)
    lib.rp_mlp_gradient.argtypes = [fp, ctypes.c_size_t, fp, fp, dp, fp]
    lib.rp_mlp_infer.argtypes = [fp, ctypes.c_size_t, fp, fp]
    sz = ctypes.c_size_t
    lib.rp_mlp_train.argtypes = [vp, sz, sz, sz, sz, sz, ctypes.c_double, ctypes.POINTER(MlpStats)]
    lib.rp_mlp_order_stats.argtypes = [vp, sz, ctypes.POINTER(MlpStats)]
    lib.rp_mlp_get_weights.argtypes = [vp, fp]
    lib.rp_mlp_set_weights.argtypes = [vp, fp]
    lib.rp_mlp_pairs.argtypes = [vp, sz, sz, fp, fp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(sz)]
    lib.rp_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(code: int):
    if code == RP_OK:
        return
    msg = _lib.rp_last_error().decode()
    if code == RP_ERR_PRECONDITION:
        raise PreconditionError(code, msg)
    if code == RP_ERR_VALIDATION:
        raise ValidationError(code, msg)
    raise RampackError(code, msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Shape:
    """ShapeTraits descriptor: 'sphere' (radius), 'spherocylinder' (radius, length between cap centres) or
    'polysphere' (hard beads: body-frame `centres`, n x 3, and `radii`, n; PolysphereTraits).
    soft: None (hard core), 'wca' (a sphere with RepulsiveLennardJonesInteraction(epsilon, sigma)), 'lj' (a sphere with
    LennardJonesInteraction, truncated at its range 3 sigma in sweeps and total energies), 'sic' (a sphere with
    SquareInverseCoreInteraction) or 'wca_segment' (a spherocylinder with WCA on the distance between the axis
    segments).  hard_core (with a soft interaction): CompoundInteraction of the shape's hard core and the soft part --
    pairs whose hard shapes overlap are forbidden, the energy is the soft part's."""
    kind: str
    radius: float = 0.5
    length: float = 0.0
    soft: Optional[str] = None
    epsilon: float = 1.0
    sigma: float = 1.0
    centres: Optional[tuple] = None
    radii: Optional[tuple] = None
    hard_core: bool = False

    def desc(self) -> ShapeDesc:
        kinds = {"sphere": SHAPE_SPHERE, "spherocylinder": SHAPE_SPHEROCYLINDER, "polysphere": SHAPE_POLYSPHERE}
        softs = {None: SOFT_NONE, "wca": SOFT_WCA, "wca_segment": SOFT_WCA_SEGMENT, "lj": SOFT_LJ, "sic": SOFT_SIC}
        if self.kind not in kinds:
            raise ValueError(f"unknown shape {self.kind}")
        if self.soft not in softs:
            raise ValueError(f"unknown soft interaction {self.soft}")
        soft = softs[self.soft] | (SOFT_HARD_CORE if self.hard_core else 0)
        d = ShapeDesc(kinds[self.kind], float(self.radius), float(self.length), soft, float(self.epsilon),
                      float(self.sigma))
        if self.kind == "polysphere":
            c = np.ascontiguousarray(self.centres, dtype=np.float64).reshape(-1, 3)
            r = np.ascontiguousarray(self.radii, dtype=np.float64).reshape(-1)
            if len(c) != len(r):
                raise ValueError("polysphere: one radius per centre")
            d.n_centres = len(c)
            d.centres = _dptr(c)
            d.radii = _dptr(r)
            d._keep = (c, r)  # the arrays must outlive the descriptor's use
        return d

    @property
    def volume(self) -> float:
        """SphereTraits::getVolume / SpherocylinderTraits::getVolume (SpherocylinderTraits.cpp:48-50); a polysphere
        of beads that do not overlap: the sum of the bead volumes (PolysphereTraits.cpp:124-131)."""
        r, l = self.radius, self.length
        if self.kind == "sphere":
            return 4.0 / 3 * np.pi * r ** 3
        if self.kind == "polysphere":
            return float(sum(4.0 / 3 * np.pi * x ** 3 for x in self.radii))
        return np.pi * r * r * l + 4.0 / 3 * np.pi * r ** 3

    @property
    def range(self) -> float:
        cut = self.sigma * 2 ** (1 / 6)
        if self.kind == "polysphere":
            hard = 2 * max(self.radii)
        else:
            hard = 2 * self.radius + (self.length if self.kind == "spherocylinder" else 0.0)
        soft = {"wca": cut, "lj": 3 * self.sigma, "sic": self.sigma, "wca_segment": self.length + cut}.get(self.soft)
        if soft is None:
            return hard
        return max(soft, hard) if self.hard_core else soft

    @property
    def total_range(self) -> float:
        """Interaction::getTotalRangeRadius (Interaction.cpp:55-70)."""
        if self.kind != "polysphere":
            return self.range
        c = np.asarray(self.centres, dtype=np.float64).reshape(-1, 3)
        return 2 * float(np.sqrt((c * c).sum(axis=1).max())) + self.range


def pair_decisions(shape: Shape, records: np.ndarray) -> np.ndarray:
    """Device pair decisions on records of 27 doubles: pos1, O1 (row-major), pos2, O2, T."""
    lib = load_library()
    rec = np.ascontiguousarray(records, dtype=np.float64).reshape(-1, 27)
    out = np.zeros(len(rec), dtype=np.uint8)
    d = shape.desc()
    _check(lib.rp_pair_decisions(ctypes.byref(d), len(rec), _dptr(rec),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
    return out


GUARD_NOALLOC_MAX_PAIRS = 1 << 20


def pair_decisions_guarded_noalloc(shape: Shape, n: int, d_records: int, d_decisions: int, d_codes: int = 0,
                                   stats: bool = True):
    """rp_pair_decisions_guarded_noalloc on device arrays given by address (records 27 n doubles, decisions and codes
    n bytes; d_codes 0 for none): allocates nothing; returns the statistics dict (or None with stats=False)."""
    lib = load_library()
    d = shape.desc()
    st = GuardStats()
    _check(lib.rp_pair_decisions_guarded_noalloc(ctypes.byref(d), int(n), ctypes.c_void_p(d_records),
                                                 ctypes.c_void_p(d_decisions), ctypes.c_void_p(d_codes or None),
                                                 ctypes.byref(st) if stats else None))
    return st.as_dict() if stats else None


def pair_decisions_guarded(shape: Shape, records: np.ndarray, verify: bool = True):
    """Guarded FP32 pair decisions (rp_pair_decisions_guarded): (decisions, codes, stats dict), or only the decisions
    with verify=False (FP64 then runs only where the FP32 bounds leave a pair undecided).  The decisions equal
    pair_decisions'; codes bits 0-1 = 0 FP32 no overlap / 1 FP32 overlap / 2 FP64 fallback, bit 2 = segment stage,
    bit 3 = differs from FP64."""
    lib = load_library()
    rec = np.ascontiguousarray(records, dtype=np.float64).reshape(-1, 27)
    out = np.zeros(len(rec), dtype=np.uint8)
    d = shape.desc()
    u8 = ctypes.POINTER(ctypes.c_uint8)
    if not verify:
        _check(lib.rp_pair_decisions_guarded(ctypes.byref(d), len(rec), _dptr(rec), out.ctypes.data_as(u8), None,
                                             None))
        return out
    codes = np.zeros(len(rec), dtype=np.uint8)
    st = GuardStats()
    _check(lib.rp_pair_decisions_guarded(ctypes.byref(d), len(rec), _dptr(rec), out.ctypes.data_as(u8),
                                         codes.ctypes.data_as(u8), ctypes.byref(st)))
    return out, codes, st.as_dict()


def box_info(box: np.ndarray):
    """rp_box_info: (inverse 3x3, heights, volume) as the device path computes them."""
    lib = load_library()
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
    inv, h, v = np.zeros(9), np.zeros(3), np.zeros(1)
    _check(lib.rp_box_info(_dptr(b), _dptr(inv), _dptr(h), _dptr(v)))
    return inv.reshape(3, 3), h, float(v[0])


def apply_moves(box, pos, rot, t, axis, angle):
    """rp_apply_moves: the sweep's move arithmetic on the device; returns (trial positions, rotation matrices, trial
    orientations, sin/cos of the angles)."""
    lib = load_library()
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (pos, rot, t, axis, angle)]
    n = len(arrs[4])
    out = [np.zeros((n, 3)), np.zeros((n, 9)), np.zeros((n, 9)), np.zeros((n, 2))]
    _check(lib.rp_apply_moves(_dptr(b), n, *[_dptr(a) for a in arrs], *[_dptr(o) for o in out]))
    return tuple(out)


MLP_PARAMS = 9217


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_gradient(theta, x, y):
    """rp_mlp_gradient: (mean binary cross-entropy, gradient) of the ordering network on a batch, on the device."""
    lib = load_library()
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
    y = np.ascontiguousarray(y, dtype=np.float32)
    grad = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = ctypes.c_double()
    _check(lib.rp_mlp_gradient(_fptr(theta), len(x), _fptr(x), _fptr(y), ctypes.byref(loss), _fptr(grad)))
    return loss.value, grad


def mlp_infer(theta, x):
    """rp_mlp_infer: overlap probabilities of pairs with the BF16 MFMA network."""
    lib = load_library()
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
    out = np.zeros(len(x), dtype=np.float32)
    _check(lib.rp_mlp_infer(_fptr(theta), len(x), _fptr(x), _fptr(out)))
    return out


def pair_energies(shape: Shape, records: np.ndarray) -> np.ndarray:
    """Device soft pair energies on records of 27 doubles: pos1, O1 (row-major), pos2, O2, T."""
    lib = load_library()
    rec = np.ascontiguousarray(records, dtype=np.float64).reshape(-1, 27)
    out = np.zeros(len(rec))
    d = shape.desc()
    _check(lib.rp_pair_energies(ctypes.byref(d), len(rec), _dptr(rec), _dptr(out)))
    return out


class GpuPacking:
    """A packing resident on the GPU with its Simulation environment (Packing + Simulation surface)."""

    def __init__(self, shape: Shape, box: np.ndarray, pos: np.ndarray, rot: np.ndarray, seed: int = 1234,
                 precision: int = PREC_FP64):
        lib = load_library()
        self.shape = shape
        self.n = len(pos)
        box = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
        if len(rot) != self.n:
            raise ValueError("positions and orientations differ in length")
        self._ctx = ctypes.c_void_p()
        d = shape.desc()
        _check(lib.rp_create(ctypes.byref(self._ctx), ctypes.byref(d), _dptr(box), self.n, _dptr(pos), _dptr(rot),
                             int(seed), int(precision)))

    def close(self):
        if self._ctx:
            _lib.rp_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_environment(self, temperature: float, pressure: float, trans_step: float,
                        rot_step: Optional[float] = None, max_trans_step: float = 0.0,
                        box_move: Optional[tuple] = None):
        """box_move: None (NVT) or (kind, step[, spec]): 'delta_triclinic' (spec 'one_side' moves one box side per
        move, TriclinicDeltaScaler's scaleTogether = false), 'log' or 'linear' (spec as the reference's
        AnisotropicVolumeScaler: 'isotropic' (default), 'anisotropic <groups>' or 'independent <groups>' with groups
        such as 'x y z' or 'xy z', and '[y]' for a side that is not scaled), or 'delta_v'."""
        m = MoveDesc(float(trans_step), float(rot_step) if rot_step else 0.0, float(max_trans_step))
        bd = None
        if box_move is not None:
            kinds = {"delta_triclinic": BOX_DELTA_TRICLINIC, "log": BOX_LOG_ISOTROPIC, "delta_v": BOX_DELTA_V,
                     "linear": BOX_LINEAR}
            bd = BoxDesc(kinds[box_move[0]], float(box_move[1]))
            spec = box_move[2] if len(box_move) > 2 else None
            if spec and box_move[0] == "delta_triclinic":
                if spec not in ("one_side", "scale_together"):
                    raise ValueError(f"unknown delta_triclinic spec {spec}")
                bd.one_side = 1 if spec == "one_side" else 0
            elif spec and box_move[0] in ("log", "linear"):
                words = spec.split()
                if words[0] == "isotropic" and len(words) == 1:
                    pass
                elif words[0] in ("anisotropic", "independent") and len(words) > 1:
                    bd.independent = 1 if words[0] == "independent" else 0
                    seen = set()
                    for g, axes in enumerate(words[1:]):
                        fixed = axes.startswith("[") and axes.endswith("]")  # sides that are not scaled
                        for a in axes.strip("[]"):
                            if a not in "xyz" or a in seen:
                                raise ValueError(f"bad scaling spec {spec}")
                            seen.add(a)
                            bd.groups["xyz".index(a)] = 3 if fixed else g
                    if seen != set("xyz"):
                        raise ValueError(f"scaling spec {spec} must name x, y and z once each")
                else:
                    raise ValueError(f"bad scaling spec {spec}")
            elif spec:
                raise ValueError(f"{box_move[0]} takes no spec")
        _check(_lib.rp_set_environment(self._ctx, float(temperature), float(pressure), ctypes.byref(m),
                                       ctypes.byref(bd) if bd is not None else None))

    def run_cycles(self, n_cycles: int, adapt_steps: bool = False) -> dict:
        c = Counters()
        _check(_lib.rp_run_cycles(self._ctx, int(n_cycles), 1 if adapt_steps else 0, ctypes.byref(c)))
        return c.as_dict()

    def record_sweep(self) -> dict:
        """rp_record_sweep: one cycle's particle moves (no box move) with the state before and after them in the
        sweep's storage order, every proposal and its acceptance flag."""
        rec = SweepRecord()
        n = self.n
        ip = ctypes.POINTER(ctypes.c_int)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        code = _lib.rp_record_sweep(self._ctx, 0, ctypes.byref(rec), None, None, None, None, None, None, None, None,
                                    None, None)
        if code != RP_ERR_PRECONDITION or rec.n_proposals == 0:
            _check(code)
        npr = int(rec.n_proposals)
        out = {"pre_pos": np.zeros((n, 3)), "pre_rot": np.zeros((n, 9)), "pre_tag": np.zeros(n, dtype=np.int32),
               "pre_cell": np.zeros(n, dtype=np.int32), "prop_k": np.zeros(npr, dtype=np.int32),
               "prop_t": np.zeros((npr, 3)), "prop_rot": np.zeros((npr, 9)), "accepted": np.zeros(npr, dtype=np.uint8),
               "post_pos": np.zeros((n, 3)), "post_rot": np.zeros((n, 9))}
        _check(_lib.rp_record_sweep(
            self._ctx, npr, ctypes.byref(rec), _dptr(out["pre_pos"]), _dptr(out["pre_rot"]),
            out["pre_tag"].ctypes.data_as(ip), out["pre_cell"].ctypes.data_as(ip), out["prop_k"].ctypes.data_as(ip),
            _dptr(out["prop_t"]), _dptr(out["prop_rot"]), out["accepted"].ctypes.data_as(u8),
            _dptr(out["post_pos"]), _dptr(out["post_rot"])))
        out.update(box=np.array(rec.box[:]).reshape(3, 3), shift=np.array(rec.shift[:]), cells=list(rec.cells),
                   period=rec.period, colour_order=list(rec.colour_order[:rec.n_colours]),
                   moves_per_cell=rec.moves_per_cell, n_samplers=rec.n_samplers, n_proposals=npr)
        return out

    def set_precision(self, precision: int):
        """PREC_FP64, PREC_GUARDED_FP32 or PREC_GUARDED_FP32_VERIFY (rp_set_precision)."""
        _check(_lib.rp_set_precision(self._ctx, int(precision)))

    def guard_stats(self) -> dict:
        """Statistics of the guarded FP32 scan (rp_get_guard_stats)."""
        st = GuardStats()
        _check(_lib.rp_get_guard_stats(self._ctx, ctypes.byref(st)))
        return st.as_dict()

    def mlp_train(self, snapshots: int, cycles_between: int, moves_per_snapshot: int, epochs: int = 3,
                  batch: int = 4096, learning_rate: float = 1e-3) -> dict:
        """rp_mlp_train: the learned ordering's network trained on pairs from this run's trajectory."""
        st = MlpStats()
        _check(_lib.rp_mlp_train(self._ctx, snapshots, cycles_between, moves_per_snapshot, epochs, batch,
                                 float(learning_rate), ctypes.byref(st)))
        return st.as_dict()

    def mlp_order_stats(self, moves: int) -> dict:
        """rp_mlp_order_stats: exact tests per trial move in natural, nearest-first and learned order."""
        st = MlpStats()
        _check(_lib.rp_mlp_order_stats(self._ctx, moves, ctypes.byref(st)))
        return st.as_dict()

    def mlp_weights(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, dtype=np.float32)
        _check(_lib.rp_mlp_get_weights(self._ctx, _fptr(out)))
        return out

    def set_mlp_weights(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        _check(_lib.rp_mlp_set_weights(self._ctx, _fptr(theta)))

    def mlp_pairs(self, moves: int):
        """rp_mlp_pairs: (features n x 12, exact decisions n, meta n x 3) of fresh sampled trial moves."""
        n = ctypes.c_size_t()
        code = _lib.rp_mlp_pairs(self._ctx, moves, 0, None, None, None, ctypes.byref(n))
        if code != RP_ERR_PRECONDITION or n.value == 0:
            _check(code)
        # sample again with room for a larger draw: each call draws fresh moves
        cap = int(n.value * 2 + 1024)
        x = np.zeros((cap, 12), dtype=np.float32)
        y = np.zeros(cap, dtype=np.float32)
        meta = np.zeros((cap, 3), dtype=np.int32)
        _check(_lib.rp_mlp_pairs(self._ctx, moves, cap, _fptr(x), _fptr(y),
                                 meta.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(n)))
        k = int(n.value)
        return x[:k], y[:k], meta[:k]

    def count_total_overlaps(self, early_exit: bool = False) -> int:
        v = ctypes.c_uint64()
        _check(_lib.rp_count_total_overlaps(self._ctx, 1 if early_exit else 0, ctypes.byref(v)))
        return int(v.value)

    def download(self):
        box = np.zeros(9)
        pos = np.zeros((self.n, 3))
        rot = np.zeros((self.n, 9))
        _check(_lib.rp_download(self._ctx, _dptr(box), _dptr(pos), _dptr(rot)))
        return box.reshape(3, 3), pos, rot

    def set_timing(self, enable):
        """True: event timing of every sweep, rebuild, scan and proposal launch; "sampled": one sweep launch per cycle
        (rotating through the colours); False: off."""
        _check(_lib.rp_set_timing(self._ctx, 2 if enable == "sampled" else (1 if enable else 0)))

    def autotune(self, cycles_per_shape: int = 10) -> dict:
        """rp_autotune: runs cycles_per_shape cycles (no step adaptation) with each candidate sweep launch shape and
        keeps the fastest; the run goes on exactly as it would have without tuning.  Returns {"candidates": [{waves,
        schedule, sweep_us_per_cycle}], "chosen": index}."""
        out = (AutotuneEntry * AUTOTUNE_MAX)()
        n, chosen = ctypes.c_size_t(), ctypes.c_size_t()
        _check(_lib.rp_autotune(self._ctx, cycles_per_shape, out, ctypes.byref(n), ctypes.byref(chosen)))
        cands = [{"waves": out[i].waves, "schedule": SCHEDULES[out[i].schedule] if 0 <= out[i].schedule < 3
                  else "default", "sweep_us_per_cycle": out[i].sweep_us_per_cycle} for i in range(n.value)]
        return {"candidates": cands, "chosen": chosen.value}

    def timing(self) -> dict:
        t = Timing()
        _check(_lib.rp_get_timing(self._ctx, ctypes.byref(t)))
        return t.as_dict()

    def fix_orientations(self) -> int:
        """Packing::renormalizeOrientations: returns the number of particles whose fix was skipped (overlap)."""
        v = ctypes.c_uint64()
        _check(_lib.rp_fix_orientations(self._ctx, ctypes.byref(v)))
        return int(v.value)

    def set_orientation_fix_every(self, every: int):
        _check(_lib.rp_set_orientation_fix_every(self._ctx, int(every)))

    def set_overlap_counting(self, enable: bool):
        _check(_lib.rp_set_overlap_counting(self._ctx, 1 if enable else 0))

    def cached_overlaps(self) -> int:
        v = ctypes.c_uint64()
        _check(_lib.rp_cached_overlaps(self._ctx, ctypes.byref(v)))
        return int(v.value)

    def relax_overlaps(self, max_cycles: int = 1 << 62):
        """Simulation::relaxOverlaps: returns (cycles run, overlaps left)."""
        done = ctypes.c_size_t()
        left = ctypes.c_uint64()
        _check(_lib.rp_relax_overlaps(self._ctx, int(max_cycles), ctypes.byref(done), ctypes.byref(left)))
        return int(done.value), int(left.value)

    def save_ramsnap(self, path: str, aux: Optional[dict] = None):
        """RamsnapWriter::write of the device-resident state."""
        _check(_lib.rp_ramsnap_save(self._ctx, os.fsencode(path), _aux_text(aux)))

    def total_energy(self) -> float:
        """Packing::getTotalEnergy of the device-resident state (0 for hard shapes)."""
        v = ctypes.c_double()
        _check(_lib.rp_total_energy(self._ctx, ctypes.byref(v)))
        return float(v.value)

    def set_move_samplers(self, samplers):
        """The environment's move samplers (after set_environment): a list of (kind, step) with kind 'rototranslation'
        (step = (translation, rotation or None for the automatic rule[, maximum translation])), 'translation'
        (step = translation or (translation, maximum)) or 'rotation' (step = rotation angle).  Each cycle makes
        len(samplers) * N moves, each with a sampler drawn uniformly."""
        kinds = {"rototranslation": MOVE_ROTOTRANSLATION, "translation": MOVE_TRANSLATION, "rotation": MOVE_ROTATION}
        arr = (MoveDesc * len(samplers))()
        for i, (kind, step) in enumerate(samplers):
            if kind not in kinds:
                raise ValueError(f"unknown move sampler {kind}")
            d = arr[i]
            d.kind = kinds[kind]
            st = step if isinstance(step, (tuple, list)) else (step,)
            if kind == "rotation":
                d.rot_step = float(st[0])
            else:
                d.trans_step = float(st[0])
                if kind == "rototranslation" and len(st) > 1 and st[1] is not None:
                    d.rot_step = float(st[1])
                mx = st[2] if kind == "rototranslation" and len(st) > 2 else (st[1] if kind == "translation"
                                                                              and len(st) > 1 else 0.0)
                d.max_trans_step = float(mx or 0.0)
        _check(_lib.rp_set_move_samplers(self._ctx, arr, len(samplers)))

    def move_samplers(self) -> list:
        """The move samplers with their current step sizes: dicts of kind, trans_step, rot_step, max_trans_step."""
        n = ctypes.c_size_t()
        _check(_lib.rp_get_move_samplers(self._ctx, None, 0, ctypes.byref(n)))
        arr = (MoveDesc * max(n.value, 1))()
        _check(_lib.rp_get_move_samplers(self._ctx, arr, n.value, None))
        names = {MOVE_ROTOTRANSLATION: "rototranslation", MOVE_TRANSLATION: "translation", MOVE_ROTATION: "rotation"}
        return [dict(kind=names[d.kind], trans_step=d.trans_step, rot_step=d.rot_step,
                     max_trans_step=d.max_trans_step) for d in arr[:n.value]]

    def try_scaling(self, new_box) -> float:
        """Packing::tryScaling (Packing.cpp:172-208): scales the device-resident state to new_box (3x3, columns are
        the box vectors) and returns +inf on an overlap (-inf / +inf by the count change in overlap-counting mode),
        else the total energy change (0 for hard shapes).  The scaled state stays current until revert_scaling."""
        b = np.ascontiguousarray(new_box, dtype=np.float64).reshape(9)
        v = ctypes.c_double()
        _check(_lib.rp_try_scaling(self._ctx, _dptr(b), ctypes.byref(v)))
        return float(v.value)

    def revert_scaling(self):
        """Packing::revertScaling (Packing.cpp:363-371): the state before the last try_scaling becomes current."""
        _check(_lib.rp_revert_scaling(self._ctx))

    def observables(self) -> dict:
        """PackingFraction and NematicOrder of the device-resident state (rp_get_observables)."""
        o = Observables()
        _check(_lib.rp_get_observables(self._ctx, ctypes.byref(o)))
        return o.as_dict()

    def packing_fraction(self) -> float:
        return self.observables()["packing_fraction"]


def _aux_text(aux) -> bytes:
    lines = []
    for k, v in (aux or {}).items():
        k = str(k)
        if not k or any(ch.isspace() for ch in k):
            raise ValueError(f"RAMSNAP aux key {k!r} must be non-empty and free of whitespace")
        lines.append(f"{k} {v}\n")
    return "".join(lines).encode()


def run_replicas(packings, n_cycles: int, adapt_steps: bool = False) -> list:
    """GpuPacking.run_cycles on every packing together (rp_run_replicas), one counter dict per packing: each colour
    pass of all packings in one launch per kernel, one host thread and one stream.  Every packing ends bit for bit
    where its own run_cycles call would take it."""
    packings = list(packings)
    ctxs = (ctypes.c_void_p * len(packings))(*[p._ctx.value for p in packings])
    out = (Counters * len(packings))()
    _check(_lib.rp_run_replicas(ctxs, len(packings), int(n_cycles), 1 if adapt_steps else 0, out))
    return [c.as_dict() for c in out]


def ramsnap_write(path: str, box: np.ndarray, pos: np.ndarray, rot: np.ndarray, aux: Optional[dict] = None):
    """RamsnapWriter::write (RamsnapWriter.cpp:8-48): a RAMSNAP snapshot of the given state."""
    lib = load_library()
    box = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
    if len(pos) != len(rot):
        raise ValueError("positions and orientations differ in length")
    _check(lib.rp_ramsnap_write(os.fsencode(path), _dptr(box), len(pos), _dptr(pos), _dptr(rot), _aux_text(aux)))


def ramsnap_read(path: str):
    """RamsnapReader::read (RamsnapReader.cpp:15-92): (box 3x3, positions n x 3, orientations n x 9, aux dict)."""
    lib = load_library()
    n = ctypes.c_size_t()
    nbytes = ctypes.c_size_t()
    _check(lib.rp_ramsnap_info(os.fsencode(path), ctypes.byref(n), ctypes.byref(nbytes)))
    box = np.zeros(9)
    pos = np.zeros((n.value, 3))
    rot = np.zeros((n.value, 9))
    buf = ctypes.create_string_buffer(nbytes.value)
    _check(lib.rp_ramsnap_read(os.fsencode(path), _dptr(box), n.value, _dptr(pos), _dptr(rot), buf, nbytes.value))
    aux = {}
    for line in buf.value.decode().splitlines():
        key, _, value = line.partition(" ")
        aux[key] = value
    return box.reshape(3, 3), pos, rot, aux


class RamtrjRecorder:
    """RamtrjRecorder (src/core/io/RamtrjRecorder.cpp:8-66), opened as SimulationRecorderFactory does
    (src/frontend/SimulationRecorderFactory.cpp:13-31): append=False truncates the file, append=True continues a
    recording of the same particle count and cycle step.  Snapshot k (k = 1, 2, ...) is recorded at cycle
    k * cycle_step; close() writes the snapshot count into the header."""

    def __init__(self, path: str, num_particles: int, cycle_step: int, append: bool = False):
        lib = load_library()
        self._rec = ctypes.c_void_p()
        self.num_particles = int(num_particles)
        _check(lib.rp_ramtrj_recorder_open(ctypes.byref(self._rec), os.fsencode(path), self.num_particles,
                                           int(cycle_step), int(bool(append))))

    def record_snapshot(self, state, cycle: int):
        """Records a GpuPacking (its device-resident state) or a (box, positions, orientations) triple."""
        if self._rec is None:
            raise PreconditionError(RP_ERR_PRECONDITION, "RAMTRJ recorder is closed")
        if isinstance(state, GpuPacking):
            _check(_lib.rp_ramtrj_record_packing(self._rec, state._ctx, int(cycle)))
            return
        box, pos, rot = state
        box = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
        if len(pos) != len(rot):
            raise ValueError("positions and orientations differ in length")
        _check(_lib.rp_ramtrj_record(self._rec, _dptr(box), len(pos), _dptr(pos), _dptr(rot), int(cycle)))

    def get_last_cycle_number(self) -> int:
        out = ctypes.c_size_t()
        _check(_lib.rp_ramtrj_recorder_last_cycle(self._rec, ctypes.byref(out)))
        return out.value

    def close(self):
        if self._rec is not None:
            rec, self._rec = self._rec, None
            _check(_lib.rp_ramtrj_recorder_close(rec))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_rec", None) is not None and _lib is not None:
            _lib.rp_ramtrj_recorder_close(self._rec)
            self._rec = None


class RamtrjPlayer:
    """RamtrjPlayer (src/core/io/RamtrjPlayer.cpp:11-128), opened as IO::loadRamtrjPlayer does
    (src/frontend/IO.cpp:32-64).  With auto_fix_particles=None the file is read strictly (a length that disagrees
    with the header is a ValidationError); with the expected particle count a truncated file is repaired in memory
    (RamtrjPlayer::AutoFix) and `auto_fix` holds the report.  Snapshots come back as (box 3x3, positions n x 3,
    orientations n x 9)."""

    def __init__(self, path: str, auto_fix_particles: Optional[int] = None):
        lib = load_library()
        self._p = ctypes.c_void_p()
        fix = RamtrjFix()
        try:
            _check(lib.rp_ramtrj_player_open(ctypes.byref(self._p), os.fsencode(path),
                                             int(auto_fix_particles or 0), ctypes.byref(fix)))
        finally:
            self.auto_fix = fix.as_dict() if auto_fix_particles else None
        h = self._header()
        self.version = (h.version_major, h.version_minor)
        self.num_particles = h.num_particles
        self.num_snapshots = h.num_snapshots
        self.cycle_step = h.cycle_step

    def _header(self) -> RamtrjHeader:
        if self._p is None:
            raise PreconditionError(RP_ERR_PRECONDITION, "RAMTRJ player is closed")
        h = RamtrjHeader()
        _check(_lib.rp_ramtrj_player_header(self._p, ctypes.byref(h)))
        return h

    def get_total_cycles(self) -> int:
        return self.cycle_step * self.num_snapshots

    def get_cycle_step(self) -> int:
        return self.cycle_step

    def get_num_molecules(self) -> int:
        return self.num_particles

    def get_current_snapshot_cycles(self) -> int:
        return self._header().current_snapshot * self.cycle_step

    def has_next(self) -> bool:
        return self._p is not None and self._header().current_snapshot < self.num_snapshots

    def _buffers(self):
        n = self.num_particles
        return np.zeros(9), np.zeros((n, 3)), np.zeros((n, 9))

    def next_snapshot(self):
        box, pos, rot = self._buffers()
        if self._p is None:
            raise PreconditionError(RP_ERR_PRECONDITION, "RAMTRJ player is closed")
        _check(_lib.rp_ramtrj_next_snapshot(self._p, _dptr(box), len(pos), _dptr(pos), _dptr(rot)))
        return box.reshape(3, 3), pos, rot

    def jump_to_snapshot(self, cycle: int):
        box, pos, rot = self._buffers()
        if self._p is None:
            raise PreconditionError(RP_ERR_PRECONDITION, "RAMTRJ player is closed")
        _check(_lib.rp_ramtrj_jump_to_snapshot(self._p, int(cycle), _dptr(box), len(pos), _dptr(pos), _dptr(rot)))
        return box.reshape(3, 3), pos, rot

    def last_snapshot(self):
        return self.jump_to_snapshot(self.num_snapshots * self.cycle_step)

    def close(self):
        if self._p is not None:
            _lib.rp_ramtrj_player_close(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_p", None) is not None and _lib is not None:
            _lib.rp_ramtrj_player_close(self._p)
            self._p = None


def mean_and_error(samples, method: str = "naive") -> tuple:
    """The mean and its standard error.  method "naive": Quantity::calculateFromSamples (src/utils/Quantity.cpp:45-62),
    the naive standard error of the mean, summed in sample order as the reference does; "blocks": the block-averaged
    standard error (block_error), which stays valid for correlated samples."""
    if method == "blocks":
        s = np.asarray(samples, dtype=np.float64)
        return (float(s.mean()) if len(s) else 0.0), block_error(s)
    if method != "naive":
        raise ValueError(f"unknown error method {method}")
    s = [float(x) for x in samples]
    if not s:
        return 0.0, 0.0
    if len(s) == 1:
        return s[0], 0.0
    acc = 0.0
    for x in s:
        acc += x
    value = acc / len(s)
    err = 0.0
    for x in s:
        err += (value - x) ** 2
    err = err / len(s) / (len(s) - 1)
    return value, math.sqrt(err)


def block_error(samples, min_blocks: int = 16) -> float:
    """Standard error of the mean of a correlated series by block averaging (H. Flyvbjerg and H. G. Petersen,
    J. Chem. Phys. 91, 461 (1989)).  Level 0 is the series itself; each further level averages neighbouring pairs
    of the level before (an odd last value is dropped).  At each level with at least min_blocks blocks the naive
    standard error of its block means is taken; the estimate is the largest of them: once blocks are longer than the
    correlation time the estimates stop growing (the plateau), and shorter blocks underestimate the error.  Series
    shorter than min_blocks get the naive error of the samples."""
    x = np.asarray(samples, dtype=np.float64).ravel()
    if len(x) < 2:
        return 0.0
    best = float(x.std(ddof=1) / np.sqrt(len(x)))
    while len(x) >= 2 * min_blocks:
        x = 0.5 * (x[0:len(x) // 2 * 2:2] + x[1:len(x) // 2 * 2:2])
        best = max(best, float(x.std(ddof=1) / np.sqrt(len(x))))
    return best


def nematic_p2(rot: np.ndarray) -> float:
    """NematicOrder (NematicOrder.cpp:9-60): largest-magnitude eigenvalue of Q = (3<a a^T> - I)/2, a = column 2."""
    a = np.asarray(rot).reshape(-1, 3, 3)[:, :, 2]
    q = 0.5 * (3 * (a.T @ a) / len(a) - np.eye(3))
    w = np.linalg.eigvalsh(q)
    return float(w[np.argmax(np.abs(w))])
