"""`casino` on the device: runs a PYON input file (docs/input-file.md) through the GPU sweep, as the reference's
`rampack casino -i input.pyon` (src/frontend/modes/CasinoMode.cpp) does on the CPU.

Supported subset (anything else stops with an error naming it):
  rampack(version 1.x, arrangement, shape, seed, runs, temperature, pressure, move_types, box_move_type,
          walls = [False] * 3, box_move_threads, domain_divisions, handle_signals)   # the last three are accepted and
                                                                                    # ignored: the device decomposes
  arrangement: lattice(cell = sc, ...) with cell_dim + n_cells, box_dim + n_shapes (LatticeDimensionsOptimizer) or
               n_cells + box_dim, transformations = [randomize_rotation(seed, axis = "random")];
               presimulated(file) of a RAMSNAP file
  shape: sphere(r, interaction = hard | wca(epsilon, sigma) | lj(epsilon, sigma) | square_inverse_core(epsilon,
         sigma)), spherocylinder(l, r), polysphere(spheres, volume, ...)
  move_types: translation(step, max_step), rotation(step), rototranslation(trans_step, rot_step = "auto",
              max_trans_step)
  box_move_type: disabled, delta_v(step), log(spec, step, scale_together), linear(spec, step, scale_together),
                 delta_triclinic(step, scale_together)
  runs: integration(run_name, thermalization_cycles, averaging_cycles, snapshot_every, temperature, pressure,
                    move_types, box_move_type, averaging_every, inline_info_every, orientation_fix_every,
                    output_last_snapshot = [ramsnap(file)], record_trajectory = [ramtrj(file)], averages_out,
                    observables, observables_out),
        overlap_relaxation(run_name, snapshot_every, temperature, pressure, move_types, box_move_type,
                           inline_info_every, orientation_fix_every, output_last_snapshot, record_trajectory,
                           observables, observables_out)
  observables: number_density, packing_fraction, compressibility_factor, energy_per_particle, nematic_order,
               temperature, pressure, optionally wrapped in scoped(observable, snapshot, averaging, inline)

Output files follow the reference's formats: averages rows "temperature pressure <name> d<name> ..." appended to
`averages_out` (docs/output-formats.md, "Observable averages"; the d<name> errors are block-averaged standard errors,
not the reference's naive ones, which ignore the correlation between successive samples), observable snapshots "cycle <values>" with a
"cycle <names>" header (ObservablesCollector.cpp:175-203), RAMSNAP and RAMTRJ files.

Usage: python -m rampack_amd.casino -i input.pyon [--verbose]
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

from . import pyon
from .pyon import Call

UNSUPPORTED = "not supported by the device casino driver"


class InputError(ValueError):
    """An input the reference would reject, or one outside the supported subset."""


def _call(v, what):
    if isinstance(v, str):
        v = Call(v)
    if not isinstance(v, Call):
        raise InputError(f"{what}: expected an object, got {v!r}")
    return v


def _float(v, what):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise InputError(f"{what}: expected a number, got {v!r}" + (" (dynamic parameters are " + UNSUPPORTED + ")"
                                                                    if isinstance(v, Call) else ""))
    return float(v)


# ------------------------------------------------------------------------------------------------ shapes

def make_shape(c):
    """ShapeTraits from a shape object (docs/shapes.md): (ra.Shape, particle volume)."""
    from . import Shape
    c = _call(c, "shape")
    if c.name == "sphere":
        r = _float(c.arg(0, "r", required=True), "sphere.r")
        inter = _call(c.arg(1, "interaction", "hard"), "sphere.interaction")
        soft = {"hard": None, "wca": "wca", "lj": "lj", "square_inverse_core": "sic"}.get(inter.name, "?")
        if soft == "?":
            raise InputError(f"interaction {inter.name}: {UNSUPPORTED}")
        kw = {}
        if soft:
            kw = dict(epsilon=_float(inter.arg(0, "epsilon", required=True), "epsilon"),
                      sigma=_float(inter.arg(1, "sigma", required=True), "sigma"))
        return Shape("sphere", radius=r, soft=soft, **kw), 4.0 / 3 * math.pi * r ** 3
    if c.name == "spherocylinder":
        length = _float(c.arg(0, "l", required=True), "spherocylinder.l")
        r = _float(c.arg(1, "r", required=True), "spherocylinder.r")
        if c.arg(2, "interaction", None) not in (None, "hard") and _call(c.arg(2, "interaction"), "").name != "hard":
            raise InputError(f"spherocylinder interaction: {UNSUPPORTED}")
        return Shape("spherocylinder", radius=r, length=length), math.pi * r * r * length + 4.0 / 3 * math.pi * r ** 3
    if c.name == "polysphere":
        spheres = c.arg(0, "spheres", required=True)
        spheres = spheres if isinstance(spheres, list) else [spheres]
        centres, radii = [], []
        for s in spheres:
            s = _call(s, "polysphere.spheres")
            pos = s.arg(0, "pos", required=True)
            r = _float(s.arg(1, "r", required=True), "sphere.r")
            for p in (pos if pos and isinstance(pos[0], list) else [pos]):
                centres.append([_float(x, "sphere.pos") for x in p])
                radii.append(r)
        inter = c.arg(6, "interaction", "hard")
        if _call(inter, "polysphere.interaction").name != "hard":
            raise InputError(f"polysphere interaction: {UNSUPPORTED}")
        volume = _float(c.arg(1, "volume", required=True), "polysphere.volume")
        return Shape("polysphere", centres=np.array(centres), radii=np.array(radii)), volume
    raise InputError(f"shape {c.name}: {UNSUPPORTED}")


# ------------------------------------------------------------------------------------------------ arrangement

def _box_matrix(v, what):
    """box_dim / cell_dim: a number (cube), 3 numbers (cuboid) or 3 box vectors (the columns of the matrix)."""
    if isinstance(v, (int, float)) and not isinstance(v, bool):
        return np.diag([float(v)] * 3)
    if isinstance(v, list) and len(v) == 3 and all(isinstance(x, (int, float)) for x in v):
        return np.diag([float(x) for x in v])
    if isinstance(v, list) and len(v) == 3 and all(isinstance(x, list) and len(x) == 3 for x in v):
        return np.array([[float(x) for x in vec] for vec in v]).T
    raise InputError(f"{what}: expected a number, 3 numbers or 3 vectors, got {v!r}")


def _heights(box):
    vol = abs(np.linalg.det(box))
    a, b, c = box.T
    return np.array([vol / np.linalg.norm(np.cross(b, c)), vol / np.linalg.norm(np.cross(c, a)),
                     vol / np.linalg.norm(np.cross(a, b))])


def _optimize_dims(n_shapes, box):
    """LatticeDimensionsOptimizer::optimize (src/frontend/LatticeDimensionsOptimizer.cpp:10-37) for one-particle
    cells."""
    h = _heights(box)
    target = float(np.cbrt(float(np.prod(h)) / n_shapes))
    dims = [int(math.floor(x / target + 0.5)) for x in h]  # std::round: halves away from zero
    while dims[0] * dims[1] * dims[2] < n_shapes:
        k = int(np.argmax(dims))
        dims[k] += 1
    return dims


def _tait_bryan(ax, ay, az):
    """Matrix<3,3>::rotation(angleX, angleY, angleZ): Rz Ry Rx."""
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def make_arrangement(c, base_dir):
    """Initial box, positions and orientations (docs/initial-arrangement.md)."""
    from . import ramsnap_read
    c = _call(c, "arrangement")
    if c.name == "presimulated":
        path = c.arg(0, "file", required=True)
        box, pos, rot, _ = ramsnap_read(os.path.join(base_dir, path))
        return box, pos, rot
    if c.name != "lattice":
        raise InputError(f"arrangement {c.name}: {UNSUPPORTED}")
    cell = _call(c.arg(0, "cell", required=True), "lattice.cell")
    if cell.name != "sc":
        raise InputError(f"unit cell {cell.name}: {UNSUPPORTED}")
    kw = c.kwargs
    if "fill_partially" in kw:
        raise InputError(f"fill_partially: {UNSUPPORTED}")
    if "cell_dim" in kw and "n_cells" in kw:
        cellm = _box_matrix(kw["cell_dim"], "cell_dim")
        dims = [int(x) for x in kw["n_cells"]]
        box = cellm * np.array(dims)[None, :]
    elif "box_dim" in kw and "n_shapes" in kw:
        box = _box_matrix(kw["box_dim"], "box_dim")
        dims = _optimize_dims(int(kw["n_shapes"]), box)
    elif "box_dim" in kw and "n_cells" in kw:
        box = _box_matrix(kw["box_dim"], "box_dim")
        dims = [int(x) for x in kw["n_cells"]]
    else:
        raise InputError("lattice: give cell_dim and n_cells, box_dim and n_shapes, or n_cells and box_dim")
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3).astype(float)
    rel = (idx + 0.5) / np.array(dims, dtype=float)
    pos = rel @ box.T
    rot = np.tile(np.eye(3).reshape(1, 9), (len(pos), 1))
    for t in kw.get("transformations", []) or []:
        t = _call(t, "transformation")
        if t.name != "randomize_rotation" or t.arg(1, "axis", "random") != "random":
            raise InputError(f"lattice transformation {t.name}: {UNSUPPORTED}")
        # RotationRandomizingTransformer.cpp:30-54: uniform rotations from Tait-Bryan angles; the draws come from
        # numpy's PCG64 seeded with `seed` (the reference's std::mt19937 stream is not reproduced)
        rng = np.random.Generator(np.random.PCG64(int(t.arg(0, "seed", required=True))))
        for i in range(len(rot)):
            ax = rng.uniform(0, 2 * math.pi)
            ay = math.asin(rng.uniform(-1, 1))
            az = rng.uniform(0, 2 * math.pi)
            rot[i] = (_tait_bryan(ax, ay, az) @ rot[i].reshape(3, 3)).ravel()
    return box, np.ascontiguousarray(pos), np.ascontiguousarray(rot)


# ------------------------------------------------------------------------------------------------ environment

def _move_sampler(c):
    c = _call(c, "move type")
    if c.name == "translation":
        mx = c.arg(1, "max_step", None)
        return ("translation", (_float(c.arg(0, "step", required=True), "step"), _float(mx, "max_step") if mx else 0))
    if c.name == "rotation":
        return ("rotation", _float(c.arg(0, "step", required=True), "step"))
    if c.name == "rototranslation":
        rs = c.arg(1, "rot_step", "auto")
        mx = c.arg(2, "max_trans_step", None)
        return ("rototranslation", (_float(c.arg(0, "trans_step", required=True), "trans_step"),
                                    None if rs == "auto" else _float(rs, "rot_step"),
                                    _float(mx, "max_trans_step") if mx else 0))
    raise InputError(f"move type {c.name}: {UNSUPPORTED}")


def _scaling_spec(spec):
    """AnisotropicVolumeScaler's spec (docs/input-file.md, class `linear`): groups of the sides x, y, z (3 = not
    scaled)."""
    named = {"isotropic": "(xyz)", "anisotropic x": "x(yz)", "anisotropic y": "y(zx)", "anisotropic z": "z(xy)",
             "anisotropic xyz": "xyz"}
    text = named.get(spec.strip(), spec.replace(" ", ""))
    groups, g, mode, seen = [None] * 3, 0, None, set()
    for ch in text:
        if ch in "([":
            if mode:
                raise InputError(f"bad scaling spec {spec!r}")
            mode = ch
        elif ch in ")]":
            if mode != {")": "(", "]": "["}[ch]:
                raise InputError(f"bad scaling spec {spec!r}")
            if mode == "(":
                g += 1
            mode = None
        elif ch in "xyz" and ch not in seen:
            seen.add(ch)
            a = "xyz".index(ch)
            if mode == "[":
                groups[a] = 3
            elif mode == "(":
                groups[a] = g
            else:
                groups[a] = g
                g += 1
        else:
            raise InputError(f"bad scaling spec {spec!r}")
    if seen != set("xyz") or mode or g > 3:
        raise InputError(f"bad scaling spec {spec!r}")
    return groups


def _box_move(c):
    c = _call(c, "box move type")
    if c.name == "disabled":
        return None
    if c.name == "delta_v":
        return ("delta_v", _float(c.arg(0, "step", required=True), "step"))
    if c.name in ("log", "linear"):
        groups = _scaling_spec(c.arg(0, "spec", required=True))
        together = c.arg(2, "scale_together", True)
        return (c.name, _float(c.arg(1, "step", required=True), "step"), (groups, not together))
    if c.name == "delta_triclinic":
        together = c.arg(1, "scale_together", True)
        return ("delta_triclinic", _float(c.arg(0, "step", required=True), "step"),
                "scale_together" if together else "one_side")
    raise InputError(f"box move type {c.name}: {UNSUPPORTED}")


# ------------------------------------------------------------------------------------------------ observables

OBSERVABLES = {"number_density": ["rho"], "packing_fraction": ["theta"], "compressibility_factor": ["Z"],
               "energy_per_particle": ["E"], "nematic_order": ["P2"], "temperature": ["T"], "pressure": ["p"]}


def _observables(lst):
    """(name, snapshot, averaging, inline) per observable; scoped(...) selects the scopes, otherwise all three."""
    out = []
    for o in lst or []:
        o = _call(o, "observable")
        if o.name == "scoped":
            inner = _call(o.arg(0, "observable", required=True), "observable")
            scopes = (bool(o.kwargs.get("snapshot", False)), bool(o.kwargs.get("averaging", False)),
                      bool(o.kwargs.get("inline", False)))
        else:
            inner, scopes = o, (True, True, True)
        if inner.name not in OBSERVABLES:
            raise InputError(f"observable {inner.name}: {UNSUPPORTED}")
        out.append((inner.name,) + scopes)
    return out


class Casino:
    def __init__(self, root, base_dir=".", log=print):
        root = _call(root, "input")
        if root.name != "rampack":
            raise InputError("the input must be a rampack(...) object")
        version = root.arg(0, "version", required=True)
        major = int(str(version[0] if isinstance(version, list) else version).split(".")[0])
        if major != 1:
            raise InputError(f"input version {version}: this driver reads the 1.x format")
        self.shape, self.v0 = make_shape(root.arg(2, "shape", required=True))
        self.seed = int(root.arg(3, "seed", required=True))
        self.box, self.pos, self.rot = make_arrangement(root.arg(1, "arrangement", required=True), base_dir)
        self.runs = root.arg(4, "runs", required=True)
        if not isinstance(self.runs, list) or not self.runs:
            raise InputError("runs: expected a non-empty array")
        walls = root.kwargs.get("walls", [False, False, False])
        if any(walls):
            raise InputError(f"walls: {UNSUPPORTED}")
        self.env = {k: root.kwargs.get(k) for k in ("temperature", "pressure", "move_types", "box_move_type")}
        self.base_dir = base_dir
        self.log = log

    def _environment(self, run):
        env = dict(self.env)
        for k in env:
            if k in run.kwargs and run.kwargs[k] is not None:
                env[k] = run.kwargs[k]
        self.env = env  # the next run starts from this environment (Simulation environment combining)
        if env["temperature"] is None or env["move_types"] is None or env["box_move_type"] is None:
            raise InputError(f"run {run.arg(0, 'run_name')}: incomplete simulation environment")
        box_move = _box_move(env["box_move_type"])
        pressure = env["pressure"]
        if box_move is not None and pressure is None:
            raise InputError("pressure is required with a box move")
        samplers = [_move_sampler(m) for m in env["move_types"]]
        return _float(env["temperature"], "temperature"), (_float(pressure, "pressure") if pressure is not None
                                                           else 0.0), samplers, box_move

    def _apply(self, pk, temperature, pressure, samplers, box_move):
        bm = None
        if box_move is not None:
            kind, step = box_move[0], box_move[1]
            spec = box_move[2] if len(box_move) > 2 else None
            if kind in ("log", "linear"):
                groups, independent = spec
                bm = (kind, step, _spec_text(groups, independent))
            else:
                bm = (kind, step) if spec is None else (kind, step, spec)
        first = samplers[0]
        trans = first[1][0] if first[0] != "rotation" else 1.0
        pk.set_environment(temperature, pressure, trans, None, box_move=bm)
        pk.set_move_samplers(samplers)

    def _values(self, pk, temperature, pressure):
        o = pk.observables()
        n = len(self.pos)
        v = o["volume"]
        return {"rho": n / v, "theta": n * self.v0 / v, "Z": pressure * v / (n * temperature),
                "E": pk.total_energy() / n, "P2": o["nematic_p2"], "T": temperature, "p": pressure}

    def run(self):
        from . import GpuPacking, RamtrjRecorder, mean_and_error
        pk = GpuPacking(self.shape, self.box, self.pos, self.rot, seed=self.seed)
        for run in self.runs:
            run = _call(run, "run")
            if run.name not in ("integration", "overlap_relaxation"):
                raise InputError(f"run type {run.name}: {UNSUPPORTED}")
            name = run.arg(0, "run_name", required=True)
            temperature, pressure, samplers, box_move = self._environment(run)
            self._apply(pk, temperature, pressure, samplers, box_move)
            relax = run.name == "overlap_relaxation"
            snapshot_every = int(run.arg(1 if relax else 3, "snapshot_every", required=True))
            inline_every = int(run.kwargs.get("inline_info_every", 100))
            pk.set_orientation_fix_every(int(run.kwargs.get("orientation_fix_every", 10000)))
            if relax and run.kwargs.get("helper_shape") is not None:
                raise InputError(f"overlap_relaxation helper_shape: {UNSUPPORTED}")
            obs = _observables(run.kwargs.get("observables", []))
            snap_names = [(o[0], k) for o in obs if o[1] for k in OBSERVABLES[o[0]]]
            avg_names = [(o[0], k) for o in obs if o[2] for k in OBSERVABLES[o[0]]]
            obs_out = run.kwargs.get("observables_out")
            obs_file = None
            if obs_out and snap_names:
                path = os.path.join(self.base_dir, obs_out)
                fresh = not os.path.exists(path)
                obs_file = open(path, "a")
                if fresh:
                    obs_file.write("cycle " + " ".join(k for _, k in snap_names) + " \n")
            recorders = []
            for t in run.kwargs.get("record_trajectory", []) or []:
                t = _call(t, "record_trajectory")
                if t.name != "ramtrj":
                    raise InputError(f"trajectory format {t.name}: {UNSUPPORTED}")
                recorders.append(RamtrjRecorder(os.path.join(self.base_dir, t.arg(0, "file", required=True)),
                                                len(self.pos), snapshot_every))
            cycle = 0
            start = time.perf_counter()

            def advance(cycles, adapt):
                nonlocal cycle
                while cycles > 0:
                    chunk = min(cycles, snapshot_every - cycle % snapshot_every)
                    c = pk.run_cycles(chunk, adapt_steps=adapt)
                    cycle += chunk
                    cycles -= chunk
                    if cycle % snapshot_every == 0:
                        if obs_file:
                            vals = self._values(pk, temperature, pressure)
                            obs_file.write(f"{cycle} " + " ".join(repr(vals[k]) for _, k in snap_names) + " \n")
                        for rec in recorders:
                            rec.record_snapshot(pk, cycle)
                    if inline_every and cycle % inline_every < chunk:
                        speed = cycle / max(time.perf_counter() - start, 1e-9)
                        self.log(f"[{name}] cycle {cycle}: moves {c['moves_accepted']}/{c['moves_attempted']}, "
                                 f"box {c['box_accepted']}/{c['box_attempted']}, {speed:.1f} cycles/s")

            samples = {k: [] for _, k in avg_names}
            if relax:
                # Simulation::relaxOverlaps (Simulation.cpp:222-284): cycles in counting mode until no overlap
                pk.set_overlap_counting(True)
                while pk.cached_overlaps() > 0:
                    advance(snapshot_every, True)
                    self.log(f"[{name}] cycle {cycle}: {pk.cached_overlaps()} overlaps left")
                pk.set_overlap_counting(False)
            else:
                therm = int(run.arg(1, "thermalization_cycles", required=True))
                avg = int(run.arg(2, "averaging_cycles", required=True))
                every = int(run.kwargs.get("averaging_every", 0))
                if avg > 0 and every <= 0:
                    raise InputError("averaging_every must be positive when averaging_cycles > 0")
                advance(therm, True)
                done = 0
                while done < avg:
                    chunk = min(every, avg - done)
                    advance(chunk, False)
                    done += chunk
                    if chunk == every and avg_names:
                        vals = self._values(pk, temperature, pressure)
                        for _, k in avg_names:
                            samples[k].append(vals[k])
                averages_out = run.kwargs.get("averages_out")
                if averages_out and avg_names:
                    path = os.path.join(self.base_dir, averages_out)
                    fresh = not os.path.exists(path)
                    with open(path, "a") as f:
                        if fresh:
                            f.write("temperature pressure " + "".join(f"{k} d{k} " for _, k in avg_names) + "\n")
                        row = [repr(temperature), repr(pressure)]
                        for _, k in avg_names:
                            m, e = mean_and_error(samples[k], method="blocks")
                            row += [repr(m), repr(e)]
                        f.write(" ".join(row) + " \n")
            if obs_file:
                obs_file.close()
            for rec in recorders:
                rec.close()
            for s in run.kwargs.get("output_last_snapshot", []) or []:
                s = _call(s, "output_last_snapshot")
                if s.name != "ramsnap":
                    raise InputError(f"snapshot format {s.name}: {UNSUPPORTED}")
                pk.save_ramsnap(os.path.join(self.base_dir, s.arg(0, "file", required=True)))
            self.log(f"[{name}] finished after {cycle} cycles")
        pk.close()


def _spec_text(groups, independent):
    words = []
    for g in range(3):
        axes = "".join("xyz"[a] for a in range(3) if groups[a] == g)
        if axes:
            words.append(axes)
    fixed = "".join("xyz"[a] for a in range(3) if groups[a] == 3)
    spec = ("independent " if independent else "anisotropic ") + " ".join(words)
    return spec + (f" [{fixed}]" if fixed else "")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m rampack_amd.casino", description=__doc__.splitlines()[0])
    ap.add_argument("-i", "--input", required=True, help="PYON input file")
    ap.add_argument("-q", "--quiet", action="store_true")
    args = ap.parse_args(argv)
    with open(args.input) as f:
        root = pyon.parse(f.read())
    log = (lambda *_: None) if args.quiet else (lambda s: print(s, flush=True))
    try:
        Casino(root, os.path.dirname(os.path.abspath(args.input)), log).run()
    except (InputError, pyon.PyonError) as e:
        print(f"input error: {e}", file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main())
