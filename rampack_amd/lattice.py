"""Start states on simple-cubic lattices (the reference's `lattice` arrangement with an sc unit cell,
Lattice.cpp:61-80, UnitCellFactory.cpp:8-18): one particle per cell at the cell centre.  Boxes are returned as
row-major 3x3 matrices whose columns are the box sides."""
from __future__ import annotations

import numpy as np


def sc_positions(n_side, spacing):
    nx, ny, nz = n_side
    sx, sy, sz = spacing
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = np.stack([(ii.ravel() + 0.5) * sx, (jj.ravel() + 0.5) * sy, (kk.ravel() + 0.5) * sz], axis=1)
    box = np.diag([nx * sx, ny * sy, nz * sz]).astype(np.float64)
    return box, np.ascontiguousarray(pos, dtype=np.float64)


def hard_sphere_lattice(n_side: int, packing_fraction: float, radius: float = 0.5):
    """n_side^3 spheres on an sc lattice in a cubic box at the given packing fraction."""
    n = n_side ** 3
    side = (n * 4.0 / 3 * np.pi * radius ** 3 / packing_fraction) ** (1.0 / 3)
    box, pos = sc_positions((n_side,) * 3, (side / n_side,) * 3)
    rot = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    return box, pos, rot


def spherocylinder_lattice(n_side, spacing, tilt_seed=None, max_tilt=0.0):
    """Spherocylinders along z (orientation = identity, primary axis = column 2) on an sc lattice.  With max_tilt
    > 0 every rod is rotated by a random angle up to max_tilt about a random horizontal axis (the caller keeps the
    spacing large enough for the tilted rods not to overlap)."""
    box, pos = sc_positions(n_side, spacing)
    n = len(pos)
    rot = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    if tilt_seed is not None and max_tilt > 0:
        rng = np.random.Generator(np.random.PCG64(tilt_seed))
        phi = rng.random(n) * 2 * np.pi
        ang = rng.random(n) * max_tilt
        ax = np.stack([np.cos(phi), np.sin(phi), np.zeros(n)], axis=1)
        c, s = np.cos(ang), np.sin(ang)
        for i in range(n):
            k = np.array([[0, -ax[i, 2], ax[i, 1]], [ax[i, 2], 0, -ax[i, 0]], [-ax[i, 1], ax[i, 0], 0]])
            rot[i] = (np.eye(3) + k * s[i] + (1 - c[i]) * (k @ k)).ravel()
    return box, pos, np.ascontiguousarray(rot)


def spherocylinder_smectic(n: int, packing_fraction: float, radius: float = 0.5, length: float = 5.0,
                           layer_gap: float = 0.2):
    """AA-stacked layers of parallel rods: n_z layers of spacing L + 2r + layer_gap, a square in-plane lattice
    whose spacing gives the packing fraction.  n must factor as n_xy^2 * n_z with the in-plane spacing >= 2r."""
    v0 = np.pi * radius ** 2 * length + 4.0 / 3 * np.pi * radius ** 3
    c = length + 2 * radius + layer_gap
    a = np.sqrt(v0 / (packing_fraction * c))
    if a < 2 * radius:
        raise ValueError("packing fraction too high for the layer spacing")
    best = None
    for nz in range(1, n + 1):
        if n % nz:
            continue
        nxy2 = n // nz
        nxy = int(round(np.sqrt(nxy2)))
        if nxy * nxy != nxy2:
            continue
        aspect = (nxy * a) / (nz * c)
        score = abs(np.log(aspect))
        if best is None or score < best[0]:
            best = (score, nxy, nz)
    if best is None:
        raise ValueError("n must be n_xy^2 * n_z")
    _, nxy, nz = best
    return spherocylinder_lattice((nxy, nxy, nz), (a, a, c))


def orthorhombic_arrange(n: int, dims, particles_in_line=None, cell=None):
    """OrthorhombicArrangingModel::arrange (OrthorhombicArrangingModel.cpp:11-70): ceil(cbrt(n)) particles per line
    by default, centred in the box, x slowest; identity orientations."""
    dims = np.asarray(dims, dtype=float)
    if particles_in_line is None:
        line = int(np.ceil(np.cbrt(n)))
        particles_in_line = (line, line, line)
        cell = dims / line
    cell = np.asarray(cell, dtype=float)
    offset = (dims - (np.asarray(particles_in_line) - 1) * cell) / 2
    pos = []
    for i in range(particles_in_line[0]):
        for j in range(particles_in_line[1]):
            for k in range(particles_in_line[2]):
                if len(pos) >= n:
                    break
                pos.append([i * cell[0] + offset[0], j * cell[1] + offset[1], k * cell[2] + offset[2]])
    pos = np.asarray(pos[:n], dtype=float)
    return np.diag(dims), pos, np.tile(np.eye(3).reshape(1, 9), (len(pos), 1))


def kinked_chain(n_beads: int = 11, diameter: float = 1.0, kink_deg: float = 30.0, kink_at: int = None):
    """Body-frame bead centres and radii of a chain of touching spheres with one kink: beads 0..kink_at lie on the
    z axis at unit spacing (times the diameter), the rest continue in the x-z plane deflected by kink_deg from the z
    axis.  The centres are shifted so that their mean is the origin.  (The deflection reading of a "30 degree kink";
    the interior angle between the two arms is 180 - kink_deg.)"""
    if kink_at is None:
        kink_at = n_beads // 2
    pts = []
    for i in range(n_beads):
        if i <= kink_at:
            pts.append((0.0, 0.0, i * diameter))
        else:
            t = np.radians(kink_deg)
            k = (i - kink_at) * diameter
            pts.append((k * np.sin(t), 0.0, kink_at * diameter + k * np.cos(t)))
    c = np.array(pts)
    c -= c.mean(axis=0)
    return c, np.full(n_beads, diameter / 2)


def rod_replica_start(n: int, packing_fraction: float, radius: float = 0.5, length: float = 5.0):
    """n parallel spherocylinders along z filling the first n sites of a k x k x m lattice, the box scaled so that
    the packing fraction is exact: the start of one replica of an equation-of-state scan (sc lattice, parallel
    orientations).  The cell is c = 1.5 (length + 2 radius) along z (end-to-end gaps wide enough for box moves
    to compress the start) and k ~ c/a times as many sites across, so
    the box is near cubic.  The unfilled sites shrink the scaled cell; a start whose scaled cell leaves rods less
    than 0.1 % apart is refused."""
    v0 = np.pi * radius ** 2 * (length + 4.0 / 3 * radius)
    c = 1.5 * (length + 2 * radius)
    a = np.sqrt(v0 / (packing_fraction * c))
    if a < 2 * radius:
        raise ValueError("packing fraction too high for a parallel lattice start")
    m = max(1, int(np.ceil(np.cbrt(n * a * a / (c * c)))))
    k = int(np.ceil(np.sqrt(n / m)))
    box, pos = sc_positions((k, k, m), (a, a, c))
    pos = pos[:n]
    scale = np.cbrt(n * v0 / packing_fraction / np.linalg.det(box))
    if c * scale < 1.001 * (length + 2 * radius) or a * scale < 1.001 * 2 * radius:
        raise ValueError("lattice too sparse for an overlap-free parallel start")
    box, pos = box * scale, np.ascontiguousarray(pos * scale)
    rot = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    return box, pos, rot


def kinked_chain_lattice(n_side, packing_fraction: float, n_beads: int = 11, kink_deg: float = 30.0,
                         z_spacing: float = 11.2):
    """Kinked chains of touching unit beads (kinked_chain), all with the identity orientation, on an n_side lattice
    (sc): the in-plane spacing a follows from the packing fraction (n v_chain / (a^2 c n_xyz)), c = z_spacing along
    the chains.  Returns (box, positions, orientations, centres, radii)."""
    c, r = kinked_chain(n_beads, 1.0, kink_deg)
    n = int(np.prod(n_side))
    v_chain = n_beads * np.pi / 6
    a = np.sqrt(n * v_chain / (packing_fraction * z_spacing * n))
    box, pos = sc_positions(n_side, (a, a, z_spacing))
    rot = np.tile(np.eye(3).reshape(1, 9), (len(pos), 1))
    return box, pos, np.ascontiguousarray(rot), c, r
