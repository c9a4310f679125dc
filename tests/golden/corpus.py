"""Deterministic generators for the parity fixtures (near-contact pair corpora and overlapping snapshots).

Only IEEE-exact numpy operations are used (+, -, *, /, sqrt, floor, comparisons on uniform doubles from PCG64), so
the same seed gives bit-identical inputs on any x86-64 host; `records_sha256` lets a test confirm that before it
compares decisions against the committed golden bitmaps.

Corpus (near-contact decision check): a random orientation pair and a random direction; the centre separation
along that direction is bisected so the FP64 segment distance of the two spherocylinder axes equals the contact
distance 2r, then offset by delta ~ U(-1e-6, 1e-6); the second particle is wrapped into a triclinic box and the
image translation T = box * k is recorded, as the reference's Packing passes it (Packing.cpp:21-44, 381-398).
"""
from __future__ import annotations

import hashlib

import numpy as np

BOX = np.array([[50.0, 1.25, -0.75],
                [0.0, 48.0, 2.125],
                [0.0, 0.0, 52.0]])


def _matvec(m, v):
    """Row-major 3x3 times vectors, ((0 + m0 v0) + m1 v1) + m2 v2 per row (no BLAS, no fused ops)."""
    out = np.empty_like(v)
    for i in range(3):
        out[:, i] = ((0.0 + m[i, 0] * v[:, 0]) + m[i, 1] * v[:, 1]) + m[i, 2] * v[:, 2]
    return out


def _inverse(m):
    det = (m[0, 0] * m[1, 1] * m[2, 2] + m[0, 1] * m[1, 2] * m[2, 0] + m[0, 2] * m[1, 0] * m[2, 1]
           - m[0, 2] * m[1, 1] * m[2, 0] - m[0, 1] * m[1, 0] * m[2, 2] - m[0, 0] * m[1, 2] * m[2, 1])
    inv = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            rows = [r for r in range(3) if r != j]
            cols = [c for c in range(3) if c != i]
            mi = m[rows[0], cols[0]] * m[rows[1], cols[1]] - m[rows[0], cols[1]] * m[rows[1], cols[0]]
            inv[i, j] = (mi if (i + j) % 2 == 0 else -mi) / det
    return inv


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotations (row-major, n x 9) from rejection-sampled unit quaternions."""
    q = np.empty((0, 4))
    while len(q) < n:
        c = rng.random((2 * (n - len(q)) + 16, 4)) * 2.0 - 1.0
        nn = (c * c).sum(axis=1)
        c = c[(nn <= 1.0) & (nn > 1e-4)]
        q = np.concatenate([q, c])
    q = q[:n]
    q = q / np.sqrt((q * q).sum(axis=1))[:, None]
    w, x, y, z = q.T
    r = np.empty((n, 9))
    r[:, 0] = 1 - 2 * (y * y + z * z); r[:, 1] = 2 * (x * y - w * z); r[:, 2] = 2 * (x * z + w * y)
    r[:, 3] = 2 * (x * y + w * z); r[:, 4] = 1 - 2 * (x * x + z * z); r[:, 5] = 2 * (y * z - w * x)
    r[:, 6] = 2 * (x * z - w * y); r[:, 7] = 2 * (y * z + w * x); r[:, 8] = 1 - 2 * (x * x + y * y)
    return r


def random_directions(rng: np.random.Generator, n: int) -> np.ndarray:
    d = np.empty((0, 3))
    while len(d) < n:
        c = rng.random((2 * (n - len(d)) + 16, 3)) * 2.0 - 1.0
        nn = (c * c).sum(axis=1)
        d = np.concatenate([d, c[(nn <= 1.0) & (nn > 1e-4)]])
    d = d[:n]
    return d / np.sqrt((d * d).sum(axis=1))[:, None]


def segment_distance2(p1, a1, p2, a2, half):
    """Squared distance of segments p +- half*a (vectorised; generator use only)."""
    u = 2 * half * a1
    v = 2 * half * a2
    w = (p1 - half * a1) - (p2 - half * a2)
    a = (u * u).sum(1); b = (u * v).sum(1); c = (v * v).sum(1); d = (u * w).sum(1); e = (v * w).sum(1)
    D = a * c - b * b
    par = D < 1e-12
    Ds = np.where(par, 1.0, D)
    sN = np.where(par, 0.0, b * e - c * d)
    sD = np.where(par, 1.0, D)
    tN = np.where(par, e, a * e - b * d)
    tD = np.where(par, c, Ds)
    m = (~par) & (sN < 0)
    sN = np.where(m, 0.0, sN); tN = np.where(m, e, tN); tD = np.where(m, c, tD)
    m = (~par) & (sN > sD)
    sN = np.where(m, sD, sN); tN = np.where(m, e + b, tN); tD = np.where(m, c, tD)
    m = tN < 0
    sN = np.where(m, np.clip(-d, 0, None), sN)
    sN = np.where(m & (-d > a), sD, sN)
    sD = np.where(m & (-d >= 0) & (-d <= a), a, sD)
    tN = np.where(m, 0.0, tN)
    m2 = (~m) & (tN > tD)
    tN = np.where(m2, tD, tN)
    s2 = -d + b
    sN = np.where(m2 & (s2 < 0), 0.0, sN)
    sN = np.where(m2 & (s2 > a), sD, sN)
    mid = m2 & (s2 >= 0) & (s2 <= a)
    sN = np.where(mid, s2, sN)
    sD = np.where(mid, a, sD)
    sc = np.where(np.abs(sN) < 1e-12, 0.0, sN / np.where(sD == 0, 1.0, sD))
    tc = np.where(np.abs(tN) < 1e-12, 0.0, tN / np.where(tD == 0, 1.0, tD))
    dp = w + sc[:, None] * u - tc[:, None] * v
    return (dp * dp).sum(1)


def _wrap(p_target, rng_rel, box):
    """Wrap target positions into the box; returns (wrapped position, translation T with wrapped + T ~ target)."""
    inv = _inverse(box)
    rel = _matvec(inv, p_target)
    k = np.floor(rel)
    relw = rel - k
    relw = np.where(relw >= 1.0, 0.0, relw)
    return _matvec(box, relw), _matvec(box, k)


def spherocylinder_contact_corpus(n: int, seed: int, radius: float = 0.5, length: float = 5.0,
                                  delta: float = 1e-6, box: np.ndarray = BOX) -> np.ndarray:
    """n x 27 records pos1, O1, pos2, O2, T with the axes' segment distance within +-delta of 2r."""
    rng = np.random.Generator(np.random.PCG64(seed))
    o1 = random_rotations(rng, n)
    o2 = random_rotations(rng, n)
    a1 = o1[:, [2, 5, 8]]
    a2 = o2[:, [2, 5, 8]]
    dirn = random_directions(rng, n)
    rel1 = rng.random((n, 3))
    p1 = _matvec(box, rel1)
    half = 0.5 * length
    lo = np.zeros(n)
    hi = np.full(n, 2 * radius + length + 0.5)
    target = (2 * radius) ** 2
    for _ in range(56):
        mid = 0.5 * (lo + hi)
        d2 = segment_distance2(p1, a1, p1 + mid[:, None] * dirn, a2, half)
        inside = d2 < target
        lo = np.where(inside, mid, lo)
        hi = np.where(inside, hi, mid)
    lam = 0.5 * (lo + hi) + (rng.random(n) * 2.0 - 1.0) * delta
    p2t = p1 + lam[:, None] * dirn
    p2, T = _wrap(p2t, rng, box)
    rec = np.concatenate([p1, o1, p2, o2, T], axis=1)
    return np.ascontiguousarray(rec)


def chunked_contact_corpus(n: int, seed: int, radius: float = 0.5, length: float = 5.0, delta: float = 1e-6,
                           chunk: int = 1000000, threads: int = 8, box: np.ndarray = BOX) -> np.ndarray:
    """A large spherocylinder corpus made of chunks of `chunk` records; chunk k is
    spherocylinder_contact_corpus(chunk, [seed, k]) (its own PCG64 stream), so the result is the same for any number
    of threads.  numpy releases the interpreter lock inside the array operations, so threads run the bisection of
    different chunks concurrently without starting processes."""
    from concurrent.futures import ThreadPoolExecutor
    rec = np.empty((n, 27))
    starts = list(range(0, n, chunk))

    def make(k):
        m = min(chunk, n - starts[k])
        rec[starts[k]:starts[k] + m] = spherocylinder_contact_corpus(m, [seed, k], radius=radius, length=length,
                                                                     delta=delta, box=box)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(make, range(len(starts))))
    return rec


def sphere_contact_corpus(n: int, seed: int, radius: float = 0.5, delta: float = 1e-6,
                          box: np.ndarray = BOX) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    o = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    dirn = random_directions(rng, n)
    p1 = _matvec(box, rng.random((n, 3)))
    lam = 2 * radius + (rng.random(n) * 2.0 - 1.0) * delta
    p2, T = _wrap(p1 + lam[:, None] * dirn, rng, box)
    return np.ascontiguousarray(np.concatenate([p1, o, p2, o, T], axis=1))


def records_sha256(rec: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(rec, dtype=np.float64).tobytes()).hexdigest()


def random_snapshot(n: int, seed: int, box: np.ndarray, kind: str = "spherocylinder", parallel: bool = False):
    """A snapshot with uniformly random positions inside `box` and random (or parallel) orientations.  Dense
    enough to overlap; used for total-overlap-count parity."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rel = 1e-9 + rng.random((n, 3)) * (1.0 - 2e-9)
    pos = _matvec(box, rel)
    if kind == "sphere" or parallel:
        rot = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    else:
        rot = random_rotations(rng, n)
    return np.ascontiguousarray(pos), np.ascontiguousarray(rot)


def jittered_lattice(n_side: int, spacing: float, jitter: float, seed: int):
    """n_side^3 particles on a simple-cubic lattice of the given spacing, each displaced by U(-jitter, jitter) per
    axis, in the cubic box of side n_side * spacing; identity orientations.  Neighbours stay about `spacing` apart,
    so soft energies are moderate and every pair contributes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = (idx + 0.5) * spacing + (rng.random((len(idx), 3)) * 2.0 - 1.0) * jitter
    box = np.diag([n_side * spacing] * 3)
    rot = np.tile(np.eye(3).reshape(1, 9), (len(idx), 1))
    return box, np.ascontiguousarray(pos), np.ascontiguousarray(rot)


def bead_set(name: str):
    """Body-frame centres and radii of the polysphere fixtures: 'kinked11' (11 touching beads of diameter 1 with a
    30 degree kink), 'dimer' (two touching beads of radius 0.5 on the z axis), 'mixed' (three beads of radii 0.5,
    0.4, 0.3 that do not overlap)."""
    if name == "kinked11":
        from rampack_amd import lattice
        return lattice.kinked_chain(11, 1.0, 30.0)
    if name == "dimer":
        return np.array([[0.0, 0.0, -0.5], [0.0, 0.0, 0.5]]), np.array([0.5, 0.5])
    return np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.9], [0.5, 0.3, -0.7]]), np.array([0.5, 0.4, 0.3])


def dimer_start(st: dict):
    """Aligned polyspheres on an orthorhombic lattice: st["n_side"] per axis at st["spacing"], identity
    orientations (the NpT dumbbell statistics' start state)."""
    from rampack_amd import lattice
    box, pos = lattice.sc_positions((st["n_side"],) * 3, st["spacing"])
    rot = np.tile(np.eye(3).reshape(1, 9), (len(pos), 1))
    return box, np.ascontiguousarray(pos), np.ascontiguousarray(rot)
