"""Known-answer tests that pin the oracle (CPU restatement) without the reference build: values restated from the
reference's own unit tests, closed forms and brute force."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib
from oracle_lib import _d

SC = 1
SPHERE = 0


def euler_rotation(ax, ay, az):
    """Matrix<3,3>::rotation(ax, ay, az) (Matrix.tpp:411-432)."""
    sx, sy, sz = math.sin(ax), math.sin(ay), math.sin(az)
    cx, cy, cz = math.cos(ax), math.cos(ay), math.cos(az)
    return np.array([[cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
                     [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
                     [-sy, sx * cy, cx * cy]])


def record(p1, o1, p2, o2, t=(0, 0, 0)):
    return np.concatenate([np.asarray(p1, float), np.asarray(o1, float).ravel(), np.asarray(p2, float),
                           np.asarray(o2, float).ravel(), np.asarray(t, float)])


# SpherocylinderTraitsTest.cpp:14-101 -- SpherocylinderTraits(3, 2): length 3, radius 2; found in Mathematica.
SC_CASES = {
    "sphere-sphere": ([[0.92669949443125, -0.3009522885099318, 0.2250683608628711],
                       [0.3235062902233933, 0.943614995716347, -0.07024542721869537],
                       [-0.191237358292679, 0.1379074323590797, 0.971807497858173]],
                      [[0.1017126282787795, -0.2093565370521866, 0.972535028491077],
                       [0.5191108031629524, 0.845122866944617, 0.127637431056981],
                       [-0.848633322046771, 0.4918711011645504, 0.1946389081120092]],
                      [1, 2, 3], [-3.5, 2.9, -0.5]),
    "sphere-cylinder": ([[0.936848795202308, -0.03157560239884583, 0.34830635403497],
                         [-0.03157560239884583, 0.984212198800577, 0.174153177017485],
                         [-0.34830635403497, -0.174153177017485, 0.921060994002885]],
                        [[0.07918504528890558, 0.3698006820181268, -0.925730621823391],
                         [-0.2008438095940726, 0.915521563787973, 0.348542723904781],
                         [0.976417683550607, 0.1583278933673811, 0.1467677942585273]],
                        [1, 2, 3], [-0.5, 4, -2]),
    "cylinder-cylinder": ([[0.921060994002885, 0.1446263415347036, 0.361565853836759],
                           [-0.1446263415347036, 0.989111861241777, -0.02722034689555686],
                           [-0.361565853836759, -0.02722034689555686, 0.931949132761108]],
                          [[0.4161880560663263, 0.5570946854720225, -0.7186327388914037],
                           [-0.902545539870646, 0.1570999152673587, -0.4009129145869377],
                           [-0.1104493116652922, 0.815453939865286, 0.5681864320017205]],
                          [1, 2, 3], [-0.5, 1.7, -0.8]),
}


@pytest.mark.parametrize("case", sorted(SC_CASES))
@pytest.mark.parametrize("scale,expected", [(0.95, 1), (1.05, 0)])
def test_spherocylinder_mathematica_cases(case, scale, expected):
    r1, r2, t1, t2 = SC_CASES[case]
    corr = euler_rotation(0, math.pi / 2, 0)
    o1 = np.asarray(r1) @ corr
    o2 = np.asarray(r2) @ corr
    p1 = np.asarray(t1, float) * scale
    p2 = np.asarray(t2, float) * scale
    rec = np.stack([record(p1, o1, p2, o2), record(p2, o2, p1, o1)])
    dec = oracle_lib.oracle().pair_decisions(SC, 2.0, 3.0, rec)
    assert list(dec) == [expected, expected]


def test_spherocylinder_periodic_case():
    """SpherocylinderTraitsTest.cpp:104-122: z = 96.6 vs 3.5 in a 100 box overlaps, 96.4 does not (length 3,
    radius 2, both along z after the correcting rotation is not applied: identity orientation)."""
    o = oracle_lib.oracle()
    box = np.diag([100.0, 100.0, 100.0]).ravel()
    for z1, expected in [(96.6, 1), (96.4, 0)]:
        p1, p2 = np.array([50, 50, z1]), np.array([50, 50, 3.5])
        trans = np.zeros(3); corr = np.zeros(3); d2 = ctypes.c_double()
        o.lib.orc_pbc(_d(box), _d(p1), _d(p2), _d(corr), _d(trans), ctypes.byref(d2))
        rec = record(p1, np.eye(3), p2, np.eye(3), trans)
        assert o.pair_decisions(SC, 2.0, 3.0, rec[None])[0] == expected


def test_sphere_periodic_345():
    """SphereTraitsTest.cpp:23-43: spheres of radius 2.5 at distance 5 (3-4-5 through the boundary) touch but do not
    overlap; moved closer they do."""
    o = oracle_lib.oracle()
    box = np.diag([10.0, 10.0, 10.0]).ravel()
    p1, p2 = np.array([1.0, 5.0, 1.0]), np.array([8.0, 1.0, 1.0])  # dx = 3 through x, dy = 4
    trans = np.zeros(3); corr = np.zeros(3); d2 = ctypes.c_double()
    o.lib.orc_pbc(_d(box), _d(p1), _d(p2), _d(corr), _d(trans), ctypes.byref(d2))
    assert d2.value == pytest.approx(25.0)
    assert list(trans) == [-10.0, 0.0, 0.0]
    rec = record(p1, np.eye(3), p2, np.eye(3), trans)
    assert o.pair_decisions(SPHERE, 2.5, 0, rec[None])[0] == 0
    assert o.pair_decisions(SPHERE, 2.51, 0, rec[None])[0] == 1


def test_triclinic_box_heights_volume():
    """TriclinicBoxTest.cpp:10-64 checks volume and heights of a sheared box; here the oracle's volume, heights
    (volume / |cross product of the other two sides|) and inverse on a sheared box against numpy."""
    o = oracle_lib.oracle()
    box = np.array([[10.0, 2.0, 1.0], [0.0, 16.0, 4.0], [0.0, 0.0, 14.0]])
    inv = np.zeros(9); h = np.zeros(3); vol = ctypes.c_double()
    o.lib.orc_box_info(_d(box.ravel()), _d(inv), _d(h), ctypes.byref(vol))
    assert vol.value == pytest.approx(abs(np.linalg.det(box)))
    sides = box.T
    for i in range(3):
        c = np.cross(sides[(i + 1) % 3], sides[(i + 2) % 3])
        assert h[i] == pytest.approx(vol.value / np.linalg.norm(c))
    np.testing.assert_allclose(inv.reshape(3, 3) @ box, np.eye(3), atol=1e-14)


def test_pbc_correction_and_translation():
    """PeriodicBoundaryConditionsTest.cpp:9-83 style checks: correction brings a point into [0,1)^3 in relative
    coordinates; translation is the minimal relative image."""
    o = oracle_lib.oracle()
    box = np.array([[10.0, 2.0, 0.0], [0.0, 10.0, 1.0], [0.0, 0.0, 10.0]]).ravel()
    p1 = np.array([-1.0, 11.0, 23.5])
    p2 = np.array([9.5, 0.5, 0.5])
    corr = np.zeros(3); trans = np.zeros(3); d2 = ctypes.c_double()
    o.lib.orc_pbc(_d(box), _d(p1), _d(p2), _d(corr), _d(trans), ctypes.byref(d2))
    b = box.reshape(3, 3)
    rel = np.linalg.solve(b, p1 + corr)
    assert np.all(rel >= 0) and np.all(rel < 1)
    assert np.allclose(np.linalg.solve(b, corr), np.round(np.linalg.solve(b, corr)))
    drel = np.linalg.solve(b, p2 + trans) - np.linalg.solve(b, p1)
    assert np.all(np.abs(drel) <= 0.5 + 1e-12)


def _brute_segment_distance2(a, b, c, d, n=401):
    s = np.linspace(0, 1, n, dtype=np.longdouble)
    p = a[None, :] + s[:, None] * (b - a)[None, :]
    q = c[None, :] + s[:, None] * (d - c)[None, :]
    diff = p[:, None, :] - q[None, :, :]
    best = (diff * diff).sum(-1).min()
    return float(best)


def test_segment_distance_brute_force_and_special_cases():
    o = oracle_lib.oracle()
    rng = np.random.Generator(np.random.PCG64(3))
    pts = rng.random((40, 12)) * 4 - 2
    out = np.zeros(len(pts))
    o.lib.orc_segment_distance2(len(pts), _d(pts), _d(out))
    for row, v in zip(pts, out):
        a, b, c, d = (np.asarray(row[3 * i:3 * i + 3], dtype=np.longdouble) for i in range(4))
        brute = _brute_segment_distance2(a, b, c, d)
        # the grid minimum is an upper bound within the grid resolution
        assert v <= brute + 1e-12
        assert v >= brute - 0.05
    # parallel segments offset by 1: distance 1; crossing segments: 0; collinear gap of 2: 4
    special = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0],
                        [-1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0],
                        [0, 0, 0, 1, 0, 0, 3, 0, 0, 4, 0, 0],
                        [0, 0, 0, 0, 0, 2, 1, 0, 1, 3, 0, 1]], dtype=float)
    out = np.zeros(len(special))
    o.lib.orc_segment_distance2(len(special), _d(special), _d(out))
    assert list(out) == [1.0, 0.0, 4.0, 1.0]


def test_rotation_matrix_is_orthonormal_and_rodrigues():
    o = oracle_lib.oracle()
    axis = np.array([1.0, 2.0, 2.0]) / 3.0
    R = np.zeros(9)
    o.lib.orc_rotation(_d(axis), 0.7, _d(R))
    R = R.reshape(3, 3)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
    np.testing.assert_allclose(R @ axis, axis, atol=1e-14)
    assert np.trace(R) == pytest.approx(1 + 2 * math.cos(0.7))


def test_is_box_upscaled():
    o = oracle_lib.oracle()
    a = np.diag([10.0, 10.0, 10.0]).ravel()
    assert o.lib.orc_is_box_upscaled(_d(a), _d(np.diag([10.1, 10.2, 10.3]).ravel())) == 1
    assert o.lib.orc_is_box_upscaled(_d(a), _d(np.diag([10.1, 9.9, 10.3]).ravel())) == 0
    sheared = np.array([[10.5, 0.1, 0], [0, 10.5, 0], [0, 0, 10.5]]).ravel()
    assert o.lib.orc_is_box_upscaled(_d(a), _d(sheared)) == 0


def test_nematic_order_limits():
    o = oracle_lib.oracle()
    aligned = np.tile(np.eye(3).ravel(), (100, 1))
    assert o.lib.orc_nematic_p2(100, _d(aligned)) == pytest.approx(1.0)
    # three perpendicular families: Q = 0
    fam = [np.eye(3), euler_rotation(math.pi / 2, 0, 0), euler_rotation(0, math.pi / 2, 0)]
    rots = np.stack([fam[i % 3].ravel() for i in range(300)])
    assert abs(o.lib.orc_nematic_p2(300, _d(rots))) < 1e-12


def test_count_equals_brute_force_on_small_boxes():
    """Neighbour-grid count == O(N^2) count over all 27 images, including grids 1 and 2 cells thick."""
    import corpus
    o = oracle_lib.oracle()
    for box, n, seed in [(np.diag([7.0, 13.0, 25.0]), 120, 1), (np.diag([12.5, 12.5, 30.0]), 200, 2),
                         (np.array([[13.0, 1.0, 0.5], [0.0, 12.6, -0.8], [0.0, 0.0, 26.0]]), 180, 3)]:
        pos, rot = corpus.random_snapshot(n, seed, box)
        count, _ = o.snapshot_stats(SC, 0.5, 5.0, box, pos, rot)
        brute = o.lib.orc_count_brute(SC, 0.5, 5.0, 0, None, None, _d(box.ravel()), n, _d(pos), _d(rot))
        assert count == brute
        assert count > 0


def test_mean_and_error_closed_form():
    """Quantity::calculateFromSamples (Quantity.cpp:45-62): empty -> (0, 0); one sample -> (x, 0); [1, 2, 3, 4] ->
    mean 2.5 and sqrt(sum (x - 2.5)^2 / (4 * 3)) = sqrt(5 / 12)."""
    import math
    import rampack_amd as ra
    assert ra.mean_and_error([]) == (0.0, 0.0)
    assert ra.mean_and_error([3.25]) == (3.25, 0.0)
    m, e = ra.mean_and_error([1.0, 2.0, 3.0, 4.0])
    assert m == 2.5
    assert abs(e - math.sqrt(5 / 12)) < 1e-16


def test_reference_packing_fraction_closed_form():
    """The reference's Packing::getPackingFraction on a triclinic box equals N v0 / |det(box)| with
    v0 = pi r^2 L + 4/3 pi r^3 (SpherocylinderTraits.cpp:48-50)."""
    import math
    import numpy as np
    import oracle_lib
    if not oracle_lib.ref_available():
        import pytest
        pytest.skip("reference build absent")
    box = np.array([[50.0, 2.0, -1.0], [0.0, 48.0, 1.5], [0.0, 0.0, 52.0]])
    phi = oracle_lib.ref().lib.ref_packing_fraction(1, 0.5, 5.0, oracle_lib._d(box.ravel()), 1000)
    v0 = math.pi * 0.25 * 5.0 + 4 / 3 * math.pi * 0.125
    assert abs(phi - 1000 * v0 / abs(np.linalg.det(box))) < 1e-15


def test_fix_rotation_matches_reference_and_orthonormalises():
    """Packing::fixRotationMatrix (Packing.cpp:1211-1216): the oracle equals the reference build bit for bit on
    perturbed rotations; the result is orthonormal to rounding, and an exact rotation matrix with entries 0 and +-1
    is a fixed point."""
    import numpy as np
    import corpus
    import oracle_lib
    rng = np.random.Generator(np.random.PCG64(3))
    rot = corpus.random_rotations(rng, 2000) + (rng.random((2000, 9)) - 0.5) * 2e-6
    mine = np.ascontiguousarray(rot.copy())
    oracle_lib.oracle().lib.orc_fix_rotation(len(mine), oracle_lib._d(mine))
    if oracle_lib.ref_available():
        theirs = np.ascontiguousarray(rot.copy())
        assert oracle_lib.ref().lib.ref_fix_rotation(len(theirs), oracle_lib._d(theirs)) == 0
        assert np.array_equal(mine, theirs)
    m = mine.reshape(-1, 3, 3)
    dev = np.abs(np.einsum("nij,nkj->nik", m, m) - np.eye(3)).max()
    assert dev < 1e-14
    perm = np.ascontiguousarray(np.array([[0, 1, 0, 0, 0, -1, -1, 0, 0]], dtype=float))
    before = perm.copy()
    oracle_lib.oracle().lib.orc_fix_rotation(1, oracle_lib._d(perm))
    assert np.array_equal(perm, before)


def test_block_error_pairs_neighbours_and_takes_the_plateau():
    """[0, 0, 1, 1] x 8: level 0 (32 values, half 0 half 1) has SE sqrt(8 / (32 * 31)) = sqrt(1 / 124); level 1
    averages neighbours into 16 alternating 0 / 1 blocks with SE sqrt(4 / (16 * 15)) = sqrt(1 / 60); level 2 would
    have 8 blocks (< 16).  The estimate is the larger, sqrt(1 / 60).  (Pairing each value with the one half a series
    away instead would give 16 blocks of 0.5 and SE 0.)"""
    import math
    import rampack_amd as ra
    x = [0.0, 0.0, 1.0, 1.0] * 8
    assert abs(ra.block_error(x) - math.sqrt(1 / 60)) < 1e-15
    m, e = ra.mean_and_error(x, method="blocks")
    assert m == 0.5 and e == ra.block_error(x)
    # alternating values: neighbours average to 0.5, so blocking adds nothing to the naive error
    y = [0.0, 1.0] * 16
    assert abs(ra.block_error(y) - math.sqrt(8 / (32 * 31))) < 1e-15
    assert ra.mean_and_error(y)[1] == ra.mean_and_error(y, method="naive")[1]


def test_block_error_on_correlated_series():
    """AR(1) series x_t = rho x_{t-1} + e_t with rho = 0.9: the standard error of the mean of n samples is
    sigma_x / sqrt(n) * sqrt((1 + rho) / (1 - rho)) for n >> correlation time, 4.36 times the naive value; the block
    estimate lands within 20 % of it (its own scatter at 16+ blocks is about 18 % at the last level), and on
    independent samples it stays within 20 % of the naive one."""
    import numpy as np
    import rampack_amd as ra
    rng = np.random.Generator(np.random.PCG64(11))
    rho, n = 0.9, 1 << 17
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    sigma_x = 1 / np.sqrt(1 - rho * rho)
    exact = sigma_x / np.sqrt(n) * np.sqrt((1 + rho) / (1 - rho))
    naive = ra.mean_and_error(x)[1]
    blocks = ra.block_error(x)
    assert naive < 0.3 * exact
    assert abs(blocks / exact - 1) < 0.2
    iid = rng.normal(size=n)
    assert abs(ra.block_error(iid) / ra.mean_and_error(iid)[1] - 1) < 0.2
