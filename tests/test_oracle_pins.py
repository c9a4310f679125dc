"""The oracle's box, move and rescale helpers pinned bit for bit against the reference's own code (oracle/_ref).

Each helper is the expected value of a device parity test, so each is compared with the compiled reference on inputs
that exercise its branches:
  * orc_box_info vs TriclinicBox (its inverse read back through absoluteToRelative, getHeights, getVolume);
  * orc_translate_wrapped vs Shape::translate with PeriodicBoundaryConditions (positions leaving the box on every
    face, several boxes of the box's size away, and staying inside);
  * orc_rotation vs Matrix<3,3>::rotation, including the sin / cos special cases 0 and 1;
  * orc_rotate vs Shape::rotate;
  * orc_rescale vs Packing::tryScaling itself (the positions it leaves, Packing.cpp:183-184).
"""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.skipif(not oracle_lib.ref_available(), reason="oracle/_ref is built only beside the "
                                                                       "reference sources")
_dp = oracle_lib._dp


def P(a):
    return a.ctypes.data_as(_dp)


def boxes(seed, n):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = [np.diag([10.0, 12.5, 7.25]), np.eye(3) * 98.91, np.diag([1.0, 1.0, 1.0])]
    for _ in range(n):
        b = np.diag(rng.uniform(5, 50, 3)) + np.triu(rng.uniform(-3, 3, (3, 3)), 1)
        if rng.random() < 0.5:
            b += np.tril(rng.uniform(-2, 2, (3, 3)), -1)
        out.append(b)
    return [np.ascontiguousarray(b.reshape(9)) for b in out]


def test_box_inverse_heights_volume_equal_reference():
    """The oracle's and the product's (rp_box_info, host code that builds every device grid) box quantities."""
    import rampack_amd as ra
    orc, ref = oracle_lib.oracle().lib, oracle_lib.ref().lib
    for b in boxes(1, 200):
        got = [np.zeros(9), np.zeros(3), np.zeros(1)]
        want = [np.zeros(9), np.zeros(3), np.zeros(1)]
        orc.orc_box_info(P(b), *[P(x) for x in got])
        ref.ref_box_info(P(b), *[P(x) for x in want])
        for g, w in zip(got, want):
            assert np.array_equal(g, w), (b, g, w)
        inv, h, v = ra.box_info(b)
        assert np.array_equal(inv.reshape(9), want[0]) and np.array_equal(h, want[1]) and v == want[2][0]


def test_translate_wrapped_equals_reference():
    orc, ref = oracle_lib.oracle().lib, oracle_lib.ref().lib
    rng = np.random.Generator(np.random.PCG64(2))
    for b in boxes(3, 40):
        m = b.reshape(3, 3)
        n = 5000
        rel = rng.random((n, 3))
        pos = np.ascontiguousarray(rel @ m.T)
        # small steps (most stay inside, some cross a face), steps of a few box lengths, exact face hits
        t = rng.uniform(-0.3, 0.3, (n, 3))
        t[: n // 5] = rng.uniform(-3, 3, (n // 5, 3)) @ m.T
        t[n // 5: n // 5 + 50] = -pos[n // 5: n // 5 + 50]
        t = np.ascontiguousarray(t)
        got, want = np.zeros((n, 3)), np.zeros((n, 3))
        orc.orc_translate_wrapped(P(b), n, P(pos), P(t), P(got))
        ref.ref_translate(P(b), n, P(pos), P(t), P(want))
        assert np.array_equal(got, want)
        crossed = np.any((np.linalg.solve(m, (pos + t).T).T < 0) | (np.linalg.solve(m, (pos + t).T).T >= 1), axis=1)
        assert crossed.sum() > n // 10


def test_rotation_and_rotate_equal_reference():
    orc, ref = oracle_lib.oracle().lib, oracle_lib.ref().lib
    rng = np.random.Generator(np.random.PCG64(4))
    n = 20000
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1)[:, None]
    axis = np.ascontiguousarray(axis)
    angle = rng.uniform(-np.pi, np.pi, n)
    angle[:4] = [0.0, np.pi / 2, -np.pi / 2, np.pi]  # sin 0 / +-1 and cos 0 / 1: Matrix's scalar special cases
    want = np.zeros((n, 9))
    ref.ref_rotation(n, P(axis), P(angle), 0, P(want))
    got = np.zeros((n, 9))
    for i in range(n):
        orc.orc_rotation(P(axis[i]), float(angle[i]), P(got[i]))
    assert np.array_equal(got, want)
    # the sampler's axis.normalized(): an axis of norm 1 is left unchanged, so both forms agree on unit axes
    want2 = np.zeros((n, 9))
    raw = np.ascontiguousarray(axis * rng.uniform(0.2, 1.0, n)[:, None])
    ref.ref_rotation(n, P(raw), P(angle), 1, P(want2))
    unit = np.ascontiguousarray(raw / np.sqrt((raw * raw).sum(axis=1))[:, None])
    got2 = np.zeros((n, 9))
    for i in range(0, n, 7):
        orc.orc_rotation(P(unit[i]), float(angle[i]), P(got2[i]))
    # np.sqrt of the same sum is the same double; numpy's row sum of three squares is ((x^2 + y^2) + z^2)
    assert np.array_equal(got2[::7], want2[::7])
    # R O
    O = np.ascontiguousarray(want[::-1])
    r1, r2 = np.zeros((n, 9)), np.zeros((n, 9))
    orc.orc_rotate(n, P(want), P(O), P(r1))
    ref.ref_rotate(n, P(want), P(O), P(r2))
    assert np.array_equal(r1, r2)


@pytest.mark.parametrize("kind", [0, 1])
def test_rescale_equals_reference_try_scaling(kind):
    orc, ref = oracle_lib.oracle().lib, oracle_lib.ref().lib
    rng = np.random.Generator(np.random.PCG64(5 + kind))
    radius, length = (0.05, 0.0) if kind == 0 else (0.05, 0.2)
    for b in boxes(6 + kind, 20):
        m = b.reshape(3, 3)
        n = 400
        pos = np.ascontiguousarray(rng.random((n, 3)) @ m.T)
        rot = np.ascontiguousarray(np.tile(np.eye(3).reshape(9), (n, 1)))
        nb = np.ascontiguousarray((m @ (np.eye(3) + rng.uniform(-0.05, 0.05, (3, 3)))).reshape(9))
        got, want = np.zeros((n, 3)), np.zeros((n, 3))
        orc.orc_rescale(P(b), P(nb), n, P(pos), P(got))
        dE = np.zeros(1)
        assert ref.ref_try_scaling(kind, radius, length, P(b), P(nb), n, P(pos), P(rot), P(want), P(dE)) == 0
        assert np.array_equal(got, want)
