"""The device's move arithmetic (rp_apply_moves: the functions k_sweep and k_proposals use) against the reference's
own code, bit for bit: Shape::translate with the periodic correction, Matrix<3,3>::rotation of the normalised axis,
Shape::rotate.  The rotation matrix depends on sin and cos of the angle, which the device takes from its own math
library: the matrix must equal the reference formula evaluated with the device's sin and cos bit for bit (the oracle's
restatement, itself pinned to Matrix::rotation), equal the reference's matrix wherever the two libraries agree, and
the libraries must agree to one ulp."""
import numpy as np
import pytest

import oracle_lib
import rampack_amd as ra

pytestmark = pytest.mark.gpu
_dp = oracle_lib._dp


def P(a):
    return a.ctypes.data_as(_dp)


def test_device_move_arithmetic_equals_reference():
    rng = np.random.Generator(np.random.PCG64(11))
    n = 100000
    box = np.array([[98.9, 0.7, -0.4], [0.0, 97.2, 1.1], [0.0, 0.0, 99.6]])
    pos = np.ascontiguousarray(rng.random((n, 3)) @ box.T)
    t = rng.uniform(-0.5, 0.5, (n, 3))
    t[: n // 10] = rng.uniform(-2.5, 2.5, (n // 10, 3)) @ box.T
    t = np.ascontiguousarray(t)
    axis = rng.uniform(-1, 1, (n, 3))
    axis = np.ascontiguousarray(axis[(axis * axis).sum(axis=1) <= 1][: n // 2])
    m = len(axis)
    angle = np.ascontiguousarray(rng.uniform(-np.pi, np.pi, m))
    angle[:3] = [0.0, np.pi / 2, np.pi]
    O = np.zeros((m, 9))
    oracle_lib.ref().lib.ref_rotation(m, P(axis[::-1].copy()), P(angle[::-1].copy()), 1, P(O))
    npos, R, NO, sc = ra.apply_moves(box, pos[:m], O, t[:m], axis, angle)
    ref = oracle_lib.ref().lib
    # translation: Shape::translate + PeriodicBoundaryConditions::getCorrection
    want = np.zeros((m, 3))
    ref.ref_translate(P(np.ascontiguousarray(box.reshape(9))), m, P(pos[:m].copy()), P(t[:m].copy()), P(want))
    assert np.array_equal(npos, want)
    # sin / cos of the device against the C library's (the oracle's std::sin / std::cos): within one ulp
    orc = oracle_lib.oracle().lib
    host = np.zeros((m, 2))
    orc.orc_sincos(m, P(angle), P(host))
    ulps = np.abs(sc - host) / np.spacing(np.abs(host))
    assert np.all(ulps <= 1.0), ulps.max()
    print(f"device sin/cos equal to the C library's for {(sc == host).all(axis=1).mean():.4%} of {m} angles",
          flush=True)
    # the Rodrigues formula with the device's sin / cos: the oracle's restatement (pinned to Matrix::rotation)
    unit = np.ascontiguousarray(axis / np.sqrt((axis * axis).sum(axis=1))[:, None])
    want_r = np.zeros((m, 9))
    for i in range(m):
        orc.orc_rotation_sc(P(unit[i]), float(sc[i, 0]), float(sc[i, 1]), P(want_r[i]))
    assert np.array_equal(R, want_r)
    # and the reference's own matrices (its sin / cos from the C library) to the rounding of those
    ref_r = np.zeros((m, 9))
    ref.ref_rotation(m, P(axis), P(angle), 1, P(ref_r))
    assert np.max(np.abs(R - ref_r)) <= 1e-15
    print(f"rotation matrices bit-identical to the reference's: {(R == ref_r).all(axis=1).mean():.4%}", flush=True)
    # Shape::rotate
    want_o = np.zeros((m, 9))
    ref.ref_rotate(m, P(np.ascontiguousarray(R)), P(O), P(want_o))
    assert np.array_equal(NO, want_o)
