"""RAMTRJ trajectories through the C ABI (host code, no GPU).

Pinned three ways:
* against the reference's own RamtrjRecorder / RamtrjPlayer (src/core/io/Ramtrj*.cpp, compiled from its sources into
  the reference build): byte-identical files, identical matrices read back, the same AutoFix reports;
* by what the format fixes on its own: the 33-byte header laid out field by field, and Euler angles of matrices
  built from known angles as Rz(c) Ry(b) Rx(a) (docs/output-formats.md), which the file must hold as (a, b, c);
* by the reference's own test of the format (test/validation_tests/SimulationIOTest.cpp:22-34, 60-128): a replay
  reproduces the recorded state to 1e-12, also after a continuation.
The reference ships no RAMTRJ fixture file, so these synthetic states stand in for one."""
import os
import struct

import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra

BOX = np.array([[12.5, 0.25, -0.125], [0.0, 11.0, 0.5], [0.0, 0.0, 13.0]])


def need_ref():
    if not oracle_lib.ref_available():
        pytest.skip("reference build absent")
    return oracle_lib.ref()


def states(n, k, seed):
    """k snapshots of n spherocylinder states inside slightly different boxes."""
    boxes, poss, rots = [], [], []
    for j in range(k):
        box = BOX * (1.0 + 0.01 * j)
        pos, rot = corpus.random_snapshot(n, seed + j, box)
        boxes.append(box)
        poss.append(pos)
        rots.append(rot)
    return np.array(boxes), np.array(poss), np.array(rots)


def record(path, step, boxes, poss, rots, append=False, first=None):
    first = step if first is None else first
    with ra.RamtrjRecorder(str(path), poss.shape[1], step, append=append) as rec:
        for j in range(len(boxes)):
            rec.record_snapshot((boxes[j], poss[j], rots[j]), first + j * step)


def rot_xyz(a, b, c):
    """Rz(c) Ry(b) Rx(a) from the textbook elementary rotations."""
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


def test_header_layout_and_euler_angles_of_known_rotations(tmp_path):
    rng = np.random.Generator(np.random.PCG64(3))
    angles = np.column_stack([rng.uniform(-np.pi, np.pi, 50), rng.uniform(-1.5, 1.5, 50),
                              rng.uniform(-np.pi, np.pi, 50)])
    rot = np.array([rot_xyz(*a).ravel() for a in angles])
    rot[0] = np.eye(3).ravel()
    angles[0] = 0
    pos = rng.random((50, 3)) * 10
    path = tmp_path / "k.ramtrj"
    record(path, 7, BOX[None], pos[None], rot[None])
    raw = path.read_bytes()
    assert raw[:7] == b"RAMTRJ\n" and raw[7] == 1 and raw[8] == 1  # minor, then major version
    assert struct.unpack("<QQQ", raw[9:33]) == (50, 1, 7)
    assert len(raw) == 33 + 9 * 8 + 50 * 48
    body = np.frombuffer(raw[33:], dtype="<f8")
    assert np.array_equal(body[:9], BOX.ravel())
    part = body[9:].reshape(50, 6)
    assert np.array_equal(part[:, :3], pos)
    assert np.array_equal(part[0, 3:], [0.0, 0.0, 0.0])  # the identity is exactly zero angles
    np.testing.assert_allclose(part[:, 3:], angles, atol=1e-13, rtol=0)


def test_files_are_byte_identical_to_the_reference_recorder(tmp_path):
    ref = need_ref()
    boxes, poss, rots = states(400, 3, 11)
    rots[0, 5] = np.eye(3).ravel()
    rots[0, 6] = rot_xyz(0.3, np.pi / 2, -1.0).ravel()  # R31 = -1: the reference's general branch still applies
    rots[0, 7] = rot_xyz(0.3, -np.pi / 2, -1.0).ravel()
    ours, theirs = tmp_path / "ours.ramtrj", tmp_path / "theirs.ramtrj"
    record(ours, 100, boxes, poss, rots)
    assert oracle_lib.ref_ramtrj_write(ref, theirs, 100, boxes, poss, rots) == 0
    assert ours.read_bytes() == theirs.read_bytes()


def test_reader_matches_the_reference_player(tmp_path):
    ref = need_ref()
    boxes, poss, rots = states(300, 4, 21)
    path = tmp_path / "t.ramtrj"
    record(path, 50, boxes, poss, rots)
    got, rb, rp, rr, info = oracle_lib.ref_ramtrj_read(ref, path, 300, 4)
    assert got == 4 and list(info[:4]) == [300, 4, 50, 200]
    with ra.RamtrjPlayer(str(path)) as pl:
        assert (pl.get_num_molecules(), pl.get_cycle_step(), pl.get_total_cycles()) == (300, 50, 200)
        j = 0
        while pl.has_next():
            box, pos, rot = pl.next_snapshot()
            assert np.array_equal(box, rb[j].reshape(3, 3))
            assert np.array_equal(pos, rp[j]) and np.array_equal(rot, rr[j])  # the same sin/cos arithmetic
            # SimulationIOTest.cpp:22-34: the replay reproduces the recorded state to 1e-12
            np.testing.assert_allclose(rot, rots[j], atol=1e-12, rtol=0)
            assert np.array_equal(pos, poss[j]) and np.array_equal(box, boxes[j])
            j += 1
            assert pl.get_current_snapshot_cycles() == 50 * j
        assert j == 4
        # jumps, as RamtrjPlayer::jumpToSnapshot / lastSnapshot
        got, jb, jp, jr, _ = oracle_lib.ref_ramtrj_read(ref, path, 300, 1, jump_cycle=100)
        box, pos, rot = pl.jump_to_snapshot(100)
        assert got == 1 and np.array_equal(rot, jr[0]) and np.array_equal(pos, poss[1])
        assert pl.get_current_snapshot_cycles() == 100
        box, pos, rot = pl.next_snapshot()  # reading continues after the jump target
        assert np.array_equal(pos, poss[2])
        box, pos, rot = pl.last_snapshot()
        assert np.array_equal(pos, poss[3]) and not pl.has_next()


def test_reference_file_read_by_this_player(tmp_path):
    ref = need_ref()
    boxes, poss, rots = states(200, 2, 31)
    path = tmp_path / "r.ramtrj"
    assert oracle_lib.ref_ramtrj_write(ref, path, 10, boxes, poss, rots) == 0
    _, rb, rp, rr, _ = oracle_lib.ref_ramtrj_read(ref, path, 200, 2)
    with ra.RamtrjPlayer(str(path)) as pl:
        for j in range(2):
            box, pos, rot = pl.next_snapshot()
            assert np.array_equal(pos, rp[j]) and np.array_equal(rot, rr[j])


def test_continuation_matches_the_reference_and_a_single_recording(tmp_path):
    """SimulationIOTest.cpp:100-128: a recording continued with append=True equals one made in a single run."""
    ref = need_ref()
    boxes, poss, rots = states(100, 4, 41)
    ours, theirs, whole = tmp_path / "a.ramtrj", tmp_path / "b.ramtrj", tmp_path / "w.ramtrj"
    record(ours, 100, boxes[:2], poss[:2], rots[:2])
    rec = ra.RamtrjRecorder(str(ours), 100, 100, append=True)
    assert rec.get_last_cycle_number() == 200
    for j in (2, 3):
        rec.record_snapshot((boxes[j], poss[j], rots[j]), 100 * (j + 1))
    rec.close()
    assert oracle_lib.ref_ramtrj_write(ref, theirs, 100, boxes[:2], poss[:2], rots[:2]) == 0
    assert oracle_lib.ref_ramtrj_write(ref, theirs, 100, boxes[2:], poss[2:], rots[2:], append=True,
                                       first_cycle=300) == 0
    record(whole, 100, boxes, poss, rots)
    assert ours.read_bytes() == theirs.read_bytes() == whole.read_bytes()


def test_continuation_from_zero_snapshots(tmp_path):
    """SimulationIOTest.cpp:130-: a recorder closed before any snapshot can be continued."""
    boxes, poss, rots = states(20, 1, 51)
    path = tmp_path / "z.ramtrj"
    ra.RamtrjRecorder(str(path), 20, 5).close()
    assert len(path.read_bytes()) == 33
    record(path, 5, boxes, poss, rots, append=True)
    with ra.RamtrjPlayer(str(path)) as pl:
        assert pl.num_snapshots == 1 and np.array_equal(pl.next_snapshot()[1], poss[0])


def test_recorder_preconditions(tmp_path):
    boxes, poss, rots = states(10, 1, 61)
    path = tmp_path / "p.ramtrj"
    with pytest.raises(ra.PreconditionError):
        ra.RamtrjRecorder(str(path), 0, 10)
    with pytest.raises(ra.PreconditionError):
        ra.RamtrjRecorder(str(path), 10, 0)
    with ra.RamtrjRecorder(str(path), 10, 10) as rec:
        with pytest.raises(ra.PreconditionError, match="expected cycle 10"):
            rec.record_snapshot((boxes[0], poss[0], rots[0]), 20)
        with pytest.raises(ra.PreconditionError):
            rec.record_snapshot((boxes[0], poss[0, :5], rots[0, :5]), 10)
        bad = rots[0].copy()
        bad[3] *= 1.0 + 1e-11  # not a rotation matrix to 1e-12 (EulerAngles.cpp:47-53)
        with pytest.raises(ra.PreconditionError, match="particle 3"):
            rec.record_snapshot((boxes[0], poss[0], bad), 10)
        rec.record_snapshot((boxes[0], poss[0], rots[0]), 10)
    with pytest.raises(ra.PreconditionError):
        rec.record_snapshot((boxes[0], poss[0], rots[0]), 20)


def test_append_validation(tmp_path):
    boxes, poss, rots = states(10, 2, 71)
    path = tmp_path / "v.ramtrj"
    record(path, 10, boxes, poss, rots)
    with pytest.raises(ra.ValidationError, match="unmatching"):
        ra.RamtrjRecorder(str(path), 11, 10, append=True)
    with pytest.raises(ra.ValidationError, match="unmatching"):
        ra.RamtrjRecorder(str(path), 10, 20, append=True)
    raw = path.read_bytes()
    path.write_bytes(raw[:-8])
    with pytest.raises(ra.ValidationError, match="broken snapshot structure"):
        ra.RamtrjRecorder(str(path), 10, 10, append=True)


def write_header(f, n, snaps, step, major=1, minor=1):
    f.write(b"RAMTRJ\n" + bytes([minor, major]) + struct.pack("<QQQ", n, snaps, step))


@pytest.mark.parametrize("cut,snaps_in_header", [(0, 3), (7, 3), (60 + 72 + 2 * 48, 3), (0, 0), (100, 0)])
def test_truncated_files_and_auto_fix_against_the_reference(tmp_path, cut, snaps_in_header):
    """A recorder that never closed leaves 0 snapshots in the header; a crash can also cut the last snapshot.  The
    strict player rejects both; the auto-fix infers the whole snapshots and reports the rest as the reference's
    AutoFix does."""
    ref = need_ref() if oracle_lib.ref_available() else None
    n = 30
    boxes, poss, rots = states(n, 3, 81)
    path = tmp_path / "c.ramtrj"
    record(path, 10, boxes, poss, rots)
    raw = bytearray(path.read_bytes())
    raw[17:25] = struct.pack("<Q", snaps_in_header)
    if cut:
        raw = raw[:-cut]
    path.write_bytes(bytes(raw))
    consistent = len(raw) == 33 + snaps_in_header * (72 + 48 * n)
    if consistent:
        ra.RamtrjPlayer(str(path)).close()
    else:
        with pytest.raises(ra.ValidationError, match="broken snapshot structure"):
            ra.RamtrjPlayer(str(path))
    pl = ra.RamtrjPlayer(str(path), auto_fix_particles=n)
    body = len(raw) - 33
    expect = dict(fixing_needed=int(not consistent), fixing_successful=1, header_snapshots=snaps_in_header,
                  inferred_snapshots=body // (72 + 48 * n), bytes_remainder=0 if consistent else body % (72 + 48 * n),
                  bytes_per_snapshot=72 + 48 * n)
    assert pl.auto_fix == expect
    assert pl.num_snapshots == expect["inferred_snapshots"]
    k = 0
    while pl.has_next():
        assert np.array_equal(pl.next_snapshot()[1], poss[k])
        k += 1
    assert k == expect["inferred_snapshots"]
    if ref is not None:
        got, _, rp, _, info = oracle_lib.ref_ramtrj_read(ref, path, n, 3, expected_n=n)
        assert got == k and list(info[4:]) == [expect["fixing_needed"], 1, k, expect["bytes_remainder"]]
        for j in range(k):
            assert np.array_equal(rp[j], poss[j])


def test_auto_fix_errors(tmp_path):
    n = 5
    boxes, poss, rots = states(n, 2, 91)
    path = tmp_path / "e.ramtrj"
    record(path, 10, boxes, poss, rots)
    raw = path.read_bytes()
    path.write_bytes(raw[:-3])
    with pytest.raises(ra.ValidationError, match=r"Expected number of molecules \(6\)"):
        ra.RamtrjPlayer(str(path), auto_fix_particles=6)
    # a v1.0 header without the particle count cannot be fixed
    v10 = tmp_path / "v10.ramtrj"
    with open(v10, "wb") as f:
        write_header(f, 0, 0, 0, minor=0)
        f.write(b"\0" * 100)
    with pytest.raises(ra.ValidationError, match="v1.0"):
        ra.RamtrjPlayer(str(v10), auto_fix_particles=n)
    if oracle_lib.ref_available():
        for p, expected in ((path, 6), (v10, n)):
            got, *_, info = oracle_lib.ref_ramtrj_read(oracle_lib.ref(), p, n, 1, expected_n=expected)
            assert got == -1 and info[4] == 0 and info[5] == 0


@pytest.mark.parametrize("header,what", [
    (b"RAMTRX\n\x01\x01" + b"\0" * 24, "magic"),
    (b"RAMTRJ\n\x01\x02" + b"\0" * 24, "versions"),
    (b"RAMTRJ\n\x02\x01" + b"\0" * 24, "versions"),
    (b"RAMTRJ\n\x01\x01" + b"\0" * 10, "num snapshots"),
    (b"RAMTRJ\n\x01\x01" + struct.pack("<QQQ", 0, 0, 1), "num particles"),
    (b"RAMTRJ\n\x01\x01" + struct.pack("<QQQ", 1, 0, 0), "cycle step"),
])
def test_broken_headers(tmp_path, header, what):
    path = tmp_path / "h.ramtrj"
    path.write_bytes(header)
    with pytest.raises(ra.ValidationError, match=what):
        ra.RamtrjPlayer(str(path))
    if oracle_lib.ref_available():
        assert oracle_lib.ref_ramtrj_read(oracle_lib.ref(), path, 1, 1)[0] == -1


def test_version_1_0_header_with_zero_fields_is_accepted(tmp_path):
    path = tmp_path / "old.ramtrj"
    with open(path, "wb") as f:
        write_header(f, 0, 0, 0, minor=0)
    with ra.RamtrjPlayer(str(path)) as pl:
        assert pl.version == (1, 0) and pl.num_snapshots == 0 and not pl.has_next()


def test_player_preconditions(tmp_path):
    boxes, poss, rots = states(8, 2, 101)
    path = tmp_path / "j.ramtrj"
    record(path, 25, boxes, poss, rots)
    with ra.RamtrjPlayer(str(path)) as pl:
        for cycle in (0, 30, 75):
            with pytest.raises(ra.PreconditionError):
                pl.jump_to_snapshot(cycle)
        pl.last_snapshot()
        with pytest.raises(ra.PreconditionError, match="no more snapshots"):
            pl.next_snapshot()
    empty = tmp_path / "empty.ramtrj"
    ra.RamtrjRecorder(str(empty), 8, 25).close()
    with ra.RamtrjPlayer(str(empty)) as pl:
        with pytest.raises(ra.PreconditionError, match="empty"):
            pl.last_snapshot()
    with pytest.raises(ra.ValidationError, match="Could not open"):
        ra.RamtrjPlayer(str(tmp_path / "missing.ramtrj"))
