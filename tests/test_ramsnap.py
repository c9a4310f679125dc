"""RAMSNAP snapshots through the C ABI (host code, no GPU): the reference's own fixture files
(test/validation_tests/data/packing_loader/*.ramsnap, copied to tests/golden/ramsnap), byte-identical output to the
reference's RamsnapWriter, round trips through the reference's RamsnapReader in both directions, the older cuboid
box line, and broken files (RamsnapException is a ValidationException)."""
import os

import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "ramsnap")


def need_ref():
    if not oracle_lib.ref_available():
        pytest.skip("reference build absent")
    return oracle_lib.ref()


@pytest.mark.parametrize("name,cycles", [("packing_finished.ramsnap", "20000"), ("packing_unfinished.ramsnap", "3000")])
def test_reads_reference_fixtures(name, cycles):
    box, pos, rot, aux = ra.ramsnap_read(os.path.join(FIXTURES, name))
    assert np.array_equal(box, np.diag([10.0, 10.0, 10.0]))
    assert np.array_equal(pos, np.zeros((1, 3)))
    assert np.array_equal(rot, np.eye(3).reshape(1, 9))
    assert aux == {"cycles": cycles}


def random_state(n, seed):
    box = np.array([[12.5, 0.25, -0.125], [0.0, 11.0, 0.5], [0.0, 0.0, 13.0]])
    pos, rot = corpus.random_snapshot(n, seed, box)
    return box, pos, rot


def test_output_is_byte_identical_to_the_reference_writer(tmp_path):
    ref = need_ref()
    box, pos, rot = random_state(300, 5)
    rot[0, :3] = [0.0, -0.0, 1e-300]  # a signed zero and a tiny value keep their text (orientations are not checked)
    aux = "cycles 1234\ntranslation_step 0.10000000000000001\nlabel two words here\n"
    ours, theirs = tmp_path / "ours.ramsnap", tmp_path / "theirs.ramsnap"
    ra.ramsnap_write(str(ours), box, pos, rot, {"cycles": "1234", "translation_step": "0.10000000000000001",
                                                 "label": "two words here"})
    assert ref.lib.ref_ramsnap_write(str(theirs).encode(), oracle_lib._d(box.ravel()), len(pos), oracle_lib._d(pos),
                                     oracle_lib._d(rot), aux.encode()) == 0
    assert ours.read_bytes() == theirs.read_bytes()


def test_round_trips_with_the_reference_reader(tmp_path):
    ref = need_ref()
    box, pos, rot = random_state(500, 6)
    path = tmp_path / "s.ramsnap"
    ra.ramsnap_write(str(path), box, pos, rot, {"k": "v"})
    b2, p2, r2 = np.zeros(9), np.zeros_like(pos), np.zeros_like(rot)
    assert ref.lib.ref_ramsnap_read(str(path).encode(), oracle_lib._d(b2), len(pos), oracle_lib._d(p2),
                                    oracle_lib._d(r2)) == len(pos)
    assert np.array_equal(b2.reshape(3, 3), box) and np.array_equal(p2, pos) and np.array_equal(r2, rot)
    # and the reference's file read by this reader
    path2 = tmp_path / "r.ramsnap"
    assert ref.lib.ref_ramsnap_write(str(path2).encode(), oracle_lib._d(box.ravel()), len(pos), oracle_lib._d(pos),
                                     oracle_lib._d(rot), b"a 1\n") == 0
    b3, p3, r3, aux = ra.ramsnap_read(str(path2))
    assert np.array_equal(b3, box) and np.array_equal(p3, pos) and np.array_equal(r3, rot) and aux == {"a": "1"}


def test_round_trip_is_exact_for_every_double(tmp_path):
    rng = np.random.Generator(np.random.PCG64(9))
    box = np.diag([7.0, 8.0, 9.0])
    pos = rng.standard_normal((200, 3)) * 10.0 ** rng.integers(-300, 300, (200, 3))
    rot = rng.standard_normal((200, 9))
    path = tmp_path / "x.ramsnap"
    ra.ramsnap_write(str(path), box, pos, rot)
    b, p, r, aux = ra.ramsnap_read(str(path))
    assert np.array_equal(p, pos) and np.array_equal(r, rot) and aux == {}


def test_snapshot_without_aux_entries(tmp_path):
    """The reference's reader rejects a file with no auxiliary entries (its box line starts at the end of the count
    line), even one its own writer produced; this reader accepts it."""
    path = tmp_path / "noaux.ramsnap"
    path.write_text("0\n4 0 0 0 5 0 0 0 6\n1\n1 1 1        1 0 0 0 1 0 0 0 1\n")
    box, pos, rot, aux = ra.ramsnap_read(str(path))
    assert np.array_equal(box, np.diag([4.0, 5.0, 6.0])) and aux == {}
    if oracle_lib.ref_available():
        b, p, r = np.zeros(9), np.zeros(3), np.zeros(9)
        assert oracle_lib.ref().lib.ref_ramsnap_read(str(path).encode(), oracle_lib._d(b), 1, oracle_lib._d(p),
                                                      oracle_lib._d(r)) == -1


def test_older_cuboid_box_line_and_empty_packing(tmp_path):
    path = tmp_path / "old.ramsnap"
    path.write_text("0\n4 5 6\n2\n1 1 1        1 0 0 0 1 0 0 0 1\n2 2 2        0 1 0 -1 0 0 0 0 1\n")
    box, pos, rot, aux = ra.ramsnap_read(str(path))
    assert np.array_equal(box, np.diag([4.0, 5.0, 6.0]))
    assert np.array_equal(pos, [[1, 1, 1], [2, 2, 2]])
    assert np.array_equal(rot[1], [0, 1, 0, -1, 0, 0, 0, 0, 1])
    empty = tmp_path / "empty.ramsnap"
    ra.ramsnap_write(str(empty), np.diag([3.0, 3.0, 3.0]), np.zeros((0, 3)), np.zeros((0, 9)), {"a": "b c"})
    box, pos, rot, aux = ra.ramsnap_read(str(empty))
    assert len(pos) == 0 and aux == {"a": "b c"}


@pytest.mark.parametrize("text,what", [
    ("x\n", "aux info size"),
    ("1\n", "aux info entry"),
    ("0\n1 2\n", "dimensions"),
    ("0\n1 0 0 0 1 0 0 0 0\n0\n", "zero box volume"),
    ("0\n1 1 1\n", "size"),
    ("0\n1 1 1\n2\n0 0 0 1 0 0 0 1 0 0 0 1\n", "shape 1/2"),
])
def test_broken_files_are_validation_errors(tmp_path, text, what):
    path = tmp_path / "broken.ramsnap"
    path.write_text(text)
    with pytest.raises(ra.ValidationError, match=what):
        ra.ramsnap_read(str(path))


def test_aux_keys_must_be_single_tokens(tmp_path):
    with pytest.raises(ValueError):
        ra.ramsnap_write(str(tmp_path / "k.ramsnap"), np.eye(3), np.zeros((0, 3)), np.zeros((0, 9)), {"a b": "c"})
