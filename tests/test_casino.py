"""The PYON reader and the casino driver's input handling (docs/input-file.md, docs/initial-arrangement.md), on
the CPU: literals and objects as the format describes them, the lattice dimensions of LatticeDimensionsOptimizer,
the scaling specs of AnisotropicVolumeScaler, and inputs outside the supported subset rejected by name."""
import numpy as np
import pytest

from rampack_amd import casino, pyon
from rampack_amd.pyon import Call


def test_pyon_literals_and_objects():
    v = pyon.parse('''
        # a comment line
        foo(1, -5, 0xFF, 0b101, 1.3, -.5, 5.13e-100, True, False, None,  # inline comment
            "esc: \\" \\\\ \\n \\t # not a comment", [1, "abc", None], {"abc": 1, "d": [1, 2]},
            bar, baz(), qux(a=1, b="x"), k = 3)''')
    assert isinstance(v, Call) and v.name == "foo"
    assert v.args[:10] == [1, -5, 255, 5, 1.3, -0.5, 5.13e-100, True, False, None]
    assert v.args[10] == 'esc: " \\ \n \t # not a comment'
    assert v.args[11] == [1, "abc", None] and v.args[12] == {"abc": 1, "d": [1, 2]}
    assert v.args[13] == Call("bar") and v.args[14] == Call("baz") and v.args[15] == Call("qux", [], {"a": 1, "b": "x"})
    assert v.kwargs == {"k": 3}
    for bad in ['foo(a=1, 2)', '"unterminated', 'foo(1', '[1, 2] 3', 'foo(a=1, a=2)', '{1: 2}']:
        with pytest.raises(pyon.PyonError):
            pyon.parse(bad)


def test_lattice_dimensions_optimizer():
    """LatticeDimensionsOptimizer.cpp:10-37: cells of the box's mean size, rounded, grown along the longest side
    until every shape fits; the whole lattice is filled (FullPopulator)."""
    assert casino._optimize_dims(1000, np.diag([100.0] * 3)) == [10, 10, 10]
    assert casino._optimize_dims(50, np.diag([10.0] * 3)) == [4, 4, 4]
    assert casino._optimize_dims(8000, np.diag([52.22] * 3)) == [20, 20, 20]
    d = casino._optimize_dims(100, np.diag([10.0, 10.0, 40.0]))
    assert np.prod(d) >= 100 and d[2] > d[0]
    box, pos, rot = casino.make_arrangement(pyon.parse('lattice(cell=sc, box_dim=[10, 10, 40], n_shapes=100)'), ".")
    assert len(pos) == np.prod(d) and np.allclose(box, np.diag([10.0, 10.0, 40.0]))
    rel = pos @ np.linalg.inv(box).T
    assert (rel > 0).all() and (rel < 1).all()


def test_lattice_forms_and_random_rotations():
    box, pos, rot = casino.make_arrangement(pyon.parse(
        'lattice(cell=sc, cell_dim=[1.5, 1.5, 6.5], n_cells=[4, 4, 2], '
        'transformations=[randomize_rotation(seed=2024)])'), ".")
    assert np.allclose(np.diag(box), [6.0, 6.0, 13.0]) and len(pos) == 32
    R = rot.reshape(-1, 3, 3)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-12)
    assert not np.allclose(R, np.eye(3))
    box, pos, _ = casino.make_arrangement(pyon.parse(
        'lattice(cell=sc, n_cells=[2, 2, 2], box_dim=[[20, 0, 0], [0, 20, 0], [10, 0, 20]])'), ".")
    assert np.allclose(box[:, 2], [10, 0, 20]) and len(pos) == 8


@pytest.mark.parametrize("spec,groups", [("isotropic", [0, 0, 0]), ("anisotropic x", [0, 1, 1]),
                                         ("anisotropic y", [1, 0, 1]), ("anisotropic xyz", [0, 1, 2]),
                                         ("xyz", [0, 1, 2]), ("(xyz)", [0, 0, 0]), ("x(yz)", [0, 1, 1]),
                                         ("(zx)[y]", [0, 3, 0])])
def test_scaling_specs(spec, groups):
    assert casino._scaling_spec(spec) == groups


def test_shapes_moves_and_unsupported_inputs():
    shape, v0 = casino.make_shape(pyon.parse('spherocylinder(l=5, r=0.5)'))
    assert shape.kind == "spherocylinder" and v0 == pytest.approx(np.pi * 0.25 * 5 + 4 / 3 * np.pi * 0.125)
    shape, _ = casino.make_shape(pyon.parse('sphere(r=0.5, interaction=wca(epsilon=1, sigma=1))'))
    assert shape.soft == "wca"
    shape, v0 = casino.make_shape(pyon.parse(
        'polysphere(spheres=[sphere(pos=[0, 0, -0.5], r=0.5), sphere(pos=[0, 0, 0.5], r=0.5)], volume=1.0472)'))
    assert shape.kind == "polysphere" and len(shape.radii) == 2 and v0 == 1.0472
    assert casino._move_sampler(pyon.parse('rototranslation(0.1)')) == ("rototranslation", (0.1, None, 0))
    assert casino._box_move(pyon.parse('delta_triclinic(0.01, scale_together=False)')) == \
        ("delta_triclinic", 0.01, "one_side")
    assert casino._box_move(pyon.parse('disabled')) is None
    for bad, what in [('sphere(r=0.5, interaction=fancy)', "interaction"), ('kmer(3, 0.5, 0.5)', "kmer")]:
        with pytest.raises(casino.InputError, match=what):
            casino.make_shape(pyon.parse(bad))
    with pytest.raises(casino.InputError, match="flip"):
        casino._move_sampler(pyon.parse('flip(every=10)'))
    with pytest.raises(casino.InputError, match="bcc"):
        casino.make_arrangement(pyon.parse('lattice(cell=bcc, box_dim=10, n_shapes=16)'), ".")
    root = pyon.parse('rampack(version="1.2.0", arrangement=lattice(cell=sc, box_dim=10, n_shapes=27), '
                      'shape=sphere(r=0.5), seed=1, runs=[integration(run_name="a", thermalization_cycles=10, '
                      'averaging_cycles=0, snapshot_every=10)], temperature=linear(1, 2))')
    c = casino.Casino(root)
    with pytest.raises(casino.InputError, match="dynamic parameters"):
        c._environment(pyon.parse('integration("a", 10, 0, 10, move_types=[translation(0.1)], '
                                  'box_move_type=disabled)'))
