"""Several move samplers (TranslationSampler, RotationSampler, RototranslationSampler; Simulation.cpp:447-491): a
translation sampler leaves orientations bit for bit and a rotation sampler positions, each cycle makes one N-move
share per sampler with the sampler drawn per move, each sampler adapts its own steps, and the reference's
"translation + rotation moves" validation case reproduces its density."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice
from test_gpu_sweep import overlaps_by_oracle, rod_start

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)


def rods():
    box, pos, _ = rod_start()
    rng = np.random.Generator(np.random.PCG64(3))
    # small random tilts keep the lattice overlap-free while making every orientation entry generic
    rot = np.stack([lattice_tilt(rng) for _ in range(len(pos))])
    return box, pos, rot


def lattice_tilt(rng):
    a = rng.normal(size=3) * 0.01
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + K + 0.5 * K @ K
    u, _, vt = np.linalg.svd(R)
    return (u @ vt).ravel()


def test_translation_sampler_keeps_orientations_bit_for_bit():
    box, pos, rot = rods()
    assert overlaps_by_oracle(SC, box, pos, rot) == 0
    pk = ra.GpuPacking(SC, box, pos, rot, seed=5)
    pk.set_environment(1.0, 2.0, 0.3, None)
    pk.set_move_samplers([("translation", 0.3)])
    c = pk.run_cycles(30, adapt_steps=True)
    b, p, r = pk.download()
    assert np.array_equal(r, rot)
    assert not np.allclose(p, pos) and c["moves_accepted"] > 0
    assert overlaps_by_oracle(SC, b, p, r) == 0
    s = pk.move_samplers()
    assert s[0]["kind"] == "translation" and s[0]["rot_step"] == 0


def test_rotation_sampler_keeps_positions_bit_for_bit():
    box, pos, rot = rods()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=6)
    pk.set_environment(1.0, 2.0, 0.3, None)
    pk.set_move_samplers([("rotation", 0.2)])
    c = pk.run_cycles(30, adapt_steps=True)
    b, p, r = pk.download()
    assert np.array_equal(p, pos)
    assert not np.allclose(r, rot) and c["moves_accepted"] > 0
    R = r.reshape(-1, 3, 3)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-12)
    assert overlaps_by_oracle(SC, b, p, r) == 0


def test_two_samplers_share_the_moves_and_adapt_separately():
    box, pos, rot = rods()
    pk = ra.GpuPacking(SC, box, pos, rot, seed=7)
    pk.set_environment(1.0, 2.0, 0.3, None, box_move=("delta_triclinic", 0.01))
    pk.set_move_samplers([("translation", 0.05), ("rotation", 3.0)])
    c = pk.run_cycles(400, adapt_steps=True)
    s = pk.move_samplers()
    print(c, s, flush=True)
    # two N-move shares per cycle (dense cells make round(N / cells) moves per share)
    assert abs(c["moves_attempted"] - 2 * 400 * len(pos)) <= 0.15 * 2 * 400 * len(pos)
    # a small translation step is accepted too often and grows; a large rotation step too rarely and shrinks
    assert s[0]["trans_step"] > 0.05 and s[1]["rot_step"] < 3.0
    b, p, r = pk.download()
    assert overlaps_by_oracle(SC, b, p, r) == 0
    with pytest.raises(ra.PreconditionError):
        pk.set_move_samplers([])
    with pytest.raises(ra.PreconditionError):
        pk.set_move_samplers([("rotation", 0.0)])
    with pytest.raises(ra.PreconditionError):
        pk.set_move_samplers([("translation", 0.1)] * 5)


def test_degenerate_spherocylinder_gas_with_translation_and_rotation_samplers():
    """SimulationTest.cpp:134-146 ("translation + rotation moves"): TranslationSampler(0.5) + RotationSampler(1)
    instead of RototranslationSampler(0.5, 1), the same expected density 0.09454087; protocol and block errors as
    the rototranslation case's test."""
    box, pos, rot = lattice.orthorhombic_arrange(50, (10, 10, 10), (5, 5, 2), (2, 2, 5))
    shape = ra.Shape("spherocylinder", radius=0.5, length=3.0)
    pk = ra.GpuPacking(shape, box, pos, rot, seed=1234)
    pk.set_environment(1.0, 0.5, 0.5, 1.0, box_move=("delta_v", 10.0))
    pk.set_move_samplers([("translation", 0.5), ("rotation", 1.0)])
    pk.run_cycles(5000, adapt_steps=True)
    rho = []
    for _ in range(600):
        pk.run_cycles(100)
        rho.append(pk.observables()["number_density"])
    m = float(np.mean(rho))
    block_err = float(np.std(np.reshape(rho, (10, -1)).mean(axis=1), ddof=1) / np.sqrt(10))
    print(f"spherocylinder gas, translation + rotation: {m} +- {block_err} (reference 0.09454087)", flush=True)
    assert abs(m - 0.09454087) <= 3 * block_err + 0.001 * 0.09454087
    assert block_err / m < 0.01
