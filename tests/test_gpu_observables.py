"""Device observables (rp_get_observables) against the reference build: PackingFraction bit-exact, NematicOrder within
the rounding of an N-term sum.  The Q tensor is also checked against numpy's float64 eigen-decomposition, and P2 on
states whose value the mathematics fixes (perfect alignment: P2 = 1; axes along x, y, z in equal numbers: Q = 0)."""
import numpy as np
import pytest

import corpus
import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
BOX = np.array([[50.0, 2.0, -1.0], [0.0, 48.0, 1.5], [0.0, 0.0, 52.0]])


def need_ref():
    if not oracle_lib.ref_available():
        pytest.skip("oracle/_ref/libref.so not built")
    return oracle_lib.ref()


def sparse_state(n, seed):
    pos, rot = corpus.random_snapshot(n, seed, BOX)
    return pos, rot


@pytest.mark.parametrize("n,seed", [(1, 1), (7, 2), (3000, 3), (20000, 4)])
def test_packing_fraction_and_p2_match_reference(n, seed):
    ref = need_ref()
    pos, rot = sparse_state(n, seed)
    pk = ra.GpuPacking(SC, BOX, pos, rot)
    o = pk.observables()
    phi_ref = ref.lib.ref_packing_fraction(1, 0.5, 5.0, oracle_lib._d(BOX.ravel()), n)
    assert o["packing_fraction"] == phi_ref
    assert o["number_density"] == pytest.approx(n / abs(np.linalg.det(BOX)), rel=1e-14)
    p2_ref = ref.lib.ref_nematic_p2(n, oracle_lib._d(rot.ravel()))
    # Q entries are means of n products in [-1, 1]: the reference's sequential sum and the device tree sum both lie
    # within n * 2^-53 of the exact mean; Cardano amplifies at most by the acos slope, bounded on these states
    assert abs(o["nematic_p2"] - p2_ref) <= 1e-9, (o["nematic_p2"], p2_ref)
    pk.close()


def test_q_tensor_matches_float64_eigendecomposition():
    pos, rot = sparse_state(20000, 5)
    pk = ra.GpuPacking(SC, BOX, pos, rot)
    o = pk.observables()
    a = rot.reshape(-1, 3, 3)[:, :, 2]
    q = 0.5 * (3 * (a.T @ a) / len(a) - np.eye(3))
    up = [q[0, 0], q[0, 1], q[0, 2], q[1, 1], q[1, 2], q[2, 2]]
    np.testing.assert_allclose(o["q_tensor"], up, atol=1e-13)
    w = np.linalg.eigvalsh(q)
    assert abs(o["nematic_p2"] - w[np.argmax(np.abs(w))]) < 1e-10
    pk.close()


def test_p2_of_aligned_rods_is_one_and_of_three_axis_mixture_is_zero():
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(6, 6, 2), spacing=(1.6, 1.6, 6.4))
    pk = ra.GpuPacking(SC, box, pos, rot)
    assert abs(pk.observables()["nematic_p2"] - 1.0) < 1e-14
    pk.close()
    # axes along x, y and z in equal numbers: <a a^T> = I/3, Q = 0
    rots = []
    for i in range(len(pos)):
        k = i % 3
        perm = np.roll(np.eye(3), k, axis=1)  # column 2 = e_{(2 + k) % 3}
        rots.append(perm.ravel())
    rot3 = np.array(rots)
    box3 = np.diag([40.0, 40.0, 40.0])
    pos3 = np.stack(np.meshgrid(*[np.arange(6) * 6.5 + 1] * 3, indexing="ij"), -1).reshape(-1, 3)[: len(rot3)]
    pk = ra.GpuPacking(SC, box3, pos3, rot3[: len(pos3)])
    o = pk.observables()
    assert np.allclose(o["q_tensor"], 0, atol=1e-15)
    assert o["nematic_p2"] == 0.0
    pk.close()


def test_spheres_have_packing_fraction_and_no_nematic_order():
    box, pos, rot = lattice.hard_sphere_lattice(10, 0.30)
    pk = ra.GpuPacking(ra.Shape("sphere", radius=0.5), box, pos, rot)
    o = pk.observables()
    assert abs(o["packing_fraction"] - 0.30) < 1e-12
    assert np.isnan(o["nematic_p2"])
    pk.close()


def test_observables_follow_the_state_during_a_run():
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(12, 12, 3), spacing=(1.45, 1.45, 6.3))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=3)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
    pk.run_cycles(50, adapt_steps=True)
    b, p, r = pk.download()
    o = pk.observables()
    assert o["packing_fraction"] == pytest.approx(len(p) * SC.volume / abs(np.linalg.det(b)), rel=1e-14)
    a = r.reshape(-1, 3, 3)[:, :, 2]
    q = 0.5 * (3 * (a.T @ a) / len(a) - np.eye(3))
    w = np.linalg.eigvalsh(q)
    assert abs(o["nematic_p2"] - w[np.argmax(np.abs(w))]) < 1e-10
    assert 0 < o["nematic_p2"] < 1
    pk.close()


def test_ramsnap_save_of_the_device_state_reads_back_exactly(tmp_path):
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(8, 8, 3), spacing=(1.6, 1.6, 6.4))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=2)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
    pk.run_cycles(20)
    path = tmp_path / "state.ramsnap"
    pk.save_ramsnap(str(path), {"cycles": 20})
    b, p, r = pk.download()
    b2, p2, r2, aux = ra.ramsnap_read(str(path))
    assert np.array_equal(b, b2) and np.array_equal(p, p2) and np.array_equal(r, r2)
    assert aux == {"cycles": "20"}
    pk.close()


def test_ramtrj_recording_of_a_running_packing_replays_it(tmp_path):
    """RamtrjRecorder::recordSnapshot on the device-resident state every 10 cycles of an NpT run (as Simulation
    records, Simulation.cpp:146-150): the file is byte-identical to the reference's recorder given the downloaded
    states, and the replay matches each state to 1e-12 (SimulationIOTest.cpp:22-34), positions and boxes exactly."""
    box, pos, rot = lattice.spherocylinder_lattice(n_side=(8, 8, 3), spacing=(1.6, 1.6, 6.4))
    pk = ra.GpuPacking(SC, box, pos, rot, seed=3)
    pk.set_environment(1.0, 2.0, 0.2, None, box_move=("delta_triclinic", 0.002))
    path = tmp_path / "run.ramtrj"
    states = []
    with ra.RamtrjRecorder(str(path), len(pos), 10) as rec:
        for k in range(1, 5):
            pk.run_cycles(10)
            rec.record_snapshot(pk, 10 * k)
            states.append(pk.download())
        assert rec.get_last_cycle_number() == 40
    with ra.RamtrjPlayer(str(path)) as pl:
        assert pl.get_total_cycles() == 40 and pl.num_snapshots == 4
        for b, p, r in states:
            b2, p2, r2 = pl.next_snapshot()
            assert np.array_equal(b, b2) and np.array_equal(p, p2)
            np.testing.assert_allclose(r2, r, atol=1e-12, rtol=0)
    if oracle_lib.ref_available():
        theirs = tmp_path / "ref.ramtrj"
        assert oracle_lib.ref_ramtrj_write(oracle_lib.ref(), theirs, 10, np.array([s[0] for s in states]),
                                           np.array([s[1] for s in states]), np.array([s[2] for s in states])) == 0
        assert path.read_bytes() == theirs.read_bytes()
    pk.close()
