"""The learned candidate ordering on the device.

* The FP32 training kernels' loss and gradient against the oracle's double-precision network (itself pinned by
  finite differences, tests/test_mlp_oracle.py).
* The BF16 MFMA inference against the same network evaluated with BF16-rounded weights and activations (to the
  FP32 accumulation order) and against the exact network (to BF16 accuracy).
* The candidate pairs of the kinked-chain workload: every label is the exact test's decision, checked against the
  bead geometry the features describe wherever the features' FP32 rounding cannot change it.
* Training on the kinked-chain workload at its specification (N = 50 000 chains of 11 touching beads kinked 30
  degrees, packing fraction 0.45, NpT p* = 3) from pairs of the run's own early trajectory: the loss falls, and the
  learned order needs fewer exact tests per trial move than the natural and nearest-first orders while every
  decision stays the same (a decision is whether any candidate overlaps, whatever the order).
"""
import math

import numpy as np
import pytest

import oracle_lib
import rampack_amd as ra
from rampack_amd import lattice
from test_mlp_oracle import random_net

pytestmark = pytest.mark.gpu


def test_training_kernels_gradient_matches_oracle():
    rng = np.random.Generator(np.random.PCG64(5))
    theta = random_net(2)
    x = rng.uniform(-1, 1, (512, 12)).astype(np.float32)
    y = (rng.random(512) < 0.2).astype(np.float32)
    loss, grad = ra.mlp_gradient(theta, x, y)
    loss_o, grad_o, _ = oracle_lib.mlp_loss_grad(theta, x, y)
    assert loss == pytest.approx(loss_o, rel=1e-5)
    scale = np.abs(grad_o).max()
    err = np.abs(grad - grad_o)
    print(f"loss {loss:.6f} (oracle {loss_o:.6f}), max |grad - oracle| {err.max():.3e} of max |grad| {scale:.3e}",
          flush=True)
    assert np.all(err <= 1e-4 * scale + 1e-4 * np.abs(grad_o))


def bf16(a):
    """Round to nearest even BF16, returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1 + np.tanh(math.sqrt(2 / math.pi) * (x + 0.044715 * x ** 3)))


def net_bf16(theta, x):
    """The inference kernel's arithmetic: BF16 inputs, weights and hidden activations, exact products, sums and
    biases in float64 (the device accumulates in FP32), the output from the unrounded last hidden layer."""
    W1 = bf16(theta[:768]).reshape(64, 12).astype(np.float64)
    W2 = bf16(theta[832:4928]).reshape(64, 64).astype(np.float64)
    W3 = bf16(theta[4992:9088]).reshape(64, 64).astype(np.float64)
    b1, b2, b3 = theta[768:832], theta[4928:4992], theta[9088:9152]
    h = gelu(bf16(x).astype(np.float64) @ W1.T + b1)
    h = gelu(bf16(h).astype(np.float64) @ W2.T + b2)
    h = gelu(bf16(h).astype(np.float64) @ W3.T + b3)
    z = h @ theta[9152:9216].astype(np.float64) + theta[9216]
    return 1 / (1 + np.exp(-z))


def test_bf16_mfma_inference():
    rng = np.random.Generator(np.random.PCG64(6))
    theta = random_net(3)
    x = rng.uniform(-1, 1, (1000, 12)).astype(np.float32)  # not a multiple of 16: a partial last tile
    p = ra.mlp_infer(theta, x)
    emu = net_bf16(theta, x)
    _, _, z = oracle_lib.mlp_loss_grad(theta, x, np.zeros(len(x), dtype=np.float32))
    exact = 1 / (1 + np.exp(-z))
    print(f"max |p - bf16 network| {np.abs(p - emu).max():.2e}, max |p - FP64 network| {np.abs(p - exact).max():.2e}",
          flush=True)
    assert np.abs(p - emu).max() < 2e-3
    assert np.abs(p - exact).max() < 0.05
    assert np.corrcoef(p, exact)[0, 1] > 0.99


def chains(n_side=(24, 20, 12), phi=0.45):
    box, pos, rot, c, r = lattice.kinked_chain_lattice(n_side, phi)
    return ra.Shape("polysphere", centres=c, radii=r), box, pos, rot, c, r


def test_candidate_pair_labels_are_the_exact_decisions():
    shape, box, pos, rot, c, r = chains()
    pk = ra.GpuPacking(shape, box, pos, rot, seed=3)
    pk.set_environment(1.0, 3.0, 0.3, None, box_move=("delta_triclinic", 0.005))
    pk.run_cycles(20)
    x, y, meta = pk.mlp_pairs(2000)
    assert len(x) > 10000 and 0 < y.mean() < 0.5
    # the bead geometry of each pair from its features: candidate beads at R_tot * pos + Q c_b in the trial frame
    R_tot = shape.total_range
    rel = x[:, :3].astype(np.float64) * R_tot
    Q = x[:, 3:].astype(np.float64).reshape(-1, 3, 3)
    cb = np.einsum("nij,bj->nbi", Q, c) + rel[:, None, :]
    d2 = ((cb[:, None, :, :] - c[None, :, None, :]) ** 2).sum(-1)  # n x a x b
    gap = d2.min(axis=(1, 2)) - 1.0  # touching beads of diameter 1
    sure = np.abs(gap) > 1e-4  # FP32 features: each position carries an error of order 1e-6 x total range
    print(f"{len(y)} pairs, overlap fraction {y.mean():.4f}, {np.count_nonzero(~sure)} within 1e-4 of contact",
          flush=True)
    assert np.array_equal(y[sure] > 0.5, gap[sure] < 0)
    # natural ranks count up within each move; moves appear in order
    assert np.all(np.diff(meta[:, 0]) >= 0)
    first = np.r_[True, meta[1:, 0] != meta[:-1, 0]]
    assert np.all(meta[first, 1] == 0)


def test_learned_order_on_the_kinked_chain_workload():
    """Kinked chains at their specification: N = 50 000, packing fraction 0.45, NpT p* = 3."""
    shape, box, pos, rot, c, r = chains((50, 40, 25), 0.45)
    assert len(pos) == 50000
    pk = ra.GpuPacking(shape, box, pos, rot, seed=2024)
    pk.set_environment(1.0, 3.0, 0.2, None, box_move=("delta_triclinic", 0.005))
    pk.run_cycles(30, adapt_steps=True)
    # 1e6 pose/overlap pairs from the early trajectory: 5 snapshots 10 cycles apart
    tr = pk.mlp_train(snapshots=5, cycles_between=10, moves_per_snapshot=800, epochs=4, batch=4096,
                      learning_rate=2e-3)
    print(f"training: {tr['pairs']} pairs, overlap fraction {tr['overlap_fraction']:.4f}, {tr['steps']} Adam steps, "
          f"loss {tr['loss_initial']:.4f} -> {tr['loss_final']:.4f}", flush=True)
    assert tr["pairs"] > 5 * 10 ** 5
    assert tr["loss_final"] < 0.8 * tr["loss_initial"]
    pk.run_cycles(20)  # later states than the training ones
    st = pk.mlp_order_stats(4000)
    print(f"order statistics over {st['moves']} trial moves ({st['moves_rejected']} rejected, "
          f"{st['candidates_per_move']:.1f} close candidates per move): exact tests per move natural "
          f"{st['tests_natural']:.1f}, nearest-first {st['tests_nearest_first']:.1f}, learned "
          f"{st['tests_learned']:.1f}", flush=True)
    assert st["moves_rejected"] > 0
    assert st["tests_learned"] < st["tests_nearest_first"] < st["tests_natural"]
    # the sweep is unaffected: the network orders tests only; the packing stays overlap-free
    pk.run_cycles(10)
    b, p, rr = pk.download()
    n, _ = oracle_lib.oracle().snapshot_stats(2, 0, 0, b, p, rr, centres=c, radii=r)
    assert n == 0
