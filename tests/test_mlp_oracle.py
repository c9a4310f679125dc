"""The oracle's network of the learned ordering (orc_mlp_loss_grad): its analytic gradient pinned by central finite
differences of its own loss, and the GELU form against the closed form values x/2 (1 + tanh(...)) at known points."""
import math

import numpy as np

import oracle_lib


def random_net(seed, scale=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    theta = np.zeros(9217, dtype=np.float32)
    offs = [(0, 768, 12), (768, 832, 1), (832, 4928, 64), (4928, 4992, 1), (4992, 9088, 64), (9088, 9152, 1),
            (9152, 9216, 64), (9216, 9217, 1)]
    for a, b, fan in offs:
        theta[a:b] = rng.normal(0, scale / math.sqrt(fan), b - a)
    return theta


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.Generator(np.random.PCG64(3))
    theta = random_net(1)
    x = rng.uniform(-1, 1, (40, 12)).astype(np.float32)
    y = (rng.random(40) < 0.3).astype(np.float32)
    loss, grad, _ = oracle_lib.mlp_loss_grad(theta, x, y)
    assert loss > 0
    # parameters of every block (weights and biases of each layer, the head)
    idx = [0, 5, 700, 770, 800, 900, 3000, 4930, 5000, 7000, 9090, 9152, 9200, 9216]
    for i in idx:
        h = 1e-2 * max(1e-3, abs(float(theta[i])))
        tp, tm = theta.astype(np.float64).copy(), theta.astype(np.float64).copy()
        # the oracle takes float parameters: step by a float-representable amount
        tp[i] = np.float32(theta[i] + h)
        tm[i] = np.float32(theta[i] - h)
        lp, _, _ = oracle_lib.mlp_loss_grad(tp.astype(np.float32), x, y)
        lm, _, _ = oracle_lib.mlp_loss_grad(tm.astype(np.float32), x, y)
        fd = (lp - lm) / (tp[i] - tm[i])
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(grad[i]), (i, fd, grad[i])


def test_oracle_logit_of_a_zero_network_is_its_output_bias():
    theta = np.zeros(9217, dtype=np.float32)
    theta[9216] = 0.75
    x = np.ones((3, 12), dtype=np.float32)
    y = np.array([0, 1, 1], dtype=np.float32)
    loss, grad, z = oracle_lib.mlp_loss_grad(theta, x, y)
    assert np.allclose(z, 0.75)
    p = 1 / (1 + math.exp(-0.75))
    assert math.isclose(loss, (-(math.log(1 - p)) - 2 * math.log(p)) / 3, rel_tol=1e-12)
    assert math.isclose(grad[9216], (3 * p - 2) / 3, rel_tol=1e-12)
