"""Pins the CPU reference of the estimator network by finite differences."""
import numpy as np

from oracle import mlp


def test_analytic_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    n_params = sum(i * o + o for i, o in mlp.DIMS)
    params = rng.normal(0, 0.1, n_params)
    x = rng.normal(0, 1, (8, 256))
    t = rng.normal(0, 1, (8, 3))
    loss, g = mlp.loss_and_grads(params, x, t, bf16_ops=False)
    idx = rng.choice(n_params, 60, replace=False)
    # include output-layer weights and biases explicitly
    idx = np.concatenate([idx, [n_params - 1, n_params - 4, n_params - 200]])
    eps = 1e-4  # central differences: O(eps^2) truncation, roundoff well below
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        l1, _ = mlp.loss_and_grads(p1, x, t, bf16_ops=False)
        l2, _ = mlp.loss_and_grads(p2, x, t, bf16_ops=False)
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5], np.float32)
    got = mlp.bf16(vals)
    # bf16 keeps 7 fraction bits: the spacing at 1.0 is 2^-7, so 1 + 2^-8 is a tie
    # that rounds to even (1.0) and 1 + 3*2^-8 rounds up to 1 + 2^-6
    assert mlp.bf16(np.array([1.0 + 2 ** -8], np.float32))[0] == 1.0
    assert mlp.bf16(np.array([1.0 + 3 * 2 ** -8], np.float32))[0] == 1.0 + 2 ** -6
    assert mlp.bf16(np.array([1.0 + 2 ** -7], np.float32))[0] == 1.0 + 2 ** -7
    assert got[-1] == -2.5
