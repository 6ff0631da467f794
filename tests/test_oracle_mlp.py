"""Pins of the network oracle (oracle/mlp.py): analytic gradient against
central finite differences of the loss, SiLU' against its difference."""
import numpy as np

import mlp


def test_silu_grad_matches_difference():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp.silu_grad(z))) < 1e-8


def test_loss_gradient_against_finite_differences():
    rng = np.random.RandomState(0)
    n_in, nh, n_out, B = 7, 5, 4, 9
    _, total = mlp.layout(n_in, nh, n_out)
    p = rng.randn(total) * 0.5
    X = rng.randn(B, n_in)
    T = rng.randn(B, n_out)
    M = (rng.rand(B, n_out) < 0.7).astype(float)
    L, g = mlp.loss_grad(p, X, T, M, nh, n_out)
    h = 1e-6
    fd = np.zeros_like(p)
    for k in range(total):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (mlp.loss_grad(p + e, X, T, M, nh, n_out)[0] - mlp.loss_grad(p - e, X, T, M, nh, n_out)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.abs(g).max())
    # masked outputs carry no gradient: a target change under M = 0 changes nothing
    T2 = T + (1 - M) * 5.0
    assert np.allclose(mlp.loss_grad(p, X, T2, M, nh, n_out)[1], g)
    assert L > 0
