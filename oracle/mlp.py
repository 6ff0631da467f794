"""CPU restatement of the learned-guess network (test infrastructure only).

The GPU library's network (csrc/ygpu_mlp.hip, ygpu_mlp_forward / the trainer)
is n_in -> nh -> nh -> nh -> n_out with SiLU after each hidden layer, packed
parameters W1 [n_in][nh], b1, W2 [nh][nh], b2, W3 [nh][nh], b3, W4 [nh][n_out],
b4, and the batch loss L = sum M (Y - T)^2 / (B n_out).  This module states the
same network and its gradient in float64.

Pinned by: tests/test_oracle_mlp.py -- central finite differences of the loss
against the analytic gradient for every parameter block, and SiLU' against
its finite difference.
"""
import numpy as np


def layout(n_in, nh, n_out):
    sizes = [(n_in, nh), (nh,), (nh, nh), (nh,), (nh, nh), (nh,), (nh, n_out), (n_out,)]
    out, off = [], 0
    for s in sizes:
        n = int(np.prod(s))
        out.append((off, s))
        off += n
    return out, off


def unpack(p, n_in, nh, n_out):
    lay, total = layout(n_in, nh, n_out)
    assert p.size == total
    return [p[o:o + int(np.prod(s))].reshape(s) for o, s in lay]


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, nh, n_out):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(np.asarray(p, dtype=np.float64), X.shape[1], nh, n_out)
    Z1 = X @ W1 + b1
    H1 = silu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = silu(Z2)
    Z3 = H2 @ W3 + b3
    H3 = silu(Z3)
    Y = H3 @ W4 + b4
    return Y, (X, Z1, H1, Z2, H2, Z3, H3)


def loss_grad(p, X, T, M, nh, n_out):
    """(loss, gradient packed like p) of L = sum M (net(X) - T)^2 / (B n_out)."""
    Y, (X, Z1, H1, Z2, H2, Z3, H3) = forward(p, X, nh, n_out)
    B = X.shape[0]
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(np.asarray(p, dtype=np.float64), X.shape[1], nh, n_out)
    R = (Y - T) * M
    L = np.sum(R * R) / (B * n_out)
    dY = 2.0 * R / (B * n_out)
    gW4, gb4 = H3.T @ dY, dY.sum(0)
    dZ3 = (dY @ W4.T) * silu_grad(Z3)
    gW3, gb3 = H2.T @ dZ3, dZ3.sum(0)
    dZ2 = (dZ3 @ W3.T) * silu_grad(Z2)
    gW2, gb2 = H1.T @ dZ2, dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * silu_grad(Z1)
    gW1, gb1 = X.T @ dZ1, dZ1.sum(0)
    g = np.concatenate([a.reshape(-1) for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return L, g
