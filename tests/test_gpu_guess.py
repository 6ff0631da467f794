"""The learned initial guess on the GPU: the network's HIP forward and backward
against the float64 oracle (oracle/mlp.py), training on the device, and the
trained network's effect on the SCF (iterations from the learned guess against
iterations from SAD, same converged energy)."""
import os
import sys
import tempfile

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(PKG, "data", "guess_mlp_alkane.bin")   # trained by guess_train.py --family alkane


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ygpu():
    import ygpu as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


def net_case(seed, n, n_in=165, nh=128, n_out=29):
    import mlp
    rng = np.random.RandomState(seed)
    _, total = mlp.layout(n_in, nh, n_out)
    p = (rng.randn(total) * 0.1).astype(np.float32)
    X = rng.randn(n, n_in).astype(np.float32)
    T = rng.randn(n, n_out).astype(np.float32)
    M = (rng.rand(n, n_out) < 0.6).astype(np.float32)
    return p, X, T, M


def test_mlp_forward_matches_oracle(ygpu):
    import mlp
    log("case: MLP forward")
    p, X, _, _ = net_case(1, 1000)
    e = ygpu.Engine(0)
    try:
        Y = ygpu.mlp_forward(e, 165, 128, 29, p, X)
    finally:
        e.close()
    ref, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64), 128, 29)
    err = np.abs(Y - ref).max()
    log("  max|dY| %.3e (max|Y| %.3e)" % (err, np.abs(ref).max()))
    assert err < 1e-5 * max(1.0, np.abs(ref).max())


def test_mlp_gradient_matches_oracle(ygpu):
    import mlp
    log("case: MLP gradient")
    p, X, T, M = net_case(2, 700)
    idx = np.random.RandomState(3).choice(700, 300, replace=False).astype(np.int32)
    e = ygpu.Engine(0)
    try:
        tr = ygpu.MLPTrainer(e, 128, X, T, M, p)
        try:
            loss, g = tr.grad(idx)
        finally:
            tr.close()
    finally:
        e.close()
    L, gref = mlp.loss_grad(p.astype(np.float64), X[idx].astype(np.float64), T[idx].astype(np.float64),
                            M[idx].astype(np.float64), 128, 29)
    err = np.abs(g - gref).max()
    log("  loss gpu %.8f oracle %.8f; max|dg| %.3e (max|g| %.3e)" % (loss, L, err, np.abs(gref).max()))
    assert abs(loss - L) < 1e-5 * L
    assert err < 1e-4 * np.abs(gref).max()


def test_mlp_training_fits_a_known_network(ygpu):
    """Targets from a fixed random network: training from another start drives the loss down."""
    import mlp
    log("case: MLP training")
    n_in, nh, n_out = 20, 128, 6
    rng = np.random.RandomState(4)
    _, total = mlp.layout(n_in, nh, n_out)
    teacher = rng.randn(total) * 0.2
    X = rng.randn(4000, n_in).astype(np.float32)
    T = mlp.forward(teacher, X.astype(np.float64), nh, n_out)[0].astype(np.float32)
    M = np.ones_like(T)
    p0 = (rng.randn(total) * 0.05).astype(np.float32)
    L0, _ = mlp.loss_grad(p0.astype(np.float64), X.astype(np.float64), T.astype(np.float64), M, nh, n_out)
    e = ygpu.Engine(0)
    try:
        tr = ygpu.MLPTrainer(e, nh, X, T, M, p0)
        try:
            losses = [tr.epoch(256, 2e-3, 100 + k) for k in range(30)]
            p = tr.params()
        finally:
            tr.close()
    finally:
        e.close()
    L1, _ = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), M, nh, n_out)
    log("  loss %.5f -> %.5f (epoch losses %.5f .. %.5f)" % (L0, L1, losses[0], losses[-1]))
    assert L1 < 0.1 * L0
    assert losses[-1] < losses[0]


def test_learned_guess_saves_iterations(ygpu):
    """Held-out alkane (a seed outside the training set): the trained network's guess converges
    to the SAD run's energy in fewer iterations."""
    import guess_train
    log("case: learned guess on a held-out alkane")
    assert os.path.exists(WEIGHTS), "trained weights missing: run guess_train.py"
    tmp = tempfile.mkdtemp(prefix="ysml_test_")
    path = guess_train.write_alkane(tmp, 14, 987654321)
    a = guess_train.solve(path, None, ene_tol=1e-9)
    b = guess_train.solve(path, WEIGHTS, ene_tol=1e-9)
    log("  SAD %d iterations E %.10f; learned %d iterations E %.10f (applied %d)"
        % (a["iters"], a["energy"], b["iters"], b["energy"], b["guess_ml"]))
    assert a["converged"] and b["converged"] and b["guess_ml"] == 1
    assert abs(a["energy"] - b["energy"]) < 1e-8
    assert b["iters"] < a["iters"]
