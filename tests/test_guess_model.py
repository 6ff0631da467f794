"""The learned guess's pair model on the CPU (host/guess_ml.c through
libtinyscf.so): pair list, invariant features, coefficient projection and
rebuild, and the rotation covariance of the rebuilt density correction."""
import numpy as np
import pytest

from conftest import system


def model_of(sys_, xyz=None):
    import ygpu
    b, m = sys_.basis, sys_.mol
    na = m["natoms"]
    ss = np.searchsorted(b["atom"], np.arange(na + 1)).astype(np.int32)
    coords = m["coords"] if xyz is None else xyz
    return ygpu.PairModel(m["Z"], coords, ss, b["am"], b["bf_start"][:b["nshells"]], sys_.nbf)


def rotation(seed):
    q = np.random.RandomState(seed).randn(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def ao_rotation(sys_, R):
    """AO-basis rotation for s (1) and p (R) shells; d rows stay zero (never written)."""
    b = sys_.basis
    U = np.zeros((sys_.nbf, sys_.nbf))
    for sh in range(b["nshells"]):
        f = b["bf_start"][sh]
        if b["am"][sh] == 0:
            U[f, f] = 1.0
        elif b["am"][sh] == 1:
            U[f:f + 3, f:f + 3] = R
    return U


@pytest.fixture(scope="module")
def c10():
    return system("c10h22")


def test_pairs_cover_the_cutoff(c10):
    pm = model_of(c10)
    xyz = c10.mol["coords"] * 0.52917720859
    Z = c10.mol["Z"]
    na = len(Z)
    want = set()
    for a in range(na):
        for b in range(a, na):
            if a == b or np.linalg.norm(xyz[a] - xyz[b]) < 6.0:
                want.add((min(a, b), max(a, b)))
    got = set((min(a, b), max(a, b)) for a, b in pm.pairs)
    assert got == want and len(got) == len(pm.pairs)
    # canonical order: carbon first in C-H pairs
    for a, b in pm.pairs:
        assert not (Z[a] == 1 and Z[b] == 6)
    assert pm.supported()


def test_features_invariant_under_rigid_motion(c10):
    pm = model_of(c10)
    R = rotation(5)
    moved = c10.mol["coords"] @ R.T + np.array([1.3, -2.0, 0.7])
    pm2 = model_of(c10, moved)
    assert np.array_equal(pm.pairs, pm2.pairs)
    X, X2 = pm.features(), pm2.features()
    assert np.max(np.abs(X - X2)) < 1e-5
    # one-hot pair type and self pairs
    t = pm.types()
    for (a, b), ty in zip(pm.pairs, t):
        za, zb = c10.mol["Z"][a], c10.mol["Z"][b]
        want = (3 if za == 6 else 4) if a == b else (0 if za == zb == 6 else 1 if za == 6 else 2)
        assert ty == want


def test_projection_of_the_rebuild_is_identity(c10):
    pm = model_of(c10)
    rng = np.random.RandomState(1)
    _, M = pm.targets(np.zeros((c10.nbf, c10.nbf)))
    Y = (rng.randn(*M.shape) * M).astype(np.float32)
    dD = pm.assemble(Y)
    assert np.allclose(dD, dD.T, atol=0)
    Y2, M2 = pm.targets(dD)
    assert np.array_equal(M, M2)
    # self s-s and p-p coefficients (shell x shell) are symmetrised by the rebuild
    t = pm.types()
    for p, ty in enumerate(t):
        if ty in (3, 4):
            ns, npp = (3, 2) if ty == 3 else (2, 1)
            S = Y[p, :ns * ns].reshape(ns, ns)
            Y[p, :ns * ns] = (0.5 * (S + S.T)).reshape(-1)
            P = Y[p, ns * ns:ns * ns + npp * npp].reshape(npp, npp)
            Y[p, ns * ns:ns * ns + npp * npp] = (0.5 * (P + P.T)).reshape(-1)
    assert np.max(np.abs(Y2 - Y)) < 1e-6
    # mask sizes per pair type: C-C 29, C-H 17, H-H 10, C self 13, H self 5
    sizes = {0: 29, 1: 17, 2: 10, 3: 13, 4: 5}
    for p, ty in enumerate(t):
        assert M[p].sum() == sizes[ty]


def test_rebuild_is_rotation_covariant(c10):
    pm = model_of(c10)
    R = rotation(11)
    pm2 = model_of(c10, c10.mol["coords"] @ R.T)
    rng = np.random.RandomState(2)
    _, M = pm.targets(np.zeros((c10.nbf, c10.nbf)))
    Y = (rng.randn(*M.shape) * M).astype(np.float32)
    U = ao_rotation(c10, R)
    dD, dD2 = pm.assemble(Y), pm2.assemble(Y)
    assert np.max(np.abs(dD2 - U @ dD @ U.T)) < 1e-5 * np.abs(dD).max()
    # and the projection of a rotated matrix gives the same coefficients
    Y1, _ = pm.targets(dD)
    Y2, _ = pm2.targets(U @ dD @ U.T)
    assert np.max(np.abs(Y1 - Y2)) < 1e-5 * np.abs(Y1).max()


def test_unsupported_element_is_reported():
    s = system("water")
    pm = model_of(s)
    assert not pm.supported()
