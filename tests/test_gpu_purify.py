"""Density build by SP2 purification (ygpu_set_density_method(YGPU_DENSITY_PURIFY)) on the GPU.

The reference forms D from the n_occ lowest eigenvectors of F' (build_density.c:37-69); the
purification path must give the same D = X P X^T, P the spectral projector, which numpy's eigh
fixes independently.  The fallbacks (no gap, nearly orthogonal occupied spaces) must hand the
build to the eigensolver, and a complete SCF must not depend on the method.
"""
import sys

import numpy as np
import pytest

from conftest import xyz_path

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ygpu():
    import ygpu as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


def s_shells(n):
    """n single-primitive s shells: a basis that only sets nbf = n for the density build."""
    return dict(nshells=n, am=np.zeros(n), nprim=np.ones(n), prim_off=np.arange(n),
                xyz=np.zeros((n, 3)), alpha=np.ones(n), coef=np.ones(n))


def spectrum(n, seed, gap=0.5):
    """Eigenvalues in [-12, 3] (as the Fock spectra of the test molecules), n_occ = n // 4, the
    virtual levels lifted by `gap`, and an orthogonal eigenbasis."""
    rng = np.random.RandomState(seed)
    nocc = max(1, n // 4)
    lam = np.sort(rng.uniform(-12.0, 3.0, n))
    lam[nocc:] += gap
    Q, _ = np.linalg.qr(rng.randn(n, n))
    return lam, Q, nocc


def sym(Q, lam):
    F = (Q * lam) @ Q.T
    return 0.5 * (F + F.T)


def run(ygpu, n, builds):
    """Density builds on one engine: builds = [(method, F, X, nocc), ...]; returns each
    (D, C, method used, SP2 steps)."""
    e = ygpu.Engine(0)
    try:
        e.set_basis(0, s_shells(n))
        out = []
        for method, F, X, nocc in builds:
            e.set_density_method(method)
            D, C, ev, idx = e.build_density(F, X, nocc)
            out.append((D, C) + e.density_last_info())
        return out
    finally:
        e.close()


@pytest.mark.parametrize("n", [40, 260, 300, 1100, 1560])
def test_purified_density_matches_the_projector(ygpu, n):
    log("case: SP2 density n=%d" % n)
    lam, Q, nocc = spectrum(n, 500 + n)
    rng = np.random.RandomState(n)
    X = np.eye(n) + 0.05 * rng.randn(n, n) / np.sqrt(n)
    F = sym(Q, lam)
    # the previous iteration: a nearby Fock matrix, whose occupied vectors start the purification
    dF = rng.randn(n, n) * 1e-3
    F1 = F + 0.5 * (dF + dF.T)
    (D1, C1, m1, k1), (D, C, m, k) = run(ygpu, n, [(ygpu.DENSITY_EIGEN, F1, X, nocc),
                                                    (ygpu.DENSITY_PURIFY, F, X, nocc)])
    assert (m1, k1) == (ygpu.DENSITY_EIGEN, 0)
    assert m == ygpu.DENSITY_PURIFY and k > 0
    Dref = X @ Q[:, :nocc] @ Q[:, :nocc].T @ X.T
    C0 = np.linalg.solve(X, C)
    errD = np.max(np.abs(D - Dref))
    errO = np.max(np.abs(C0.T @ C0 - np.eye(nocc)))
    log("  %d SP2 steps, max|dD| %.2e, orthonormality %.2e" % (k, errD, errO))
    assert errD < 1e-11
    assert errO < 1e-12
    assert np.max(np.abs(C @ C.T - D)) < 1e-12


def test_purification_falls_back_without_a_gap(ygpu):
    """A degenerate pair across the occupation boundary: SP2 leaves it half occupied and never
    reaches idempotency, so the build goes to the eigensolver."""
    n = 300
    lam, Q, nocc = spectrum(n, 7, gap=0.5)
    lam[nocc] = lam[nocc - 1]
    X = np.eye(n)
    F = sym(Q, lam)
    (_, _, m1, _), (D, C, m, k) = run(ygpu, n, [(ygpu.DENSITY_EIGEN, F, X, nocc),
                                                (ygpu.DENSITY_PURIFY, F, X, nocc)])
    log("  method %d after %d SP2 steps" % (m, k))
    assert m == ygpu.DENSITY_EIGEN and k > 0
    assert abs(np.trace(D) - nocc) < 1e-10
    assert np.max(np.abs(D @ D - D)) < 1e-11
    # every level below the pair is occupied, every level above it empty
    P_low = Q[:, :nocc - 1] @ Q[:, :nocc - 1].T
    assert np.max(np.abs(P_low @ D @ P_low - P_low)) < 1e-11


def test_purification_falls_back_on_a_swapped_level(ygpu):
    """The previous occupied space holds the new LUMO instead of the new HOMO (a level
    crossing): V_prev^T P V_prev is singular, and the eigensolver builds the density."""
    n = 300
    lam, Q, nocc = spectrum(n, 11)
    lam1 = lam.copy()
    lam1[nocc - 1], lam1[nocc] = lam[nocc] + 1.0, lam[nocc - 1]
    X = np.eye(n)
    (_, _, _, _), (D, C, m, k) = run(ygpu, n, [(ygpu.DENSITY_EIGEN, sym(Q, lam1), X, nocc),
                                               (ygpu.DENSITY_PURIFY, sym(Q, lam), X, nocc)])
    log("  method %d after %d SP2 steps" % (m, k))
    assert m == ygpu.DENSITY_EIGEN
    assert np.max(np.abs(D - Q[:, :nocc] @ Q[:, :nocc].T)) < 1e-11


def test_scf_does_not_depend_on_the_density_method(ygpu):
    """C10H22: a complete SCF with purified densities (the default) against the eigensolver in
    every iteration: same iterations, energy within 1e-10, and the same final orbital energies
    (the last density build of both is the eigensolver's)."""
    a = ygpu.solve(xyz_path("c10h22"), ene_tol=1e-11, purify=1)
    b = ygpu.solve(xyz_path("c10h22"), ene_tol=1e-11, purify=0)
    log("  purify: %d iterations E %.12f; eigensolver: %d iterations E %.12f"
        % (a["iters"], a["energy"], b["iters"], b["energy"]))
    assert a["converged"] and b["converged"]
    assert abs(a["iters"] - b["iters"]) <= 1
    assert abs(a["energy"] - b["energy"]) < 1e-10
    assert np.max(np.abs(a["eigval"] - b["eigval"])) < 1e-8
    assert np.max(np.abs(a["D"] - b["D"])) < 1e-7
