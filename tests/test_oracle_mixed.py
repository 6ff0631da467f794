"""Pins of the BF16-split emulation (oracle/mixed.py) used to check the GPU's
mixed-precision K build: rounding against hand-worked bf16 bit patterns,
the representation and product error bounds of the split."""
import numpy as np
import pytest

import mixed


def test_bf16_round_known_patterns():
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20,
                  -1.0 - 2 ** -8 - 2 ** -20, 3.140625, 0.0, 2.0 ** -130], dtype=np.float32)
    want = np.array([1.0, 1.0 + 2 ** -7,
                     1.0,                 # tie, 1.0 has an even last bit
                     1.0 + 2 ** -6,       # tie between 1+2^-7 (odd) and 1+2^-6 (even)
                     1.0 + 2 ** -7,       # above the tie
                     -1.0 - 2 ** -7,      # sign-magnitude rounding
                     3.140625,            # 8 significant bits: exact
                     0.0,
                     2.0 ** -130],        # subnormal with one bit set: exact
                    dtype=np.float32)
    got = mixed.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # every result has its low 16 bits clear (a bf16 value)
    r = np.random.RandomState(0).randn(1000).astype(np.float32)
    assert np.all((mixed.bf16_round(r).view(np.uint32) & 0xFFFF) == 0)


def test_split_represents_float32_to_2_pow_minus_17():
    x = np.random.RandomState(1).randn(100000) * np.exp(np.random.RandomState(2).randn(100000) * 5)
    x32 = x.astype(np.float32).astype(np.float64)
    hi, lo = mixed.split(x)
    err = np.abs(x32 - (hi + lo)) / np.abs(x32)
    assert err.max() <= 2.0 ** -17
    assert np.all(np.abs(lo) <= 2.0 ** -8 * np.abs(hi))


def test_split_matmul_error_bound():
    rng = np.random.RandomState(3)
    a, b = rng.randn(40, 300), rng.randn(300, 30)
    exact = a.astype(np.float32).astype(np.float64) @ b.astype(np.float32).astype(np.float64)
    got = mixed.split_matmul(a, b)
    # per product |a b - (hh + hl + lh)| <= ~3 * 2^-17 |a||b|
    bound = 3 * 2.0 ** -17 * (np.abs(a) @ np.abs(b))
    assert np.all(np.abs(got - exact) <= bound)
    # and the dropped lo*lo term is what makes it inexact (the four-term sum is exact)
    ah, al = mixed.split(a)
    bh, bl = mixed.split(b)
    four = ah @ bh + ah @ bl + al @ bh + al @ bl
    assert np.max(np.abs(four - (ah + al) @ (bh + bl))) < 1e-12 * np.max(np.abs(exact))


@pytest.mark.parametrize("nbf,nocc,naux", [(7, 3, 11), (12, 5, 20)])
def test_k_build_matches_split_products(nbf, nocc, naux):
    """k_build against a direct statement of X_i and K on a random screened mask."""
    rng = np.random.RandomState(nbf)
    keep = rng.rand(nbf, nbf) < 0.7
    keep = keep | keep.T
    np.fill_diagonal(keep, True)
    mask = -np.ones((nbf, nbf), dtype=np.int64)
    mask[keep] = np.arange(keep.sum())
    B = rng.randn(keep.sum(), naux)
    C = rng.randn(nbf, nocc)
    K, X32 = mixed.k_build(B, mask, C)
    # the exact K of the screened tensor: K_ij = sum_{s,p} X_isp X_jsp, X_i = C[js]^T B[rows]
    X = np.zeros((nbf, nocc, naux))
    for i in range(nbf):
        js = np.nonzero(keep[i])[0]
        X[i] = C[js].T @ B[mask[i, js]]
    Kx = X.reshape(nbf, -1) @ X.reshape(nbf, -1).T
    assert np.max(np.abs(X32 - X)) <= 1e-4 * np.max(np.abs(X))
    assert np.max(np.abs(K - Kx)) <= 1e-4 * np.max(np.abs(Kx))
    assert np.allclose(K, K.T, rtol=0, atol=1e-12 * np.abs(K).max())


def test_three_part_split_known_patterns():
    """Hand-worked parts: x = 1 + 2^-9 + 2^-18 + 2^-23 splits into 1, 2^-9 and 2^-18 + 2^-23 (each
    part the bf16 rounding of the remainder); x = 1 + 2^-8 (a tie) gives hi = 1, mid = 2^-8."""
    x = np.array([1.0 + 2 ** -9 + 2 ** -18 + 2 ** -23, 1.0 + 2 ** -8, -3.0 - 2 ** -12], dtype=np.float32)
    p0, p1, p2 = mixed.split(x, 3)
    assert np.array_equal(p0, [1.0, 1.0, -3.0])
    assert np.array_equal(p1, [2.0 ** -9, 2.0 ** -8, -2.0 ** -12])
    assert np.array_equal(p2, [2.0 ** -18 + 2.0 ** -23, 0.0, 0.0])
    # two parts are the first two of three
    h, lo = mixed.split(x, 2)
    assert np.array_equal(h, p0) and np.array_equal(lo, p1)


def test_three_parts_hold_float32_exactly():
    x = np.random.RandomState(4).randn(100000) * np.exp(np.random.RandomState(5).randn(100000) * 5)
    x32 = x.astype(np.float32).astype(np.float64)
    p0, p1, p2 = mixed.split(x, 3)
    assert np.array_equal(p0 + p1 + p2, x32)
    assert np.all(np.abs(p1) <= 2.0 ** -8 * np.abs(p0)) and np.all(np.abs(p2) <= 2.0 ** -8 * np.abs(p1))


def test_six_term_product_error_bound():
    """BF16x6: |ab - sum_{a+b<3} p_a q_b| <= ~3 2^-24 |a||b| per product (the dropped
    mid*lo, lo*mid, lo*lo terms), 2^7 times tighter than BF16x3's bound; the nine-term sum is
    the exact float32 product."""
    rng = np.random.RandomState(6)
    a, b = rng.randn(40, 300), rng.randn(300, 30)
    exact = a.astype(np.float32).astype(np.float64) @ b.astype(np.float32).astype(np.float64)
    got = mixed.split_matmul(a, b, 3)
    bound = 3 * 2.0 ** -23 * (np.abs(a) @ np.abs(b))
    assert np.all(np.abs(got - exact) <= bound)
    got3 = mixed.split_matmul(a, b, 2)
    assert np.max(np.abs(got - exact)) < np.max(np.abs(got3 - exact)) / 16
    pa, pb = mixed.split(a, 3), mixed.split(b, 3)
    nine = sum(pa[i] @ pb[j] for i in range(3) for j in range(3))
    assert np.max(np.abs(nine - exact)) < 1e-12 * np.max(np.abs(exact))
