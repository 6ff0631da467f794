"""Spherical-harmonic basis option: the per-shell transforms T of ygpu_sph_coefficients
(C-ABI, no device needed).  No reference counterpart (the reference's Simint path is
Cartesian), so T is pinned by what fixes it mathematically:
  - every column is a harmonic polynomial (zero Laplacian) and the 2l+1 columns are
    independent: they span exactly the solid harmonics of degree l, the kernel of the
    Laplacian on degree-l polynomials, whose dimension is 2l+1;
  - the columns are orthonormal in the overlap metric of the x^l-normalised Cartesian
    components (angular integrals (a+a'-1)!!(b+b'-1)!!(c+c'-1)!!/(2l-1)!!), computed here from
    the closed form, not from the library;
  - textbook members: d0 ~ 2z^2 - x^2 - y^2, d(+2) ~ x^2 - y^2, f0 ~ 2z^3 - 3x^2 z - 3y^2 z,
    f(+3) ~ x^3 - 3xy^2 (m = -l..l column order);
  - s and p: the identity."""
import ctypes
import os

import numpy as np
import pytest

from conftest import PKG


def lib():
    path = os.path.join(PKG, "libyatscf_gpu.so")
    if not os.path.exists(path):
        pytest.skip("libyatscf_gpu.so not built")
    return ctypes.CDLL(path)


def comps(l):
    return [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]


def T_of(l):
    L = lib()
    nc, ns = (l + 1) * (l + 2) // 2, 2 * l + 1
    T = np.zeros(nc * ns)
    assert L.ygpu_sph_coefficients(l, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return T.reshape(nc, ns)


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def metric(l):
    c = comps(l)
    M = np.zeros((len(c), len(c)))
    for i, (a, b, e) in enumerate(c):
        for j, (a2, b2, e2) in enumerate(c):
            if (a + a2) % 2 or (b + b2) % 2 or (e + e2) % 2:
                continue
            M[i, j] = dfact(a + a2 - 1) * dfact(b + b2 - 1) * dfact(e + e2 - 1) / dfact(2 * l - 1)
    return M


def laplacian(l, col):
    """coefficients of the Laplacian of sum_c col[c] x^a y^b z^c over the degree l-2 monomials"""
    out = {}
    for (a, b, e), v in zip(comps(l), col):
        for k, p in ((0, a), (1, b), (2, e)):
            if p >= 2:
                m = [a, b, e]
                m[k] -= 2
                out[tuple(m)] = out.get(tuple(m), 0.0) + v * p * (p - 1)
    return out


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4])
def test_columns_are_the_solid_harmonics(l):
    T = T_of(l)
    assert T.shape == ((l + 1) * (l + 2) // 2, 2 * l + 1)
    assert np.linalg.matrix_rank(T) == 2 * l + 1
    for s in range(2 * l + 1):
        lap = laplacian(l, T[:, s])
        assert all(abs(v) < 1e-12 for v in lap.values()), (l, s, lap)


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4])
def test_columns_orthonormal_in_the_component_metric(l):
    T = T_of(l)
    G = T.T @ metric(l) @ T
    assert np.abs(G - np.eye(2 * l + 1)).max() < 1e-13


def test_s_and_p_are_the_identity():
    assert np.array_equal(T_of(0), np.eye(1))
    assert np.array_equal(T_of(1), np.eye(3))


def test_textbook_members():
    d = T_of(2)                                   # xx xy xz yy yz zz; m = -2..2
    c = {k: i for i, k in enumerate(comps(2))}
    d0 = d[:, 2]
    assert abs(d0[c[(0, 0, 2)]] / d0[c[(2, 0, 0)]] + 2.0) < 1e-14
    assert abs(d0[c[(0, 2, 0)]] - d0[c[(2, 0, 0)]]) < 1e-14
    d2 = d[:, 4]
    assert abs(d2[c[(2, 0, 0)]] + d2[c[(0, 2, 0)]]) < 1e-14 and abs(d2[c[(0, 0, 2)]]) < 1e-14
    f = T_of(3)
    c = {k: i for i, k in enumerate(comps(3))}
    f0 = f[:, 3]
    assert abs(f0[c[(2, 0, 1)]] / f0[c[(0, 0, 3)]] + 1.5) < 1e-14
    assert abs(f0[c[(0, 2, 1)]] / f0[c[(0, 0, 3)]] + 1.5) < 1e-14
    f3 = f[:, 6]
    assert abs(f3[c[(1, 2, 0)]] / f3[c[(3, 0, 0)]] + 3.0) < 1e-14
    nz = [k for k, i in c.items() if abs(f3[i]) > 1e-14]
    assert sorted(nz) == [(1, 2, 0), (3, 0, 0)]


def test_bad_degree_is_an_error():
    L = lib()
    T = np.zeros(200)
    assert L.ygpu_sph_coefficients(5, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) != 0
    assert L.ygpu_sph_coefficients(-1, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) != 0
