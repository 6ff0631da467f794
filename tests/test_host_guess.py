"""The host side of the initial guess (TinySCF.c:282-364): D -> C_occ by
pivoted Cholesky, with the reference's quirk A1 (pivot ignored, diagonal of
the first n_occ rows dropped, TinySCF.c:295-319) and without it.  The host
engine's own pivoted Cholesky (host/TinySCF.c) against the oracle, which
calls LAPACK's dpstrf as the reference does, and against what the
factorisation must satisfy on a density of known rank."""
import ctypes

import numpy as np
import pytest

from conftest import system

dp = ctypes.POINTER(ctypes.c_double)


def d2cocc(D, n_occ, quirk_a1):
    import ygpu
    L = ygpu.host_lib()
    L.tinyscf_d2cocc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
    L.tinyscf_d2cocc.restype = ctypes.c_int
    D = np.ascontiguousarray(D, dtype=np.float64)
    n = D.shape[0]
    C = np.zeros((n, n_occ))
    assert L.tinyscf_d2cocc(n, n_occ, int(quirk_a1), D.ctypes.data_as(dp), C.ctypes.data_as(dp)) == 0
    return C


@pytest.mark.parametrize("name", ["water", "c10h22"])
@pytest.mark.parametrize("quirk_a1", [True, False])
def test_d2cocc_matches_the_oracle(name, quirk_a1):
    """The SAD density of each workload: the same C_occ as the oracle's dpstrf path, to rounding
    (the pivots coincide, equal SAD diagonals included)."""
    import oracle
    s = system(name)
    o = oracle.OracleSCF(s, quirk_a1=quirk_a1)
    o.setup()
    n, no = s.nbf, s.n_occ
    D = o.mat("D_mat").copy()
    ref = o.d("Cocc_mat", (n, no)).copy()
    C = d2cocc(D, no, quirk_a1)
    assert np.max(np.abs(C - ref)) < 1e-15 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n,n_occ,rank", [(12, 3, 3), (40, 7, 7), (40, 7, 5), (64, 10, 10)])
def test_d2cocc_factor_of_a_known_rank_density(n, n_occ, rank):
    """D = G G^T of rank r <= n_occ: without the quirk the factor reproduces D exactly and its
    columns beyond the rank are zero; with the quirk the output is the unpivoted copy of the
    pivoted factor's lower triangle with the first n_occ rows' diagonals dropped."""
    rng = np.random.RandomState(n + 31 * rank)
    G = rng.randn(n, rank)
    D = G @ G.T
    C = d2cocc(D, n_occ, False)
    assert np.max(np.abs(C @ C.T - D)) < 1e-12 * np.abs(D).max()
    assert np.all(C[:, rank:] == 0.0)
    # the quirk: the pivoted factor L (rows in pivot order, C[piv[i]] = L[i] without it) is copied
    # unpivoted, rows >= n_occ whole, rows < n_occ only below the diagonal
    Cq = d2cocc(D, n_occ, True)
    assert np.all(np.triu(Cq[:n_occ]) == 0.0)
    rows = {C[r].tobytes() for r in range(n)}
    assert all(Cq[i].tobytes() in rows for i in range(n_occ, n))
    for i in range(n_occ):
        assert any(np.array_equal(Cq[i, :i], C[r, :i]) for r in range(n))
