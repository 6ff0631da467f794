"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on the
same inputs.  Bars: FP64 Fock matrix within 1e-10 max-abs, converged total
energy within 1e-8 Eh; integral and tensor stages within the tolerances
written in each test."""
import contextlib
import sys

import numpy as np
import pytest

from conftest import system, xyz_path

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ygpu():
    import ygpu as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


class OracleState:
    def __init__(self, name):
        import oracle
        self.name = name
        self.sys = system(name)
        self.o = oracle.OracleSCF(self.sys, niters=60)
        self.o.setup()
        self.o.build_df_tensor()
        n = self.sys.nbf
        self.D0 = self.o.mat("D_mat").copy()
        self.C0 = self.o.d("Cocc_mat", (n, self.sys.n_occ)).copy()
        self.o.build_fock()
        self.F0 = self.o.mat("F_mat").copy()
        self.J0 = self.o.mat("J_mat").copy()
        self.K0 = self.o.mat("K_mat").copy()


_states = {}


def oracle_state(name):
    if name not in _states:
        _states[name] = OracleState(name)
    return _states[name]


@contextlib.contextmanager
def make_engine(ygpu, st=None):
    """An engine that is always destroyed when the block ends, also on failure."""
    e = ygpu.Engine(0)
    try:
        if st is not None:
            e.set_system(st.sys)
        yield e
    finally:
        e.close()


def set_oracle_sparsity(e, st):
    o = st.o
    e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                   o.d("df_sp_scrval"))


def test_dgemm_all_transposes(ygpu):
    log("case: dgemm transposes")
    rng = np.random.RandomState(7)
    with make_engine(ygpu) as e:
        # fewer than 64 output tiles with k >= 64 take the split-K path: (130, 70, 200), (257, 129, 67)
        # (two chunks, the second short), (260, 260, 260) (the SCF's n x n products), (300, 41, 1000)
        # (a C_occ product with deep k); (520, 520, 100) has 81 tiles and stays in one pass
        for (m, n, k) in [(1, 1, 1), (17, 33, 5), (64, 64, 16), (130, 70, 200), (257, 129, 67), (260, 260, 260),
                          (300, 41, 1000), (520, 520, 100)]:
            for ta in (0, 1):
                for tb in (0, 1):
                    A = rng.randn(*((k, m) if ta else (m, k)))
                    B = rng.randn(*((n, k) if tb else (k, n)))
                    C0 = rng.randn(m, n)
                    ref = 1.5 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.5 * C0
                    got = e.dgemm(ta, tb, A, B, alpha=1.5, beta=-0.5, C=C0)
                    assert np.max(np.abs(got - ref)) < 1e-12 * max(1, k), (m, n, k, ta, tb)
                    ref0 = 2.0 * ((A.T if ta else A) @ (B.T if tb else B))
                    got0 = e.dgemm(ta, tb, A, B, alpha=2.0, beta=0.0)
                    assert np.max(np.abs(got0 - ref0)) < 1e-12 * max(1, k), (m, n, k, ta, tb, "beta 0")
                    # the split partials are added in a fixed order: a repeat is bit-identical
                    assert np.array_equal(e.dgemm(ta, tb, A, B, alpha=2.0, beta=0.0), got0)


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_one_electron_and_schwarz(ygpu, name):
    log("case: S/H/X/Schwarz", name)
    st = oracle_state(name)
    with make_engine(ygpu, st) as e:
        S, H = e.ovlp_hcore()
        assert np.max(np.abs(S - st.o.mat("S_mat"))) < 1e-12
        assert np.max(np.abs(H - st.o.mat("Hcore_mat"))) < 1e-10
        X, mine = e.sym_inv_sqrt(S)
        assert mine > 0
        # X = S^-1/2 is unique: X S X = I
        assert np.max(np.abs(X @ S @ X - np.eye(len(S)))) < 1e-9
        assert np.max(np.abs(X - st.o.mat("X_mat"))) < 1e-8 * max(1.0, np.abs(X).max())
        sp, bf, df = e.schwarz(quirk_a5=True)
        ref_bf = st.o.d("bf_pair_scrval")
        assert np.allclose(bf, ref_bf, rtol=1e-10, atol=1e-20)
        assert np.allclose(sp, st.o.d("sp_scrval"), rtol=1e-10, atol=1e-20)
        assert np.allclose(df, st.o.d("df_sp_scrval"), rtol=1e-12, atol=0)


def metric(st):
    """J_PQ from independent oracle integral calls (the oracle overwrites its
    own copy with J^{-1/2}, as the reference does)."""
    import oracle
    if not hasattr(st, "J"):
        b = st.sys.auxbasis
        shells = []
        for i in range(b["nshells"]):
            k0, m = b["prim_off"][i], b["nprim"][i]
            shells.append(oracle.Shell(int(b["am"][i]), b["xyz"][i], b["alpha"][k0:k0 + m], b["coef"][k0:k0 + m]))
        n, bs = st.sys.naux, b["bf_start"]
        J = np.zeros((n, n))
        for P, A in enumerate(shells):
            for Q, B in enumerate(shells):
                J[bs[P]:bs[P + 1], bs[Q]:bs[Q + 1]] = oracle.eri2(A, B)
        st.J = J
    return st.J


def oracle_spread(st):
    """max|dF| between the oracle's F (MKL eigen-based J^{-1/2}) and the same
    oracle with an equally valid FP64 factor of the metric (Cholesky): the
    intrinsic FP64 precision of the DF step for this molecule."""
    if not hasattr(st, "spread"):
        o, n, nocc, naux = st.o, st.sys.nbf, st.sys.n_occ, st.sys.naux
        nnz = o.geti("num_bf_pair_scr")
        pqA = o.d("pqA").reshape(nnz, naux)
        L = np.linalg.cholesky(metric(st))
        Bv = o.d("df_tensor").reshape(nnz, naux)
        B0 = Bv.copy()
        Bv[:] = np.linalg.solve(L, pqA.T).T
        o.mat("D_mat")[:] = st.D0
        o.d("Cocc_mat", (n, nocc))[:] = st.C0
        o.build_fock()
        st.spread = float(np.max(np.abs(o.mat("F_mat") - st.F0)))
        Bv[:] = B0
        o.build_fock()
    return st.spread


@pytest.mark.parametrize("name", ["water", "c10h22"])
@pytest.mark.parametrize("mode", ["cholesky", "invsqrt"])
def test_df_tensor(ygpu, name, mode):
    """3-centre integrals, the metric factor W and the fitted operator B B^T against the oracle.
    invsqrt: W = J^-1/2, the reference's factor (symmetric, W J W = I); cholesky: W = L^-1 with
    J = L L^T (lower triangular, W J W^T = I).  B B^T = (mn|P) J^-1 (Q|kl) either way."""
    log("case: DF tensor", name, mode)
    st = oracle_state(name)
    with make_engine(ygpu, st) as e:
        set_oracle_sparsity(e, st)
        e.set_metric_factor(ygpu.METRIC_CHOLESKY if mode == "cholesky" else ygpu.METRIC_INVSQRT)
        r = e.build_df_tensor(want_pqA=True, want_Jpq=True, want_B=True)
    pqA_ref = st.o.d("pqA").reshape(r["pqA"].shape)
    err3 = np.max(np.abs(r["pqA"] - pqA_ref))
    log("  3c max abs err %.3e (max |value| %.3e)" % (err3, np.abs(pqA_ref).max()))
    assert err3 < 1e-11
    # the metric factor by its defining property, as good as the oracle's own J^-1/2
    J = metric(st)
    W, Wref = r["Jpq"], st.o.d("Jpq").reshape(r["Jpq"].shape)
    res_ref = np.max(np.abs(Wref @ J @ Wref - np.eye(len(J))))
    ev = np.linalg.eigvalsh(J)
    kappa_eps = ev[-1] / ev[0] * np.finfo(float).eps
    log("  metric condition number %.3e" % (ev[-1] / ev[0]))
    if mode == "invsqrt":
        res = np.max(np.abs(W @ J @ W - np.eye(len(J))))
        log("  |W J W - I|: gpu %.3e oracle %.3e; rel |dW| %.3e" % (res, res_ref,
            np.max(np.abs(W - Wref)) / np.abs(Wref).max()))
        assert np.max(np.abs(W - W.T)) < 1e-9 * np.abs(W).max()
    else:
        res = np.max(np.abs(W @ J @ W.T - np.eye(len(J))))
        log("  |W J W^T - I|: gpu %.3e (oracle's J^-1/2: %.3e)" % (res, res_ref))
        assert np.all(np.triu(W, 1) == 0.0)
        # the factor is unique: W = L^-1 of numpy's Cholesky factor
        Wnp = np.linalg.inv(np.linalg.cholesky(J))
        log("  rel |W - L^-1 (numpy)| %.3e" % (np.max(np.abs(W - Wnp)) / np.abs(Wnp).max()))
        assert np.max(np.abs(W - Wnp)) < 10 * kappa_eps * np.abs(Wnp).max()
    assert res < 10 * res_ref + 10 * kappa_eps
    # the fitted Coulomb operator B B^T on random vectors
    B_ref = st.o.d("df_tensor").reshape(r["B"].shape)
    x = np.random.RandomState(3).randn(B_ref.shape[0], 4)
    g, gref = r["B"] @ (r["B"].T @ x), B_ref @ (B_ref.T @ x)
    errG = np.max(np.abs(g - gref)) / np.abs(gref).max()
    log("  fitted operator rel err %.3e" % errG)
    assert errG < 1e-10


def test_metric_factors_agree_beyond_2048(ygpu):
    """C30H62 (naux 2910 > 2048, padded row length 2944): the reference's J^-1/2 (rocSOLVER
    dsyevd on the padded matrix) and the Cholesky factor describe the same J^-1 (W_is W_is =
    W_ch^T W_ch) and the same fitted operator B B^T."""
    log("case: metric factors at naux 2910")
    T = ygpu.TinySCF(xyz_path("c30h62"), niters=1)
    try:
        T.setup()
        mask, lid, rid = T.i("bf_pair_mask").copy(), T.i("uniq_sp_lid").copy(), T.i("uniq_sp_rid").copy()
        sp, df = T.d("sp_scrval").copy(), T.d("df_sp_scrval").copy()
    finally:
        T.close()
    s = system("c30h62")
    out = {}
    for mode in (ygpu.METRIC_INVSQRT, ygpu.METRIC_CHOLESKY):
        e = ygpu.Engine(0)
        try:
            e.set_system(s)
            e.set_sparsity(mask, lid, rid, sp, df)
            e.set_metric_factor(mode)
            out[mode] = e.build_df_tensor(want_Jpq=True, want_B=True)
        finally:
            e.close()
    Wi, Wc = out[ygpu.METRIC_INVSQRT]["Jpq"], out[ygpu.METRIC_CHOLESKY]["Jpq"]
    Ji_a, Ji_b = Wi @ Wi, Wc.T @ Wc
    rel = np.max(np.abs(Ji_a - Ji_b)) / np.abs(Ji_b).max()
    Bi, Bc = out[ygpu.METRIC_INVSQRT]["B"], out[ygpu.METRIC_CHOLESKY]["B"]
    x = np.random.RandomState(5).randn(Bi.shape[0], 3)
    gi, gc = Bi @ (Bi.T @ x), Bc @ (Bc.T @ x)
    relG = np.max(np.abs(gi - gc)) / np.abs(gc).max()
    log("  rel |W_is^2 - W_ch^T W_ch| %.3e; fitted operator rel diff %.3e; W_is symmetric %.1e"
        % (rel, relG, np.max(np.abs(Wi - Wi.T)) / np.abs(Wi).max()))
    assert np.max(np.abs(Wi - Wi.T)) < 1e-9 * np.abs(Wi).max()
    assert np.all(np.triu(Wc, 1) == 0.0)
    assert rel < 1e-5
    assert relG < 1e-10


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_fock_same_inputs(ygpu, name):
    """Same D, C_occ and the oracle's own B: F within 1e-10."""
    log("case: Fock with oracle B", name)
    st = oracle_state(name)
    with make_engine(ygpu, st) as e:
        set_oracle_sparsity(e, st)
        e.set_df_tensor(st.o.d("df_tensor"))
        r = e.build_fock(st.D0, st.C0, st.o.mat("Hcore_mat"))
        errJ = np.max(np.abs(r["J"] - st.J0))
        errK = np.max(np.abs(r["K"] - st.K0))
        errF = np.max(np.abs(r["F"] - st.F0))
        log("  max|dJ| %.3e max|dK| %.3e max|dF| %.3e" % (errJ, errK, errF))
        assert errF < 1e-10 and errJ < 1e-10 and errK < 1e-10


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_fock_device_tensor(ygpu, name):
    """End to end on the device (integrals, metric, B, J/K/F) from the same D, C_occ.
    Bar: 1e-10, or twice the oracle's own spread between two valid FP64 metric
    factorisations when the metric's conditioning makes that larger."""
    log("case: Fock with device B", name)
    st = oracle_state(name)
    with make_engine(ygpu, st) as e:
        set_oracle_sparsity(e, st)
        e.build_df_tensor()
        r = e.build_fock(st.D0, st.C0, st.o.mat("Hcore_mat"))
    errF = np.max(np.abs(r["F"] - st.F0))
    bar = max(1e-10, 2.0 * oracle_spread(st))
    log("  max|dF| %.3e (bar %.3e, oracle self-spread %.3e)" % (errF, bar, oracle_spread(st)))
    assert errF < bar


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_density(ygpu, name):
    log("case: density", name)
    st = oracle_state(name)
    o = st.o
    n, nocc = st.sys.nbf, st.sys.n_occ
    X = o.mat("X_mat").copy()
    Fo = X.T @ st.F0 @ X
    with make_engine(ygpu, st) as e:
        D, C, ev, idx = e.build_density(Fo, X, nocc)
        evref = np.linalg.eigvalsh(Fo)
        assert np.max(np.abs(ev - evref)) < 1e-10
        w, V = np.linalg.eigh(Fo)
        Cref = X @ V[:, :nocc]
        Dref = Cref @ Cref.T
        errD = np.max(np.abs(D - Dref))
        log("  max|dD| %.3e" % errD)
        assert errD < 1e-10
        assert np.max(np.abs(C @ C.T - D)) < 1e-12
        assert np.all(idx == np.arange(n))


def s_shells(n):
    """n single-primitive s shells: a basis that only sets nbf = n for the density build."""
    return dict(nshells=n, am=np.zeros(n), nprim=np.ones(n), prim_off=np.arange(n),
                xyz=np.zeros((n, 3)), alpha=np.ones(n), coef=np.ones(n))


def wilkinson_plus(m):
    """Eigenvalues of Wilkinson's W_{2m+1}^+ (diagonal |m - i|, off-diagonal 1), ascending: its
    largest eigenvalues come in pairs that agree to many digits."""
    k = 2 * m + 1
    W = np.diag(np.abs(m - np.arange(k)).astype(float)) + np.diag(np.ones(k - 1), 1) + np.diag(np.ones(k - 1), -1)
    return np.linalg.eigvalsh(W)


def hard_occupied_levels(n, kind, rng):
    """Occupied-level patterns that stress inverse iteration (ascending, in [-12, 0)):
    wilkinson: -W_21^+ repeated (near-degenerate pairs); glued: W_21^+ copies 1e-10 apart
    (clusters of near-equal levels); graded: levels 10^-k from 1e-9 to 1; tight: a run of 24
    levels within 1e-13 at the top of the occupied block."""
    if kind in ("wilkinson", "glued"):
        w = -wilkinson_plus(10)[::-1]                   # 21 levels, pairs at the bottom
        w = (w - w.min()) / (w.max() - w.min()) * 10.0 - 11.0
        reps = (n + len(w) - 1) // len(w)
        if kind == "wilkinson":
            lev = np.concatenate([w + 0.045 * r for r in range(reps)])
        else:
            lev = np.concatenate([w + 1e-10 * r for r in range(reps)])
        return np.sort(lev[:n])
    if kind == "graded":
        return -np.sort(np.logspace(-9, 0, n))[::-1] * 11.0
    if kind == "tight":
        lev = np.sort(rng.uniform(-12.0, -1.0, n))
        t = min(24, n)
        lev[n - t:] = -0.5 + 1e-13 * np.arange(t)
        return lev
    raise ValueError(kind)


def spectrum_case(n, kind, seed):
    """Symmetric F' and n_occ with a gap of at least 0.5 above the occupied levels."""
    rng = np.random.RandomState(seed)
    nocc = max(1, n // 4)
    lam = np.sort(rng.uniform(-12.0, 3.0, n))
    lam[nocc:] += 0.5
    if kind in ("wilkinson", "glued", "graded", "tight"):
        lam[:nocc] = hard_occupied_levels(nocc, kind, rng)
        lam[nocc:] = np.sort(rng.uniform(0.5, 3.5, n - nocc))
    if kind == "cluster" and nocc >= 4:
        lam[1:4] = lam[1]                       # an exactly degenerate occupied triple
    if kind == "diag":
        return np.diag(lam[rng.permutation(n)]), nocc
    Q, _ = np.linalg.qr(rng.randn(n, n))
    F = (Q * lam) @ Q.T
    return 0.5 * (F + F.T), nocc


# 272 is the largest size of the one-workgroup register-resident tridiagonal reduction;
# 273 .. 2048 take the blocked reduction with rocSOLVER dstedc; the register back-transform
# has variants for n <= 512, 1024, 1536 and 2048 (1600 and 2048 reach the last)
@pytest.mark.parametrize("n,kind", [(1, "dense"), (2, "dense"), (3, "dense"), (17, "cluster"), (31, "dense"),
                                    (32, "dense"), (33, "diag"), (64, "cluster"), (200, "diag"), (255, "dense"),
                                    (271, "cluster"), (272, "dense"), (273, "dense"), (300, "cluster"),
                                    (600, "cluster"), (1600, "cluster"), (2048, "dense"),
                                    (200, "wilkinson"), (272, "glued"), (240, "graded"), (260, "tight"),
                                    (600, "wilkinson"), (700, "glued"), (900, "tight")])
def test_density_eigensolver_sizes(ygpu, n, kind):
    log("case: density eigensolver n=%d %s" % (n, kind))
    F, nocc = spectrum_case(n, kind, 1000 + n)
    rng = np.random.RandomState(n)
    X = np.eye(n) + 0.05 * rng.randn(n, n) / np.sqrt(n)
    with make_engine(ygpu) as e:
        e.set_basis(0, s_shells(n))
        D, C, ev, idx = e.build_density(F, X, nocc)
    w, V = np.linalg.eigh(F)
    Dref = X @ V[:, :nocc] @ V[:, :nocc].T @ X.T
    scale = max(1.0, np.abs(w).max())
    errE, errD = np.max(np.abs(ev - w)), np.max(np.abs(D - Dref))
    # orthonormal occupied orbitals in the orthogonal basis: C0 = X^-1 C
    C0 = np.linalg.solve(X, C)
    errO = np.max(np.abs(C0.T @ C0 - np.eye(nocc)))
    log("  max|d eps| %.2e  max|dD| %.2e  orthonormality %.2e" % (errE, errD, errO))
    assert errE < 1e-12 * scale * max(1.0, n / 16)
    assert errD < 1e-11
    assert errO < 1e-12
    assert np.max(np.abs(C @ C.T - D)) < 1e-12


@pytest.mark.parametrize("n,kind", [(200, "dense"), (260, "tight")])
def test_density_small_route_residual_fallback(ygpu, monkeypatch, n, kind):
    """With the residual bound of the one-workgroup route at 0 every inverse-iteration vector
    fails its check, and the build must come from the general eigensolver (rocSOLVER dsyevd)
    with the same eigenvalues and density."""
    log("case: density small route forced to the general eigensolver n=%d %s" % (n, kind))
    monkeypatch.setenv("YGPU_TRIDIAG_RESID", "0")
    F, nocc = spectrum_case(n, kind, 1000 + n)
    X = np.eye(n)
    with make_engine(ygpu) as e:
        e.set_basis(0, s_shells(n))
        D, C, ev, idx = e.build_density(F, X, nocc)
    w, V = np.linalg.eigh(F)
    Dref = V[:, :nocc] @ V[:, :nocc].T
    errE, errD = np.max(np.abs(ev - w)), np.max(np.abs(D - Dref))
    log("  max|d eps| %.2e  max|dD| %.2e" % (errE, errD))
    assert errE < 1e-12 * max(1.0, np.abs(w).max()) * max(1.0, n / 16)
    assert errD < 1e-11
    assert np.max(np.abs(C.T @ C - np.eye(nocc))) < 1e-12


# S^-1/2 and metrics: one-workgroup reduction + dstedc up to 272, the blocked reduction +
# dstedc + the register back-transform up to 2048, dsyevd above
@pytest.mark.parametrize("n", [1, 5, 64, 260, 272, 273, 990, 1500, 2048, 2049])
def test_sym_inv_sqrt_sizes(ygpu, n):
    log("case: symmetric inverse square root n=%d" % n)
    rng = np.random.RandomState(77 + n)
    Q, _ = np.linalg.qr(rng.randn(n, n))
    lam = np.logspace(-4, 1, n)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    with make_engine(ygpu) as e:
        W, mn = e.sym_inv_sqrt(A)
    Wref = (Q / np.sqrt(lam)) @ Q.T
    w_np, V_np = np.linalg.eigh(A)              # LAPACK's own error level on the same matrix
    W_np = (V_np / np.sqrt(w_np)) @ V_np.T
    res = np.max(np.abs(W @ A @ W - np.eye(n)))
    res_np = np.max(np.abs(W_np @ A @ W_np - np.eye(n)))
    rel = np.max(np.abs(W - Wref)) / np.abs(Wref).max()
    rel_np = np.max(np.abs(W_np - Wref)) / np.abs(Wref).max()
    log("  |W A W - I| %.2e (numpy %.2e)  rel |dW| %.2e (numpy %.2e)  min eig %.6e"
        % (res, res_np, rel, rel_np, mn))
    assert abs(mn - lam[0]) < 1e-12 * lam[-1]
    assert np.array_equal(W, W.T)
    assert res < 10 * res_np + 1e-12
    assert rel < 10 * rel_np + 1e-13


@pytest.mark.parametrize("name,a2", [("water", True), ("c10h22", True), ("c10h22", False)])
def test_full_scf_energy(ygpu, name, a2):
    """Drop-in TinySCF engine (host C + device) against the oracle SCF; a2=False takes the
    DIIS slot to replace as the largest stored residual instead of the largest ever."""
    import oracle
    log("case: full SCF", name, "quirk_a2", a2)
    s = system(name)
    o = oracle.OracleSCF(s, niters=60, quirk_a2=a2)
    o.setup()
    o.build_df_tensor()
    o.do_scf()
    e_ref = o.getd("total_energy")
    it_ref = o.geti("iter")
    T = ygpu.TinySCF(xyz_path(name), niters=60)
    try:
        T.seti("quirk_a2", int(a2))
        T.setup()
        T.build_df_tensor()
        T.do_scf()
        e_gpu = T.getd("total_energy")
        log("  E gpu %.10f oracle %.10f dE %.3e iters %d/%d" % (e_gpu, e_ref, e_gpu - e_ref, T.geti("iter"), it_ref))
        assert T.geti("converged") == 1
        assert abs(e_gpu - e_ref) < 1e-8
    finally:
        T.close()


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_device_diis(ygpu, name):
    """Device DIIS history (DIIS.c:42-122) against the same algebra in NumPy."""
    log("case: device DIIS", name)
    st = oracle_state(name)
    o = st.o
    S, X = o.mat("S_mat").copy(), o.mat("X_mat").copy()
    rng = np.random.RandomState(11)
    Fs = []
    for k in range(3):
        P = rng.randn(*st.F0.shape) * 1e-3
        Fs.append(st.F0 + P + P.T)
    D = st.D0
    with make_engine(ygpu, st) as e:
        e.set_overlap(S, X)
        Fo0 = e.diis_transform(Fs[0])
        assert np.max(np.abs(Fo0 - X.T @ Fs[0] @ X)) < 1e-11
        Rs = []
        for k, F in enumerate(Fs):
            dots, Fo = e.diis_update(k, k + 1, F, D)
            FDS = F @ D @ S
            R = X.T @ (FDS - FDS.T) @ X
            Rs.append(R)
            ref = np.array([np.sum(R * Rj) for Rj in Rs])
            assert np.allclose(dots, ref, rtol=1e-11, atol=1e-18), (k, dots, ref)
            assert np.max(np.abs(Fo - X.T @ F @ X)) < 1e-11
        coef = np.array([0.2, -0.5, 1.3])
        Fx = e.diis_extrapolate(coef)
        ref = sum(c * (X.T @ F @ X) for c, F in zip(coef, Fs))
        assert np.max(np.abs(Fx - ref)) < 1e-11


def test_device_resident_handoffs(ygpu):
    """The SCF loop's device-resident calls (NULL host inputs: D and C_occ of the last density
    build, F of the last Fock build, F_orth of the last DIIS step) give bit-identical results to
    the same sequence with every input copied from the host."""
    log("case: device-resident hand-offs c10h22")
    st = oracle_state("c10h22")
    o = st.o
    S, X, H = o.mat("S_mat").copy(), o.mat("X_mat").copy(), o.mat("Hcore_mat").copy()
    nocc = st.C0.shape[1]
    out = []
    for resident in (False, True):
        seq = []
        with make_engine(ygpu, st) as e:
            set_oracle_sparsity(e, st)
            e.set_df_tensor(o.d("df_tensor"))
            e.set_overlap(S, X)
            r = e.build_fock(st.D0, st.C0, H)                       # iteration 0: the host guess
            Fo = e.diis_transform(None if resident else r["F"])
            D, C, ev, _ = e.build_density(None if resident else Fo, None, nocc)
            seq += [r["F"], r["J"], r["K"], Fo, D, C, ev]
            r = e.build_fock(None if resident else D, None if resident else C, H, n_occ=nocc)
            Fo = e.diis_transform(None if resident else r["F"])
            D1, C1, ev, _ = e.build_density(None if resident else Fo, None, nocc)
            seq += [r["F"], r["J"], r["K"], Fo, D1, C1, ev]
            r = e.build_fock(None if resident else D1, None if resident else C1, H, n_occ=nocc)
            dots, Fo = e.diis_update(0, 1, None if resident else r["F"], None if resident else D1,
                                     want_forth=not resident)
            err = e.diis_error(0)
            Fx = e.diis_extrapolate(np.array([1.0]))
            D2, C2, ev, _ = e.build_density(None if resident else Fx, None, nocc)
            seq += [r["F"], r["J"], r["K"], dots, np.array([err]), Fx, D2, C2, ev]
        out.append(seq)
    for k, (a, b) in enumerate(zip(*out)):
        assert np.array_equal(a, b), k
    # and the DIIS error is max |R| of the residual just stored
    R = X.T @ (out[0][-9] @ out[0][11] @ S - S @ out[0][11] @ out[0][-9]) @ X
    assert abs(out[0][-5][0] - np.abs(R).max()) < 1e-12 * np.abs(R).max() + 1e-15


@pytest.mark.parametrize("n_occ", [5, 100, 171, 250, 300, 330])
def test_k_build_tile_variants(ygpu, n_occ):
    """K = sum_P (B_P C)(B_P C)^T for synthetic occupied blocks of every width the X-build
    kernels specialise on (all occupied rows of a basis function in one workgroup up to 320,
    row tiles beyond), FP64 and mixed, against the closed form in NumPy on the oracle's tensor."""
    log("case: K tile variants n_occ", n_occ)
    st = oracle_state("c10h22")
    o, n = st.o, st.sys.nbf
    nnz, naux = o.geti("num_bf_pair_scr"), st.sys.naux
    B = o.d("df_tensor").reshape(nnz, naux)
    mask = o.i("bf_pair_mask").reshape(n, n)
    C = np.random.RandomState(n_occ).randn(n, n_occ) * 0.3
    # X_i = C(partners)^T B(rows of i); K = sum over (s, P) of X_i X_j
    X = np.zeros((n, n_occ, naux))
    for i in range(n):
        js = np.nonzero(mask[i] >= 0)[0]
        X[i] = C[js].T @ B[mask[i, js]]
    Xf = X.reshape(n, -1)
    K = Xf @ Xf.T
    S = np.abs(Xf) @ np.abs(Xf).T
    with make_engine(ygpu, st) as e:
        set_oracle_sparsity(e, st)
        e.set_df_tensor(o.d("df_tensor"))
        r64 = e.build_fock(st.D0, C, st.o.mat("Hcore_mat"))
        e.set_fock_precision(ygpu.FOCK_MIXED)
        rmx = e.build_fock(st.D0, C, st.o.mat("Hcore_mat"))
        e.set_fock_precision(ygpu.FOCK_MIXED6)
        rm6 = e.build_fock(st.D0, C, st.o.mat("Hcore_mat"))
    e64 = np.abs(r64["K"] - K).max() / np.abs(K).max()
    emx = (np.abs(rmx["K"] - K) / S).max()
    em6 = (np.abs(rm6["K"] - K) / S).max()
    log("  FP64 max|dK|/max|K| %.3e; BF16x3 max|dK|/S %.3e; BF16x6 max|dK|/S %.3e" % (e64, emx, em6))
    assert e64 < 1e-12
    assert emx < 2.0 ** -14
    assert em6 < 2.0 ** -17


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_initial_guess_matches_oracle(ygpu, name):
    """The host engine's SAD density (SAD blocks scaled by 0.5 (Z - Q) / Z, TinySCF.c:324-364)
    and its iteration-0 C_occ (D2Cocc with quirk A1, TinySCF.c:282-322) equal the oracle's."""
    log("case: initial guess", name)
    st = oracle_state(name)
    n, nocc = st.sys.nbf, st.sys.n_occ
    T = ygpu.TinySCF(xyz_path(name), niters=1)
    try:
        T.setup()
        D, C = T.mat("D_mat").copy(), T.d("Cocc_mat", (n, nocc)).copy()
    finally:
        T.close()
    log("  max|dD| %.2e max|dC_occ| %.2e" % (np.abs(D - st.D0).max(), np.abs(C - st.C0).max()))
    assert np.array_equal(D, st.D0)
    assert np.max(np.abs(C - st.C0)) < 1e-15 * max(1.0, np.abs(st.C0).max())
