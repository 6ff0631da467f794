"""Mixed-precision Fock builds on the GPU (ygpu_set_fock_precision).

K: BF16-split MFMA products with FP32 accumulation, checked against the
bit-exact CPU emulation of the same split (oracle/mixed.py; the only
difference left is the GPU's FP32 accumulation) and against the FP64 K.
J: incremental from the FP32 tensor; the first build is full FP64.
SCF: the mixed phase followed by FP64 iterations converges to the FP64
oracle energy within 1e-8 Eh."""
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


class Ref:
    """Two consecutive (D, C_occ) states of the oracle SCF and their FP64 J, K, F."""

    def __init__(self, name):
        import oracle
        self.sys = s = system(name)
        o = oracle.OracleSCF(s, niters=60)
        o.setup()
        o.build_df_tensor()
        n, nocc = s.nbf, s.n_occ
        self.o = o
        self.H = o.mat("Hcore_mat").copy()
        self.states = []
        for _ in range(2):
            D = o.mat("D_mat").copy()
            C = o.d("Cocc_mat", (n, nocc)).copy()
            o.build_fock()
            self.states.append(dict(D=D, C=C, J=o.mat("J_mat").copy(), K=o.mat("K_mat").copy(),
                                    F=o.mat("F_mat").copy()))
            o.diis()
            o.build_density()
        nnz = o.geti("num_bf_pair_scr")
        self.B = o.d("df_tensor").reshape(nnz, s.naux).copy()
        self.mask = o.i("bf_pair_mask").reshape(n, n).copy()


_refs = {}


def ref(name):
    if name not in _refs:
        _refs[name] = Ref(name)
    return _refs[name]


@pytest.mark.parametrize("name", ["water", "c10h22"])
@pytest.mark.parametrize("nparts", [2, 3])
def test_mixed_fock_against_emulation(ygpu, name, nparts):
    """nparts 2: YGPU_FOCK_MIXED (BF16x3); 3: YGPU_FOCK_MIXED6 (BF16x6)."""
    import mixed
    log("case: mixed Fock", name, "BF16 parts", nparts)
    r = ref(name)
    o = r.o
    e = ygpu.Engine(0)
    try:
        e.set_system(r.sys)
        e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                       o.d("df_sp_scrval"))
        e.set_df_tensor(o.d("df_tensor"))
        e.set_fock_precision(ygpu.FOCK_MIXED if nparts == 2 else ygpu.FOCK_MIXED6)
        out = []
        for st in r.states:
            out.append(e.build_fock(st["D"], st["C"], r.H))
        e.set_fock_precision(ygpu.FOCK_FP64)
        back = e.build_fock(r.states[1]["D"], r.states[1]["C"], r.H)
    finally:
        e.close()
    for k, (st, got) in enumerate(zip(r.states, out)):
        Kem, X32 = mixed.k_build(r.B, r.mask, st["C"], nparts)
        # S_ij = sum_k |X_ik| |X_jk| >= |K_ij|: the scale of rounding errors made in the K product;
        # T_ij = sum_k (A_ik |X_jk| + |X_ik| A_jk) with A = |C|^T |B| summed like X: the scale of
        # the FP32 accumulation errors of X carried into K
        Xa = np.abs(X32.reshape(X32.shape[0], -1).astype(np.float64))
        A = mixed.abs_x(r.B, r.mask, st["C"]).reshape(Xa.shape)
        S = Xa @ Xa.T + 1e-300
        T = A @ Xa.T + Xa @ A.T + 1e-300
        kmax = np.abs(st["K"]).max()
        e_emul = np.abs(got["K"] - Kem) / T
        e_fp64 = np.abs(got["K"] - st["K"]) / S
        d_J = np.abs(got["J"] - st["J"]).max()
        log("  build %d: max|K| %.3e  max|dK| vs emulation %.3e (max |dK|/T %.2e), vs FP64 %.3e (max |dK|/S %.2e),"
            " emulation vs FP64 %.3e; max|dJ| %.3e (max|J| %.3e)"
            % (k, kmax, np.abs(got["K"] - Kem).max(), e_emul.max(), np.abs(got["K"] - st["K"]).max(), e_fp64.max(),
               np.abs(Kem - st["K"]).max(), d_J, np.abs(st["J"]).max()))
        # the GPU differs from the exact-split emulation only by its FP32 accumulation: one
        # rounding per MFMA (three per 16-deep step) in X and in each 1024-long chunk of K,
        # ~1.5e-6 T when emulated step by step for c10h22; a dropped split term shows at ~1e-4 T
        assert e_emul.max() < 2.0 ** -17
        # the method's own error: BF16x3 <= 3 2^-17 per product plus the accumulation; BF16x6
        # ~2^-23 per product, the FP32 accumulation (~1e-6 of S) left
        assert e_fp64.max() < (2.0 ** -14 if nparts == 2 else 2.0 ** -17)
        assert np.allclose(got["K"], got["K"].T, rtol=0, atol=0)
        if k == 0:
            assert d_J < 1e-10            # first mixed build: full FP64 J
        else:
            # J_prev + J(D - D_prev) from the FP32 tensor: error ~ 2^-24 |dD| |B|^2
            assert d_J < 1e-7 * np.abs(st["J"]).max()
        F = r.H + 2 * got["J"] - got["K"]
        assert np.abs(got["F"] - F).max() < 1e-12 * max(1.0, np.abs(F).max())
    # back to FP64: the reference arithmetic again
    assert np.abs(back["F"] - r.states[1]["F"]).max() < 1e-10


@pytest.mark.parametrize("name", ["water", "c10h22"])
def test_diis_error_is_max_abs_residual(ygpu, name):
    log("case: DIIS error", name)
    r = ref(name)
    o = r.o
    S, X = o.mat("S_mat").copy(), o.mat("X_mat").copy()
    st = r.states[1]
    e = ygpu.Engine(0)
    try:
        e.set_system(r.sys)
        e.set_overlap(S, X)
        e.diis_update(0, 1, st["F"], st["D"])
        err = e.diis_error(0)
    finally:
        e.close()
    FDS = st["F"] @ st["D"] @ S
    R = X.T @ (FDS - FDS.T) @ X
    log("  max|R| gpu %.6e numpy %.6e" % (err, np.abs(R).max()))
    assert abs(err - np.abs(R).max()) <= 1e-12 * max(1.0, np.abs(R).max())


def test_mixed_scf_reaches_fp64_energy(ygpu):
    """Mixed phase until the DIIS error < 1e-5 (or stalls), then FP64 to convergence."""
    import oracle
    name = "c10h22"
    log("case: mixed SCF", name)
    s = system(name)
    o = oracle.OracleSCF(s, niters=60)
    o.setup()
    o.build_df_tensor()
    o.do_scf()
    e_ref = o.getd("total_energy")
    T = ygpu.TinySCF(xyz_path(name), niters=60)
    try:
        T.seti("mixed", 1)
        T.setup()
        T.build_df_tensor()
        T.do_scf()
        e = T.getd("total_energy")
        sw = T.geti("mixed_switch_iter")
        log("  E mixed %.10f oracle %.10f dE %.3e iters %d (oracle %d), FP64 from iteration %d"
            % (e, e_ref, e - e_ref, T.geti("iter"), o.geti("iter"), sw))
        assert T.geti("converged") == 1
        assert sw >= 3
        assert abs(e - e_ref) < 1e-8
    finally:
        T.close()


def test_difference_density_fock(ygpu):
    """YGPU_FOCK_INCR on C10H22: K = K_ref + K(D - D_ref) with the difference in BF16x3 products
    of the aligned orbitals.  States are consecutive SCF iterations of the oracle (occupied
    orbitals from its eigensolver, S-orthonormal).  (1) The next state against its FP64 K: the
    error is a small fraction of the change K - K_ref (it scales with the difference, not with
    K).  (2) The reference orbitals rotated within their occupied space (the same density): the
    alignment undoes the rotation and K equals K_ref to rounding.  (3) Without a reference the
    call makes a full FP64 build."""
    import oracle
    log("case: difference-density Fock build c10h22")
    s = system("c10h22")
    o = oracle.OracleSCF(s, niters=60)
    o.setup()
    o.build_df_tensor()
    n, nocc = s.nbf, s.n_occ
    H, S, X = o.mat("Hcore_mat").copy(), o.mat("S_mat").copy(), o.mat("X_mat").copy()
    states = []
    for _ in range(4):
        D, C = o.mat("D_mat").copy(), o.d("Cocc_mat", (n, nocc)).copy()
        o.build_fock()
        states.append(dict(D=D, C=C, K=o.mat("K_mat").copy()))
        o.diis()
        o.build_density()
    sa, sb = states[2], states[3]
    assert np.abs(sa["C"].T @ S @ sa["C"] - np.eye(nocc)).max() < 1e-10
    R, _ = np.linalg.qr(np.random.RandomState(9).randn(nocc, nocc))
    e = ygpu.Engine(0)
    try:
        e.set_system(s)
        e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                       o.d("df_sp_scrval"))
        e.set_df_tensor(o.d("df_tensor"))
        e.set_overlap(S, X)
        e.set_fock_precision(ygpu.FOCK_INCR)
        first = e.build_fock(sa["D"], sa["C"], H)               # no reference yet: full FP64
        m_first = e.fock_last_info()[0]
        e.set_fock_precision(ygpu.FOCK_FP64)
        ref0 = e.build_fock(sa["D"], sa["C"], H)                # the reference
        e.set_fock_precision(ygpu.FOCK_INCR)
        inc = e.build_fock(sb["D"], sb["C"], H)
        m_inc, a_inc = e.fock_last_info()
        e.set_fock_precision(ygpu.FOCK_FP64)
        e.build_fock(sa["D"], sa["C"], H)                       # back to the reference
        e.set_fock_precision(ygpu.FOCK_INCR)
        rot = e.build_fock(sa["D"], sa["C"] @ R, H)
        m_rot, a_rot = e.fock_last_info()
    finally:
        e.close()
    assert m_first == ygpu.FOCK_FP64 and np.array_equal(first["K"], ref0["K"])
    assert m_inc == ygpu.FOCK_INCR and m_rot == ygpu.FOCK_INCR
    dK = np.abs(sb["K"] - sa["K"]).max()
    err = np.abs(inc["K"] - sb["K"]).max()
    err_rot = np.abs(rot["K"] - ref0["K"]).max()
    log("  max|K_b - K_a| %.3e; difference build vs FP64 K_b: %.3e (%.2e of the change; alignment %.1e); "
        "rotated reference: %.3e (alignment %.1e)" % (dK, err, err / dK, a_inc, err_rot, a_rot))
    assert a_inc < 1e-12 and a_rot < 1e-12
    assert err < 2.0 ** -12 * dK
    assert err_rot < 1e-12 * np.abs(ref0["K"]).max()
    assert np.allclose(inc["K"], inc["K"].T, rtol=0, atol=0)


def test_j_stream_placement_gives_identical_fock(ygpu, monkeypatch):
    """The J build runs ahead of the K build on the main stream, or (YGPU_J_BESIDE=1, read when
    an engine is created) on the second stream beside it: the same kernels on the same data, so
    FP64, mixed (first and incremental J) and difference builds agree bit for bit."""
    log("case: J stream placement c10h22")
    r = ref("c10h22")
    o = r.o
    s = r.sys
    S, X = o.mat("S_mat").copy(), o.mat("X_mat").copy()

    def builds(beside):
        monkeypatch.setenv("YGPU_J_BESIDE", "1" if beside else "0")
        e = ygpu.Engine(0)
        try:
            e.set_system(s)
            e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                           o.d("df_sp_scrval"))
            e.set_df_tensor(o.d("df_tensor"))
            e.set_overlap(S, X)
            out = []
            # the last build: the reference's own density with its orbitals rotated (the
            # eigensolver's S-orthonormal orbitals of state 1), so the alignment runs
            R, _ = np.linalg.qr(np.random.RandomState(5).randn(s.n_occ, s.n_occ))
            st1 = r.states[1]
            for mode, D, C in [(ygpu.FOCK_MIXED, r.states[0]["D"], r.states[0]["C"]),
                               (ygpu.FOCK_MIXED, st1["D"], st1["C"]), (ygpu.FOCK_FP64, st1["D"], st1["C"]),
                               (ygpu.FOCK_INCR, st1["D"], st1["C"] @ R)]:
                e.set_fock_precision(mode)
                out.append(e.build_fock(D, C, r.H))
            assert e.fock_last_info()[0] == ygpu.FOCK_INCR
            return out
        finally:
            e.close()

    a, b = builds(False), builds(True)
    for k, (x, y) in enumerate(zip(a, b)):
        for name in ("J", "K", "F"):
            assert np.array_equal(x[name], y[name]), (k, name, np.abs(x[name] - y[name]).max())


def test_repeated_solves_reuse_device_blocks(ygpu):
    """Consecutive mixed-precision solves in one process: the second takes its large device
    buffers (FP32 tensor, operand blocks, X, k-chunk partials) from the cache of blocks the
    first engine gave back, the third runs after the cache went back to the driver; all three
    reach the same energy in the same iterations (the blocks hold the previous solve's data,
    which no kernel may read)."""
    name = "c10h22"
    log("case: repeated mixed solves", name)
    runs = []
    for k in range(3):
        if k == 2:
            ygpu.release_cached_memory()
        r = ygpu.solve(xyz_path(name), ene_tol=1e-10, mixed=True)
        log("  solve %d: %.3f s, %d iterations, E %.12f" % (k, r["seconds"], r["iters"], r["energy"]))
        runs.append(r)
    for r in runs[1:]:
        assert r["converged"] and r["iters"] == runs[0]["iters"]
        assert abs(r["energy"] - runs[0]["energy"]) < 1e-10
